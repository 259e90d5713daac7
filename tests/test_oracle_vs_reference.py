"""Pin the oracle against hvd-core itself.

oracle/_ref/libhvdref.so is hvd-core's own FFT-free engine code (lockin.cpp,
variational.cpp, motion.cpp, ntsc_geometry.cpp, colour.cpp, gradients.h)
compiled from the reference checkout.  The FP32 build of the oracle follows
hvd-core operation for operation, so it must agree with it BIT FOR BIT; the
FP64 oracle (the GPU's checker) must agree within the float rounding of the
reference.  Skipped where the reference build is unavailable (the committed
golden vectors in tests/golden/ carry the same pins there).
"""
import math

import numpy as np
import pytest

from conftest import TOL_IRE

import oracle as O
from hvd_amd import synth


def _rng(seed=0):
    return np.random.default_rng(seed)


@pytest.mark.parametrize("hw", [(1, 1), (1, 7), (7, 1), (2, 2), (64, 48)])
def test_gradient_operators_bit_exact(orc32, ref, hw):
    rng = _rng(1)
    a = rng.normal(size=hw).astype(np.float32)
    for op in range(4):
        assert np.array_equal(orc32.real_op(op, a), ref.real_op(op, a)), op
    c = (rng.normal(size=hw) + 1j * rng.normal(size=hw)).astype(np.complex64)
    for op in range(8):
        assert np.array_equal(orc32.cplx_op(op, c), ref.cplx_op(op, c)), op


@pytest.mark.parametrize("n", [1, 2, 5, 263, 313])
def test_tridiag_bit_exact(orc32, ref, n):
    rng = _rng(n)
    d = rng.normal(size=n).astype(np.float32)
    a = rng.uniform(0, 2, size=n).astype(np.float32)
    assert np.array_equal(orc32.tridiag_smooth(d, a, 25.0), ref.tridiag_smooth(d, a, 25.0))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_lockin_ntsc(orc32, orc64, ref, seed):
    f = synth.ire_from_raw(synth.ntsc_field(seed)[0])
    want = ref.burst_lockin_phase(f, synth.NTSC)
    assert np.array_equal(orc32.burst_lockin_phase(f, synth.NTSC), want)
    got64 = orc64.burst_lockin_phase(f, synth.NTSC)
    assert np.abs(got64 - want).max() < 2e-4  # float32 theta of |theta| <= 830 rad


@pytest.mark.parametrize("seed", [3, 4])
def test_lockin_pal(orc32, orc64, ref, seed):
    f = synth.ire_from_raw(synth.pal_field(seed)[0])
    wt, wv = ref.burst_lockin_phase_pal(f, synth.PAL)
    t32, v32 = orc32.burst_lockin_phase_pal(f, synth.PAL)
    assert np.array_equal(t32, wt) and np.array_equal(v32, wv)
    t64, v64 = orc64.burst_lockin_phase_pal(f, synth.PAL)
    assert np.array_equal(v64, wv)
    assert np.abs(t64 - wt).max() < 3e-4


def test_amplitude_and_acc(orc32, orc64, ref):
    for g, raw in ((synth.NTSC, synth.ntsc_field(5)[0]), (synth.PAL, synth.pal_field(6)[0])):
        f = synth.ire_from_raw(raw)
        a = ref.burst_amplitude_ire(f, g)
        assert orc32.burst_amplitude_ire(f, g) == a
        assert abs(orc64.burst_amplitude_ire(f, g) - a) < 1e-4
    for m in (0.0, 0.3, 5.0, 9.99, 20.0, 21.43, 39.9, 50.0):
        assert orc32.acc_gain(m, 20.0) == ref.acc_gain(m, 20.0)
        assert orc32.acc_gain(m, 21.43) == ref.acc_gain(m, 21.43)


def test_carrier_bit_exact(orc32, ref):
    f = synth.ire_from_raw(synth.pal_field(7)[0])
    th, vs = ref.burst_lockin_phase_pal(f, synth.PAL)
    th, vs = th[22:310], vs[22:310]
    assert np.array_equal(orc32.make_carrier(th, vs, synth.PAL), ref.make_carrier(th, vs, synth.PAL))
    f = synth.ire_from_raw(synth.ntsc_field(7)[0])
    th = ref.burst_lockin_phase(f, synth.NTSC)[21:]
    assert np.array_equal(orc32.make_carrier(th, None, synth.NTSC), ref.make_carrier(th, None, synth.NTSC))


def _scene(h, w):
    # variational_v2_test.cpp:64-95: diagonal chroma + luma texture, corrupted chi0
    r, c = np.mgrid[0:h, 0:w].astype(np.float32)
    car = np.exp(1j * (0.5 * math.pi * c)).astype(np.complex64)
    y_true = 30 + 20 * np.sin(0.35 * (r + c))
    diag = np.sin(0.4 * (c - r))
    chi_true = 10 * diag - 6j * diag
    s = (y_true + (chi_true * car).real).astype(np.float32)
    e = 3 * np.sin(0.9 * c) * np.cos(0.7 * r)
    return s, car, (chi_true + e - 1j * e).astype(np.complex64)


@pytest.mark.parametrize("kw", [dict(cg_iterations=60, fast=0), dict(cg_iterations=60, fast=1),
                                dict(cg_iterations=60, fast=0, diag_prior=0.7),
                                dict(cg_iterations=60, fast=0, cg_tol=0.2),
                                dict(cg_iterations=2, fast=1)])
def test_refine_scene_bit_exact(orc32, orc64, ref, kw):
    s, car, chi0 = _scene(48, 56)
    cfg = O.Cfg.make(**kw)
    yr, cr = ref.variational_refine(s, car, chi0, cfg)
    y32, c32, _, _ = orc32.variational_refine(s, car, chi0, cfg)
    assert np.array_equal(y32, yr) and np.array_equal(c32.astype(np.complex64), cr.astype(np.complex64))
    y64, c64, _, _ = orc64.variational_refine(s, car, chi0, cfg)
    assert np.abs(y64 - yr).max() <= TOL_IRE and np.abs(c64 - cr).max() <= TOL_IRE


@pytest.mark.parametrize("std", ["ntsc", "pal"])
def test_refine_field_bit_exact(orc32, orc64, ref, std):
    """A full active field with AUTO aniso, through hvd-core's own lock-in and
    carrier; chi0 from the oracle's init (hvd-core's needs FFTW)."""
    g = synth.NTSC if std == "ntsc" else synth.PAL
    raw = (synth.ntsc_field if std == "ntsc" else synth.pal_field)(9)[0]
    f = synth.ire_from_raw(raw)
    fal, lal = g.first_active_field_line, g.first_active_field_line + g.lines
    if g.uses_vswitch:
        th, vs = ref.burst_lockin_phase_pal(f, g)
        car = ref.make_carrier(th[fal:lal], vs[fal:lal], g)
    else:
        th = ref.burst_lockin_phase(f, g)
        car = ref.make_carrier(th[fal:lal], None, g)
    s = f[fal:lal, g.active_video_start:g.active_video_end]
    _, chi0 = orc64.holographic_init(s, car, g.sample_rate)
    chi0 = chi0.astype(np.complex64)
    cfg = O.Cfg.make(cg_iterations=40, irls_outer=4, fast=0, cg_tol=1e-12, is_pal=int(g.uses_vswitch))
    yr, cr = ref.variational_refine(s, car, chi0, cfg)
    y32, c32, _, _ = orc32.variational_refine(s, car, chi0, cfg)
    assert np.array_equal(y32, yr) and np.array_equal(c32.astype(np.complex64), cr.astype(np.complex64))
    y64, c64, _, _ = orc64.variational_refine(s, car, chi0, cfg)
    assert np.abs(y64 - yr).max() <= TOL_IRE and np.abs(c64 - cr).max() <= TOL_IRE


def test_box_blur_bit_exact(orc32, ref):
    a = (_rng(3).normal(size=(40, 50)) * 50).astype(np.float32)
    for r in (0, 1, 2, 5):
        assert np.array_equal(orc32.box_blur(a, r), ref.box_blur(a, r))


@pytest.mark.parametrize("shift,tile,search", [((2, -3), 32, 8), ((-1, 5), 32, 16), ((0, 0), 16, 8)])
def test_estimate_motion_bit_exact(orc32, orc64, ref, shift, tile, search):
    sc = synth.make_scene(synth.NTSC, _rng(4), canvas_pad=16)
    Y = sc.y.astype(np.float32)
    dy, dx = shift
    A = Y[16:16 + 128, 16:16 + 160]
    B = Y[16 - dy:16 - dy + 128, 16 - dx:16 - dx + 160]
    got = orc32.estimate_motion(B, A, tile, search)
    want = ref.estimate_motion(B, A, tile, search)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    # FP64 oracle: same integer vectors on confident tiles
    g64 = orc64.estimate_motion(B, A, tile, search)
    conf = want[2] > 0.5
    assert np.array_equal(np.round(g64[0])[conf], np.round(want[0])[conf])
    assert np.array_equal(np.round(g64[1])[conf], np.round(want[1])[conf])
