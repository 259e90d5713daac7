"""hvd_gpu_variational_refine_noalloc against hvd_gpu_variational_refine.

The no-allocation entry point runs the same solver kernels with the workspace
and descriptors in the library's fixed per-GPU scratch; its results must be
the existing entry point's bit for bit, on two calls in a row (the second
smaller, so it runs over what the first left in the scratch), and a call that
needs more scratch than there is must fail before touching its outputs.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import gpu, synth  # noqa: E402

DEV = "cuda:0"


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _cplx(z):
    return _dev(np.stack([z.real, z.imag], -1).astype(np.float32))


def _field_inputs(orc64, g, seed):
    """A synthetic field's composite plane, lock-in theta (and V-switch) and
    FP64 init chi0, as the refine receives them."""
    raw = (synth.ntsc_field if g is synth.NTSC else synth.pal_field)(seed)[0]
    s, _, th, vs = orc64.prepare_field(synth.ire_from_raw(raw), g)
    fal = g.first_active_field_line
    th_act = th[fal:fal + g.lines].astype(np.float32)
    vs_act = vs[fal:fal + g.lines] if g.uses_vswitch else None
    car = orc64.make_carrier(th_act, vs_act, g)
    _, chi0 = orc64.holographic_init(s.astype(np.float32), car, g.sample_rate)
    return dict(kind="field", g=g, s=s.astype(np.float32), th=th_act, vs=vs_act, chi0=chi0.astype(np.complex64))


def _scene_inputs(h, w):
    r, c = np.mgrid[0:h, 0:w].astype(np.float64)
    car = np.exp(1j * (0.5 * math.pi * c + 0.3 * r))
    chi_true = 10 * np.sin(0.4 * (c - r)) - 6j * np.cos(0.2 * c)
    s = (30 + 20 * np.sin(0.35 * (r + c)) + (chi_true * car).real).astype(np.float32)
    chi0 = (chi_true + 2.0 * np.sin(0.9 * c)).astype(np.complex64)
    return dict(kind="scene", s=s, car=car.astype(np.complex64), chi0=chi0)


def _planes(cases):
    planes, outs = [], []
    for cs in cases:
        dchi = _cplx(cs["chi0"])
        dy = torch.zeros(cs["s"].shape, device=DEV)
        if cs["kind"] == "field":
            g = cs["g"]
            pl = gpu.Engine.plane(_dev(cs["s"]), dchi, dy, theta=_dev(cs["th"]),
                                  vswitch=None if cs["vs"] is None else _dev(cs["vs"]),
                                  a0=g.active_video_start, w0=gpu.geometry_phase_per_sample(g),
                                  sample_rate=g.sample_rate)
        else:
            pl = gpu.Engine.plane(_dev(cs["s"]), dchi, dy, carrier=_cplx(cs["car"]))
        planes.append(pl)
        outs.append((dy, dchi))
    return planes, outs


def _run(engine, cases, params, no_alloc):
    planes, outs = _planes(cases)
    iters, aniso = engine.variational_refine(planes, params, want_stats=True, no_alloc=no_alloc)
    torch.cuda.synchronize()
    return [(y.cpu().numpy(), c.cpu().numpy()) for y, c in outs], iters, aniso


def _same(engine, cases, params):
    want, it_w, an_w = _run(engine, cases, params, no_alloc=False)
    got, it_g, an_g = _run(engine, cases, params, no_alloc=True)
    assert np.array_equal(it_w, it_g) and np.array_equal(an_w, an_g)
    for (yw, cw), (yg, cg) in zip(want, got):
        assert np.isfinite(yg).all()
        assert np.array_equal(yw, yg) and np.array_equal(cw, cg)


def test_noalloc_matches_refine_two_calls(engine, orc64):
    params = gpu.SolverParams.make(cg_iterations=100, irls_outer=4, fast=0, cg_tol=1e-12)
    print("noalloc call 1: 3 NTSC fields + 1 scene 67x130", flush=True)
    first = [_field_inputs(orc64, synth.NTSC, s) for s in (3, 4, 5)] + [_scene_inputs(67, 130)]
    _same(engine, first, params)
    print("noalloc call 2: 1 NTSC field + 1 scene 9x1 (smaller)", flush=True)
    second = [_field_inputs(orc64, synth.NTSC, 6), _scene_inputs(9, 1)]
    _same(engine, second, params)


def test_noalloc_pal_and_defaults(engine, orc64):
    params = gpu.SolverParams.make(cg_iterations=2, irls_outer=4, fast=1, is_pal=1)
    print("noalloc: 2 PAL fields, hvd-core defaults", flush=True)
    _same(engine, [_field_inputs(orc64, synth.PAL, s) for s in (7, 8)], params)


def test_noalloc_rejects_oversize(engine):
    # 8 PAL-sized planes need about 85 MB of solver workspace, more than the
    # 63 MiB the scratch leaves for it: refused before any work is queued
    h, w = synth.PAL.lines, synth.PAL.width
    cases = [_scene_inputs(h, w) for _ in range(8)]
    planes, outs = _planes(cases)
    before = [c.clone() for _, c in outs]
    params = gpu.SolverParams.make(cg_iterations=4, irls_outer=2, fast=0, chroma_aniso=0.5)
    with pytest.raises(gpu.HvdGpuError, match="no-allocation"):
        engine.variational_refine(planes, params, no_alloc=True)
    torch.cuda.synchronize()
    for b, (_, c) in zip(before, outs):
        assert torch.equal(b, c)
    # the context still works afterwards, both ways
    _same(engine, [_scene_inputs(48, 56)], params)
