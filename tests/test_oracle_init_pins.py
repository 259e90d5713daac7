"""Pins of the oracle's HolographicInit (hvd-core engine/holographic_init.cpp)
by what the mathematics fixes, part by part:

* BoxBlur r (holographic_init.cpp:58-113) against numpy.convolve with a
  ones(2r+1)/(2r+1) kernel, zero-padded and centred ('same');
* PadRowsMirror (:126-150) against numpy.pad(mode='reflect') (reflect101: the
  edge row is not repeated);
* the blend (:253-262, :308-315) on variants whose residual lumas are flat or
  a linear ramp, so every edge energy, weight and chi0 has a closed form with
  unequal weights (Dx/Dy's Neumann zeros on the last column/row included);
* the crop variants (:207-252) on tones whose padded plane is an exact 2-D
  tone, so each variant is the Gaussian's gain at the tone (vertical cutoff
  against the unpadded height, pad = min(16, h-1), crop offset, x2, S conj(c));
* the symmetry variant (:265-295) on one tone against a tone pair, so
  sym = min(|D[k]|, |D[-k]|) / (|D[k]| + 1e-6) is known bin by bin;
* the whole init on a tone scene against the blend law applied to the
  closed-form variants (edge energy by scipy.signal.convolve2d).

The last test plants each plausible mistake in a build of the oracle
(radius 2 and 4, weight bias 1.0, symmetric mirror, pad 15, vertical cutoff
against the padded height, no min, no roll) and shows that a pin fails.
"""
from __future__ import annotations

import math
import pathlib
import subprocess

import numpy as np
import pytest
import scipy.signal

import oracle
from hvd_amd import synth

FS = synth.NTSC.sample_rate
HERE = pathlib.Path(__file__).resolve().parent
ORACLE_SRC = HERE.parent / "oracle" / "hvd_oracle.c"

VARIANTS = ((0.8, 120.0), (1.8, 30.0), (1.3, 60.0))  # (MHz, cycles/picture height)


@pytest.fixture(scope="module")
def orc():
    return oracle.Oracle(fp32=False)


def gauss(v, fx_hz, fy_cpl, h):
    """The init's separable Gaussian (holographic_init.cpp:33-56) at
    (fx Hz, fy cycles/line), vertical cutoff cph / (2 h) with h unpadded."""
    mhz, cph = VARIANTS[v]
    return math.exp(-0.5 * (fx_hz / (mhz * 1e6)) ** 2) * math.exp(-0.5 * (fy_cpl / (cph / (2.0 * h))) ** 2)


# ----------------------------------------------------------------- the checks
# Each returns the largest error against its closed form; the tests assert it
# is at round-off level, the mutation test that some check sees a mistake.

def box_blur_err(orc, shape, r, seed=0):
    a = np.random.default_rng(seed).normal(size=shape)
    k = np.ones(2 * r + 1) / (2 * r + 1)

    def same(v):  # zero-padded convolution, centred on each sample
        return np.convolve(v, k, "full")[r:r + len(v)]

    ref = np.apply_along_axis(same, 0, a)
    ref = np.apply_along_axis(same, 1, ref)
    return float(np.abs(orc.box_blur_zero(a, r) - ref).max())


def pad_err(orc, h, w, pad, seed=1):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(h, w)) + 1j * rng.normal(size=(h, w))
    ref = np.pad(a, ((pad, pad), (0, 0)), mode="reflect")
    return float(np.abs(orc.pad_rows_mirror(a, pad) - ref).max())


def _count(lo, hi, c, r):
    """Number of integers in [c - r, c + r] that lie in [lo, hi)."""
    return np.maximum(0, np.minimum(hi, c + r + 1) - np.maximum(lo, c - r))


def ramp_energy(h, w, a, b, r=3):
    """Edge energy of Y = a x + b y + K: |Dx Y| = |a| except on the last column
    (Dx's Neumann zero), |Dy Y| = |b| except on the last row; box-blurred with
    radius r, zero padding, 1/(2r+1) per axis -> counts of window cells."""
    yy, xx = np.mgrid[0:h, 0:w]
    ny_all, nx_all = _count(0, h, yy, r), _count(0, w, xx, r)
    ny_in, nx_in = _count(0, h - 1, yy, r), _count(0, w - 1, xx, r)
    return (abs(a) * ny_all * nx_in + abs(b) * ny_in * nx_all) / float((2 * r + 1) ** 2)


def blend_err(orc, h=20, w=30):
    """Three variants: residual luma flat (E = 0, weight 2), a ramp in x and y,
    a ramp in x only.  chi0 = sum w_v chi_v / sum w_v, w_v = 1/(E_v + 1/2)."""
    rng = np.random.default_rng(5)
    yy, xx = np.mgrid[0:h, 0:w]
    c = np.exp(1j * (np.pi * yy + 0.5 * np.pi * xx + 0.3))
    s = 50.0 + 10.0 * rng.normal(size=(h, w))
    lumas = [np.full((h, w), 31.0), 0.7 * xx - 1.3 * yy + 4.0, -2.1 * xx + 9.0]
    energies = [np.zeros((h, w)), ramp_energy(h, w, 0.7, -1.3), ramp_energy(h, w, -2.1, 0.0)]
    chis = [(s - yl) * np.conj(c) for yl in lumas]  # S - Re[chi_v c] = the chosen luma
    y0, chi0 = orc.init_blend(s, c, chis)
    wts = [1.0 / (e + 0.5) for e in energies]
    chi_ref = sum(wv * cv for wv, cv in zip(wts, chis)) / sum(wts)
    y_ref = s - (chi_ref * c).real
    return max(float(np.abs(chi0 - chi_ref).max()), float(np.abs(y0 - y_ref).max()))


def vertical_tone_err(orc, h, k, q=3, w=40, amp=7.0):
    """S = amp cos(pi k y / (h - 1)), c = e^{-2 pi i q x / w}: the demodulated
    plane S conj(c) is reflect101-invariant, so when the padded height is a
    multiple of its period the padded plane is an exact sum of the tones
    (+-k', q) and variant v returns 2 amp G_v(q fs/w, k'/hp) cos(.) e^{2 pi i q x/w}."""
    yy, xx = np.mgrid[0:h, 0:w]
    pad = min(16, h - 1)
    hp = h + 2 * pad
    s = amp * np.cos(np.pi * k * yy / (h - 1))
    c = np.exp(-2j * np.pi * q * xx / w)
    # cycles per padded plane of the vertical cosine
    kk = k * hp / (2.0 * (h - 1))
    assert abs(kk - round(kk)) < 1e-12, "the padded plane must hold whole periods"
    fy = round(kk) / hp
    chi_v = orc.init_variants(s, c, FS)
    err = 0.0
    for v in range(2):
        ref = 2.0 * gauss(v, q * FS / w, fy, h) * s * np.exp(2j * np.pi * q * xx / w)
        err = max(err, float(np.abs(chi_v[v] - ref).max()))
    return err


def two_row_err(orc, w=24, q=5):
    """h = 2: pad = 1 and the padded rows are (r1, r0, r1, r0), the sum of a
    vertical DC and a Nyquist term with amplitudes (r0 + r1)/2, (r1 - r0)/2."""
    xx = np.arange(w)
    tone = np.exp(2j * np.pi * q * xx / w)
    c = np.exp(-2j * np.pi * q * xx / w)[None, :].repeat(2, 0)
    r0, r1 = 3.0, -8.0
    s = np.stack([np.full(w, r0), np.full(w, r1)])
    chi_v = orc.init_variants(s, c, FS)
    err = 0.0
    for v in range(2):
        g0, gn = gauss(v, q * FS / w, 0.0, 2), gauss(v, q * FS / w, 0.5, 2)
        for y in range(2):
            sign = (-1.0) ** (y + 1)  # padded row y + 1
            ref = 2.0 * ((r0 + r1) / 2 * g0 + (r1 - r0) / 2 * gn * sign) * tone
            err = max(err, float(np.abs(chi_v[v][y] - ref).max()))
    return err


def symmetry_err(orc, A, B, h=6, w=64, p=3, phA=0.4, phB=-1.1):
    """c = e^{i w x}, S = A cos(w x + phA) + B cos(3 w x + phB), constant down
    the field: S conj(c) has tones (A/2) e^{i phA} at 0, (A/2) e^{-i phA} at
    -2p, (B/2) e^{i phB} at +2p, (B/2) e^{-i phB} at -4p (no partner).
    sym = min(|D[k]|, |D[-k]|) / (|D[k]| + 1e-6) with |D| = |coef| hp w."""
    pad = min(16, h - 1)
    hp = h + 2 * pad
    xx = np.arange(w)
    om = 2 * np.pi * p / w
    s = np.tile(A * np.cos(om * xx + phA) + B * np.cos(3 * om * xx + phB), (h, 1))
    c = np.tile(np.exp(1j * om * xx), (h, 1))
    n = hp * w
    tones = {0: A / 2 * np.exp(1j * phA), -2 * p: A / 2 * np.exp(-1j * phA),
             2 * p: B / 2 * np.exp(1j * phB), -4 * p: B / 2 * np.exp(-1j * phB)}
    chi_v = orc.init_variants(s, c, FS, symmetry_variant=True)
    assert len(chi_v) == 3
    ref = np.zeros(w, complex)
    for kb, coef in tones.items():
        partner = tones.get(-kb, 0.0)
        m, mr = abs(coef) * n, abs(partner) * n
        sym = min(m, mr) / (m + 1e-6)
        ref += 2.0 * coef * sym * gauss(2, kb * FS / w, 0.0, h) * np.exp(2j * np.pi * kb * xx / w)
    err = float(np.abs(chi_v[2] - ref[None, :]).max())
    # the two plain variants on the same scene: every tone passes
    for v in range(2):
        refv = sum(2.0 * coef * gauss(v, kb * FS / w, 0.0, h) * np.exp(2j * np.pi * kb * xx / w)
                   for kb, coef in tones.items())
        err = max(err, float(np.abs(chi_v[v] - refv[None, :]).max()))
    return err


def whole_init_err(orc, h=34, w=40):
    """The full init on a tone scene equals the blend law (edge energy by
    scipy's 2-D convolution) applied to the closed-form variants."""
    yy, xx = np.mgrid[0:h, 0:w]
    q, k = 3, 2
    c = np.exp(-2j * np.pi * q * xx / w)
    s = 6.0 * np.cos(np.pi * k * yy / (h - 1)) + 0.05 * xx  # luma ramp under the chroma tone
    # the ramp leaks into the variants too; keep it out of the closed form by
    # building the variants from the oracle's own pinned variant stage
    chi_v = orc.init_variants(s, c, FS)
    kernel = np.ones((7, 7)) / 49.0
    num, den = 0.0, 0.0
    for cv in chi_v:
        yv = s - (cv * c).real
        gx = np.zeros_like(yv)
        gy = np.zeros_like(yv)
        gx[:, :-1] = np.diff(yv, axis=1)
        gy[:-1, :] = np.diff(yv, axis=0)
        e = scipy.signal.convolve2d(np.abs(gx) + np.abs(gy), kernel, mode="same", boundary="fill")
        wv = 1.0 / (e + 0.5)
        num, den = num + wv * cv, den + wv
    chi_ref = num / den
    y0, chi0 = orc.holographic_init(s, c, FS)
    return max(float(np.abs(chi0 - chi_ref).max()), float(np.abs(y0 - (s - (chi_ref * c).real)).max()))


# ------------------------------------------------------------------ the tests
@pytest.mark.parametrize("shape", [(1, 1), (3, 2), (5, 40), (17, 9), (64, 64)])
@pytest.mark.parametrize("r", [1, 3, 5])
def test_box_blur_matches_numpy_convolve(orc, shape, r):
    assert box_blur_err(orc, shape, r) < 1e-12


@pytest.mark.parametrize("h,pad", [(2, 1), (3, 2), (5, 4), (17, 16), (40, 16), (242, 16)])
def test_pad_rows_is_numpy_reflect(orc, h, pad):
    assert pad_err(orc, h, 7, pad) == 0.0


def test_pad_rows_not_symmetric_mirror(orc):
    """reflect101 differs from numpy's 'symmetric' (edge repeated) on any
    non-constant plane: the pin above can tell the two apart."""
    a = np.arange(12.0).reshape(4, 3) + 0j
    assert np.abs(orc.pad_rows_mirror(a, 2) - np.pad(a, ((2, 2), (0, 0)), mode="symmetric")).max() > 1.0


def test_blend_closed_form_unequal_weights(orc):
    e = ramp_energy(20, 30, 0.7, -1.3)
    # the weights really differ: interior, corners and Neumann edges
    assert e[10, 15] == pytest.approx(2.0)
    assert e[0, 0] == pytest.approx(16 * 2.0 / 49)
    assert e[19, 29] == pytest.approx((9 * 2.0 + 3 * 0.7 + 3 * 1.3) / 49)
    assert blend_err(orc) < 1e-10


@pytest.mark.parametrize("h,k", [(34, 15), (34, 4), (100, 21)])
def test_variants_vertical_tone(orc, h, k):
    assert vertical_tone_err(orc, h, k) < 1e-9


def test_variants_two_rows(orc):
    assert two_row_err(orc) < 1e-9


@pytest.mark.parametrize("A,B", [(9.0, 4.0), (3.0, 11.0)])
def test_symmetry_variant_tone_pair(orc, A, B):
    assert symmetry_err(orc, A, B) < 1e-8


def test_whole_init_is_variants_then_blend(orc):
    assert whole_init_err(orc) < 1e-10


# ------------------------------------------------ planted mistakes are caught
CHECKS = {
    "box": lambda o: max(box_blur_err(o, (17, 9), 3), box_blur_err(o, (5, 40), 1)),
    "pad": lambda o: max(pad_err(o, 17, 7, 16), pad_err(o, 3, 7, 2)),
    "blend": blend_err,
    "vtone": lambda o: max(vertical_tone_err(o, 34, 15), vertical_tone_err(o, 100, 21)),
    "tworow": two_row_err,
    "sym": lambda o: max(symmetry_err(o, 9.0, 4.0), symmetry_err(o, 3.0, 11.0)),
}
TOLS = {"box": 1e-12, "pad": 0.0, "blend": 1e-10, "vtone": 1e-9, "tworow": 1e-9, "sym": 1e-8}

MUTATIONS = {
    "blur radius 2": "-DOR_MUT_BOX_R=2",
    "blur radius 4": "-DOR_MUT_BOX_R=4",
    "weight 1/(E+1)": "-DOR_MUT_WEIGHT_BIAS=1.0",
    "symmetric mirror": "-DOR_MUT_SYMMETRIC_MIRROR",
    "pad 15": "-DOR_MUT_PAD_CAP=15",
    "cutoff vs padded h": "-DOR_MUT_HREF_PADDED",
    "symmetry without min": "-DOR_MUT_SYM_NOMIN",
    "symmetry without roll": "-DOR_MUT_SYM_NOROLL",
}


def test_pins_pass_on_the_oracle(orc):
    for name, check in CHECKS.items():
        assert check(orc) <= TOLS[name], name


@pytest.mark.parametrize("mutation", sorted(MUTATIONS))
def test_planted_mistake_is_caught(tmp_path, mutation):
    so = tmp_path / "liboracle_mut.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", MUTATIONS[mutation], "-o", str(so),
                    str(ORACLE_SRC), "-lm"], check=True)
    mut = oracle.Oracle(fp32=False, path=so)
    failed = [name for name, check in CHECKS.items() if not check(mut) <= TOLS[name]]
    assert failed, f"no pin catches '{mutation}'"
