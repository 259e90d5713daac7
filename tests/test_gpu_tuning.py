"""GPU: the solver's launch shape (strip height, solve-group size) and its autotuner.

The launch shape changes how the CG's FP64 sums are ordered and nothing
else: every per-pixel value is formed by the same operations whichever strip
and line forms it.  Results are bit-identical whatever the grouping, and
across strip heights for these inputs (the sums' order reaches the float
step sizes only when a sum lies within its last bits of a float rounding
boundary); any strip height stays within TOL_IRE of the FP64 oracle.
"""
import math

import numpy as np
import pytest

from conftest import TOL_IRE

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import gpu, synth  # noqa: E402
import oracle as O  # noqa: E402

DEV = "cuda:0"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def cplx_dev(z):
    return dev(np.stack([z.real, z.imag], -1).astype(np.float32))


def scene(h, w, k):
    r, c = np.mgrid[0:h, 0:w].astype(np.float64)
    car = np.exp(1j * (0.5 * math.pi * c + 0.3 * r))
    y_true = 30 + 20 * np.sin(0.35 * (r + c) + k)
    diag = np.sin(0.4 * (c - r) + 0.5 * k)
    chi_true = 10 * diag - 6j * diag
    s = (y_true + (chi_true * car).real).astype(np.float32)
    e = 3 * np.sin(0.9 * c + k) * np.cos(0.7 * r)
    chi0 = (chi_true + e - 1j * e).astype(np.complex64).astype(np.complex128)
    return s, car.astype(np.complex64).astype(np.complex128), chi0


CFG = dict(cg_iterations=40, irls_outer=4, fast=0, chroma_aniso=0.7)


def refine(eng, scenes):
    planes, outs = [], []
    for s, car, chi0 in scenes:
        dchi, dy = cplx_dev(chi0), torch.zeros(s.shape, device=DEV)
        planes.append(gpu.Engine.plane(dev(s), dchi, dy, carrier=cplx_dev(car)))
        outs.append((dy, dchi))
    eng.variational_refine(planes, gpu.SolverParams.make(**CFG))
    return [(dy.cpu().numpy(), dchi.cpu().numpy()) for dy, dchi in outs]


@pytest.fixture()
def eng():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    e = gpu.Engine(0)  # its own context: tunings stay out of the shared engine
    yield e
    e.close()


def test_autotune_records_and_is_used(eng, orc64):
    h, w, n = 37, 211, 5
    rows, gpx, ns = eng.autotune_solver(h, w, False, n, iterations=4)
    print(f"autotune {h}x{w}x{n}: rows {rows}, group {gpx} px, {ns:.3f} ns/px-it", flush=True)
    assert 2 <= rows <= 32 and gpx >= h * w and ns > 0.0
    scenes = [scene(h, w, k) for k in range(n)]
    got = refine(eng, scenes)  # runs with the recorded tuning
    for (s, car, chi0), (y, chi) in zip(scenes, got):
        y_ref, chi_ref, _, _ = orc64.variational_refine(s, car, chi0, O.Cfg.make(**CFG))
        assert np.abs(y - y_ref).max() <= TOL_IRE
        assert np.abs((chi[..., 0] + 1j * chi[..., 1]) - chi_ref).max() <= TOL_IRE
    # the same answer with the recorded tuning replaced by a fixed one
    eng.set_solver_tuning(rows, gpx)
    fixed = refine(eng, scenes)
    for (y1, c1), (y2, c2) in zip(got, fixed):
        assert np.array_equal(y1, y2) and np.array_equal(c1, c2)


@pytest.mark.parametrize("rows", [2, 5, 16, 32])
def test_strip_heights_agree(eng, orc64, rows):
    scenes = [scene(70, 150, k) for k in range(3)]
    eng.set_solver_tuning(rows, 0)
    got = refine(eng, scenes)
    for (s, car, chi0), (y, chi) in zip(scenes, got):
        y_ref, chi_ref, _, _ = orc64.variational_refine(s, car, chi0, O.Cfg.make(**CFG))
        assert np.abs(y - y_ref).max() <= TOL_IRE
        assert np.abs((chi[..., 0] + 1j * chi[..., 1]) - chi_ref).max() <= TOL_IRE


def test_strip_heights_bit_identical(eng):
    # strips that start on different lines of the same plane: a line formed
    # at the top of a strip or inside its walk gets the same bits
    scenes = [scene(40, 150, 0)]
    res = {}
    for rows in (2, 3, 4, 8, 13):
        eng.set_solver_tuning(rows, 0)
        res[rows] = refine(eng, scenes)[0]
    for rows in (3, 4, 8, 13):
        same = np.array_equal(res[rows][0], res[2][0]) and np.array_equal(res[rows][1], res[2][1])
        print(f"strip rows {rows} vs 2: {'same bits' if same else 'DIFF'}", flush=True)
        assert same


def test_grouping_is_bit_neutral(eng):
    # fields solved one group at a time or all together: same bits at a fixed strip height
    raws = [synth.ntsc_field(60 + k)[0] for k in range(5)]
    p = gpu.SolverParams.make(cg_iterations=20, irls_outer=2, fast=0)
    px = 242 * 760
    res = []
    for group in (px, 2 * px, 100 * px):
        eng.set_solver_tuning(8, group)
        ys, chis, _ = eng.decode_fields_u16([dev(r.view(np.int16)) for r in raws], [synth.NTSC] * 5, p)
        res.append([(y.cpu().numpy(), c.cpu().numpy()) for y, c in zip(ys, chis)])
    for other in res[1:]:
        for (y1, c1), (y2, c2) in zip(res[0], other):
            assert np.array_equal(y1, y2) and np.array_equal(c1, c2)


def test_tuning_rejects_bad_values(eng):
    for rows in (1, 33, -2):
        with pytest.raises(gpu.HvdGpuError):
            eng.set_solver_tuning(rows, 0)
    with pytest.raises(gpu.HvdGpuError):
        eng.set_solver_tuning(0, -5)
    with pytest.raises(gpu.HvdGpuError):
        eng.autotune_solver(1, 100, False, 4)
    with pytest.raises(gpu.HvdGpuError):
        eng.autotune_solver(40, 100, False, 0)
