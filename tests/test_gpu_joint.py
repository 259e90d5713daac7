"""GPU parity of the batched per-pixel solves that the 3-D window runs:
VariationalRefineJoint (variational.cpp:558-891; hvd_gpu_variational_refine_joint)
against the FP64 oracle's joint solve, which is bit-exact with hvd-core's
own in FP32 (tests/test_oracle_sequence3d.py), and
hvd_gpu_variational_refine_temporal_batch against one call per field.
Float results are held to TOL_IRE (or, on single-line planes, twice the
FP64-vs-FP32 spread), iteration counts where the two precisions agree."""
import numpy as np
import pytest

from conftest import TOL_IRE
from temporal_cases import temporal_scene

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import gpu  # noqa: E402
import oracle as O  # noqa: E402

DEV = "cuda:0"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(DEV)


def cdev(z):
    return dev(np.stack([z.real, z.imag], -1))


def chost(t):
    a = t.cpu().numpy()
    return a[..., 0].astype(np.float64) + 1j * a[..., 1]


def _case(h, w, seed, nn=2, anchor=True):
    s, car, chi0, nbrs = temporal_scene(h, w, seed, n_nbr=nn)
    rng = np.random.default_rng(seed + 100)
    y0 = (s - (chi0 * car).real + rng.normal(0, 1.0, s.shape)).astype(np.float32)
    an = None
    if anchor:
        an = (y0 + rng.normal(0, 0.5, s.shape).astype(np.float32), chi0 + 0.3 - 0.2j,
              rng.uniform(0.1, 1.0, s.shape).astype(np.float32))
    return s, car, chi0, y0, nbrs, an


def _gpu_joint(engine, cases, kw, nu, eps_t, nu_a):
    planes, outs, nbl, anl = [], [], [], []
    for s, car, chi0, y0, nbrs, an in cases:
        dchi, dy = cdev(chi0), dev(y0)
        planes.append(gpu.Engine.plane(dev(s), dchi, dy, carrier=cdev(car)))
        outs.append((dy, dchi))
        nbl.append([(dev(a), cdev(b), dev(c)) for a, b, c in nbrs])
        anl.append(None if an is None else (dev(an[0]), cdev(an[1]), dev(an[2])))
    it = engine.variational_refine_joint(planes, nbl, anl, gpu.SolverParams.make(**kw), nu, eps_t, nu_a,
                                         want_stats=True)
    return [(y.cpu().numpy(), chost(c)) for y, c in outs], it


KW = dict(cg_iterations=40, irls_outer=4, fast=0, chroma_aniso=0.7)


@pytest.mark.parametrize("kw", [KW, dict(cg_iterations=2, irls_outer=4, fast=1),
                                dict(cg_iterations=60, irls_outer=3, fast=0),
                                dict(KW, diag_prior=0.7)])
def test_joint_matches_oracle(engine, orc64, orc32, kw):
    cases = [_case(40, 72, 3), _case(24, 40, 4, nn=3), _case(33, 50, 5, nn=0), _case(20, 30, 6, anchor=False)]
    got, it = _gpu_joint(engine, cases, kw, 0.8, 1.5, 1.0)
    cfg = O.Cfg.make(**kw)
    for k, (s, car, chi0, y0, nbrs, an) in enumerate(cases):
        y6, c6, _, i6 = orc64.joint_refine(s, car, y0, chi0, cfg, nbrs, 0.8, 1.5, an, 1.0)
        i3 = orc32.joint_refine(s, car, y0, chi0, cfg, nbrs, 0.8, 1.5, an, 1.0)[3]
        assert np.abs(got[k][0] - y6).max() <= TOL_IRE
        assert np.abs(got[k][1] - c6).max() <= TOL_IRE
        if np.array_equal(i6, i3):
            assert np.array_equal(it[k], i6)


def test_joint_single_line_and_column(engine, orc64, orc32):
    cases = [_case(1, 33, 8), _case(17, 1, 9), _case(1, 1, 10)]
    got, _ = _gpu_joint(engine, cases, KW, 1.0, 1.0, 0.5)
    cfg = O.Cfg.make(**KW)
    for k, (s, car, chi0, y0, nbrs, an) in enumerate(cases):
        y6, c6, _, _ = orc64.joint_refine(s, car, y0, chi0, cfg, nbrs, 1.0, 1.0, an, 0.5)
        y3, c3, _, _ = orc32.joint_refine(s, car, y0, chi0, cfg, nbrs, 1.0, 1.0, an, 0.5)
        tol = max(TOL_IRE, 2.0 * max(np.abs(y6 - y3).max(), np.abs(c6 - c3).max()))
        assert np.abs(got[k][0] - y6).max() <= tol
        assert np.abs(got[k][1] - c6).max() <= tol


def test_temporal_batch_matches_single(engine):
    cases = [temporal_scene(40, 72, 21, n_nbr=2), temporal_scene(24, 40, 22, n_nbr=3),
             temporal_scene(30, 64, 23, n_nbr=1)]
    p = gpu.SolverParams.make(**KW)
    single = []
    for s, car, chi0, nbrs in cases:
        dchi, dy = cdev(chi0), torch.zeros(s.shape, device=DEV)
        engine.variational_refine_temporal(gpu.Engine.plane(dev(s), dchi, dy, carrier=cdev(car)),
                                           [(dev(a), cdev(b), dev(c)) for a, b, c in nbrs], p, 0.9, 1.2)
        single.append((dy.cpu().numpy(), dchi.cpu().numpy()))
    planes, outs, nbl = [], [], []
    for s, car, chi0, nbrs in cases:
        dchi, dy = cdev(chi0), torch.zeros(s.shape, device=DEV)
        planes.append(gpu.Engine.plane(dev(s), dchi, dy, carrier=cdev(car)))
        outs.append((dy, dchi))
        nbl.append([(dev(a), cdev(b), dev(c)) for a, b, c in nbrs])
    engine.variational_refine_temporal_batch(planes, nbl, p, 0.9, 1.2)
    for (y1, c1), (dy, dchi) in zip(single, outs):
        assert np.array_equal(y1, dy.cpu().numpy()) and np.array_equal(c1, dchi.cpu().numpy())
