"""GPU parity of VariationalRefine with temporal neighbours (variational.cpp:
198-556 with NeighborTerm, variational.h:54-58) against the oracle's 3-D
solve, which is pinned bit-exact to hvd-core's own solve in
test_oracle_temporal.py.  Float results are held to TOL_IRE; the entry point
is called through the C ABI (hvd_gpu_variational_refine_temporal)."""
import numpy as np
import pytest

from conftest import TOL_IRE
from temporal_cases import temporal_scene

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import gpu  # noqa: E402
import oracle as O  # noqa: E402

DEV = "cuda:0"
CFG = dict(cg_iterations=40, irls_outer=4, fast=0, chroma_aniso=0.7)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(DEV)


def cdev(z):
    return dev(np.stack([z.real, z.imag], -1))


def chost(t):
    a = t.cpu().numpy()
    return a[..., 0].astype(np.float64) + 1j * a[..., 1]


def reference(orc64, orc32, s, car, chi0, nbrs, kw, nu, eps_t):
    """The FP64 oracle's answer and the tolerance around it: TOL_IRE, or on
    a system so ill-conditioned that float rounding alone moves the answer
    further (single-line and single-column fields), twice the distance
    between the FP64 evaluation and hvd-core's own FP32 bits (orc32 is
    bit-exact with hvd-core).  The iteration counts are compared where the
    two evaluations agree on them."""
    cfg = O.Cfg.make(**kw)
    y6, c6, _, i6 = orc64.variational_refine_nbr(s, car, chi0, cfg, nbrs, nu, eps_t)
    y3, c3, _, i3 = orc32.variational_refine_nbr(s, car, chi0, cfg, nbrs, nu, eps_t)
    spread = max(np.abs(y6 - y3).max(), np.abs(c6 - c3).max())
    return y6, c6, (i6 if np.array_equal(i6, i3) else None), max(TOL_IRE, 2.0 * spread)


def run_gpu(engine, s, car, chi0, nbrs, cfg, nu, eps_t):
    dchi, dy = cdev(chi0), torch.zeros(s.shape, device=DEV)
    pl = gpu.Engine.plane(dev(s), dchi, dy, carrier=cdev(car))
    nb = [(dev(a), cdev(b), dev(c)) for a, b, c in nbrs]
    iters = engine.variational_refine_temporal(pl, nb, gpu.SolverParams.make(**cfg), nu, eps_t, want_stats=True)
    return dy.cpu().numpy(), chost(dchi), iters


@pytest.mark.parametrize("hw,n_nbr,nu,eps_t,seed", [((24, 40), 2, 0.5, 2.0, 1), ((24, 40), 2, 1.5, 0.7, 1),
                                                    ((33, 70), 1, 1.0, 1.0, 5), ((17, 129), 3, 0.8, 1.5, 6),
                                                    ((1, 9), 2, 1.0, 1.0, 7), ((9, 1), 2, 1.0, 1.0, 8),
                                                    ((1, 1), 1, 1.0, 1.0, 9)])
def test_temporal_matches_oracle(engine, orc64, orc32, hw, n_nbr, nu, eps_t, seed):
    h, w = hw
    s, car, chi0, nbrs = temporal_scene(h, w, seed, n_nbr=n_nbr)
    y_ref, c_ref, it_ref, tol = reference(orc64, orc32, s, car, chi0, nbrs, CFG, nu, eps_t)
    if min(h, w) > 1:
        assert tol == TOL_IRE
    y, c, it = run_gpu(engine, s, car, chi0, nbrs, CFG, nu, eps_t)
    assert np.abs(y - y_ref).max() <= tol
    assert np.abs(c - c_ref).max() <= tol
    if it_ref is not None:
        assert np.array_equal(it, it_ref)


@pytest.mark.parametrize("kw", [dict(cg_iterations=2, irls_outer=4, fast=1),       # hvd-core's defaults
                                dict(cg_iterations=60, irls_outer=3, fast=0),       # AUTO aniso
                                dict(cg_iterations=40, irls_outer=4, fast=0, cg_tol=0.2, lambda_c=2.0)])
def test_temporal_configs(engine, orc64, orc32, kw):
    s, car, chi0, nbrs = temporal_scene(48, 96, 11)
    y_ref, c_ref, it_ref, tol = reference(orc64, orc32, s, car, chi0, nbrs, kw, 0.7, 1.2)
    assert tol == TOL_IRE
    y, c, it = run_gpu(engine, s, car, chi0, nbrs, kw, 0.7, 1.2)
    assert np.abs(y - y_ref).max() <= TOL_IRE
    assert np.abs(c - c_ref).max() <= TOL_IRE
    if it_ref is not None:
        assert np.array_equal(it, it_ref)


def test_temporal_field_size(engine, orc64):
    # an NTSC field's active grid (263 lines x 758 samples)
    s, car, chi0, nbrs = temporal_scene(263, 758, 12)
    y_ref, c_ref, _, _ = orc64.variational_refine_nbr(s, car, chi0, O.Cfg.make(**CFG), nbrs, 0.5, 2.0)
    y, c, _ = run_gpu(engine, s, car, chi0, nbrs, CFG, 0.5, 2.0)
    assert np.abs(y - y_ref).max() <= TOL_IRE
    assert np.abs(c - c_ref).max() <= TOL_IRE


def test_temporal_reduces_to_2d(engine):
    # no strength: the 2-D solve's own bits; zero confidence: every temporal
    # term is zero, so the 3-D kernels land on the 2-D answer
    s, car, chi0, nbrs = temporal_scene(40, 64, 13, n_nbr=2)
    dchi, dy = cdev(chi0), torch.zeros(s.shape, device=DEV)
    engine.variational_refine([gpu.Engine.plane(dev(s), dchi, dy, carrier=cdev(car))], gpu.SolverParams.make(**CFG))
    y2, c2 = dy.cpu().numpy(), chost(dchi)
    y, c, _ = run_gpu(engine, s, car, chi0, nbrs, CFG, 0.0, 1.0)
    assert np.array_equal(y, y2) and np.array_equal(c, c2)
    off = [(a, b, np.zeros_like(cf)) for a, b, cf in nbrs]
    y, c, _ = run_gpu(engine, s, car, chi0, off, CFG, 1.0, 1.0)
    assert np.abs(y - y2).max() <= TOL_IRE and np.abs(c - c2).max() <= TOL_IRE


def test_temporal_neighbours_change_the_answer(engine):
    s, car, chi0, nbrs = temporal_scene(24, 40, 1)
    y2, c2, _ = run_gpu(engine, s, car, chi0, nbrs, CFG, 0.0, 1.0)
    y3, c3, _ = run_gpu(engine, s, car, chi0, nbrs, CFG, 1.5, 0.7)
    assert np.abs(c3 - c2).max() > 1.0


def test_temporal_argument_errors(engine):
    s, car, chi0, nbrs = temporal_scene(8, 8, 14, n_nbr=1)
    with pytest.raises(gpu.HvdGpuError, match="temporal_eps"):
        run_gpu(engine, s, car, chi0, nbrs, CFG, 1.0, 0.0)
    with pytest.raises(gpu.HvdGpuError, match="neighbour count"):
        run_gpu(engine, s, car, chi0, nbrs * 17, CFG, 1.0, 1.0)


@pytest.mark.parametrize("diag_prior", [0.0, 0.5])
def test_stationary_start_exits_on_curvature(engine, orc64, diag_prior):
    """A scene that starts at its minimum (flat composite 37.5 IRE, zero
    chroma, neighbours that agree): the gradient and so the first direction
    are zero, hh = 0 and every IRLS stage takes the curvature exit before its
    first step (variational.cpp's `hh <= 1e-12` break).  The 2-D and 3-D
    solves leave chi untouched and count the iterations as the oracle does."""
    h, w = 40, 64
    kw = dict(CFG, diag_prior=diag_prior)
    s = np.full((h, w), 37.5)
    r, c = np.mgrid[0:h, 0:w]
    car = np.exp(1j * (0.5 * np.pi * c + 0.3 * r))
    chi0 = np.zeros((h, w), np.complex128)
    nbrs = [(s.copy(), car.copy(), np.full((h, w), 0.8)) for _ in range(2)]
    y_ref, c_ref, _, it_ref = orc64.variational_refine(s, car, chi0, O.Cfg.make(**kw))
    assert not it_ref.any() and not np.abs(c_ref).any()
    y6, c6, _, it6 = orc64.variational_refine_nbr(s, car, chi0, O.Cfg.make(**kw), nbrs, 1.0, 1.0)
    assert not it6.any() and not np.abs(c6).any()
    dchi, dy = cdev(chi0), torch.zeros(s.shape, device=DEV)
    it, _ = engine.variational_refine([gpu.Engine.plane(dev(s), dchi, dy, carrier=cdev(car))],
                                      gpu.SolverParams.make(**kw), want_stats=True)
    assert np.array_equal(it[0], it_ref)
    assert not chost(dchi).any() and np.array_equal(dy.cpu().numpy(), y_ref.astype(np.float32))
    y, cc, it = run_gpu(engine, s, car, chi0, nbrs, kw, 1.0, 1.0)
    assert np.array_equal(np.asarray(it).reshape(-1), it6.reshape(-1))
    assert not cc.any() and np.array_equal(y, y6.astype(np.float32))
