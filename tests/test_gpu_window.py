"""GPU parity of the 3-D window driver (hvd_gpu_decode_window, hvd-core's
DecodeFieldWindowWithInits, engine/sequence.cpp:371-1114) against the
oracle's restatement (oracle/sequence3d.py), itself bit-exact with hvd-core's
own driver built from source (tests/test_oracle_sequence3d.py), and of
VerifyMotion against the FP32 oracle.  The driver's decisions (sigma, the
resolved strength and gates) must agree; Y and chi are held to TOL_IRE."""
import numpy as np
import pytest

from conftest import TOL_IRE
from sequence_cases import window

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import gpu  # noqa: E402
import sequence3d as S3  # noqa: E402

DEV = "cuda:0"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(DEV)


def cdev(z):
    return dev(np.stack([z.real, z.imag], -1))


def chost(t):
    a = t.cpu().numpy()
    return a[..., 0].astype(np.float64) + 1j * a[..., 1]


def run_gpu(engine, fields, parities, inits, cfg: S3.WindowCfg, pal=False):
    planes, outs = [], []
    for (s, car), chi0 in zip(fields, inits):
        dchi, dy = cdev(chi0), torch.zeros(s.shape, device=DEV)
        planes.append(gpu.Engine.plane(dev(s), dchi, dy, carrier=cdev(car)))
        outs.append((dy, dchi))
    sp = gpu.SolverParams.make(lambda_c=cfg.lambda_c, chroma_aniso=cfg.chroma_aniso, diag_prior=cfg.diag_prior,
                               cg_iterations=cfg.cg_iterations, irls_outer=cfg.irls_outer, cg_tol=cfg.cg_tol,
                               fast=int(cfg.fast))
    wp = gpu.WindowParams.make(temporal_strength=cfg.temporal_strength, temporal_eps=cfg.temporal_eps,
                               nr_anchor=cfg.nr_anchor, nr_eps=cfg.nr_eps, coherence_gate=cfg.coherence_gate,
                               odd_gate_floor=cfg.odd_gate_floor, passes=cfg.passes, nr_radius=cfg.nr_radius,
                               mc_tile=cfg.mc_tile, mc_search=cfg.mc_search, bidirectional=int(cfg.bidirectional),
                               extended_temporal=int(cfg.extended_temporal), trajectory_fit=int(cfg.trajectory_fit),
                               psi_init=int(cfg.psi_init), output_fidelity=int(cfg.output_fidelity),
                               parallel_across_fields=int(cfg.parallel_across_fields))
    d = engine.decode_window(planes, parities, pal, sp, wp)
    torch.cuda.synchronize()
    return [y.cpu().numpy() for y, _ in outs], [chost(c) for _, c in outs], d


CASES = {
    "2d": dict(cg_iterations=40, temporal_strength=-1.0),
    "3d_fixed": dict(cg_iterations=40, temporal_strength=2.0),
    "auto": dict(cg_iterations=40),
    "slow": dict(cg_iterations=40, temporal_strength=2.0, fast=False),
    "psi": dict(cg_iterations=40, temporal_strength=1.0, psi_init=True),
    "sequential": dict(cg_iterations=40, temporal_strength=1.0, parallel_across_fields=False, output_fidelity=False),
    "defaults": dict(),
}


def _check(orc32, engine, fields, parities, inits, cfg, pal=False):
    d_ref = {}
    ry, rc = S3.decode_window_with_inits(orc32, fields, parities, inits, cfg, is_pal=pal, diag=d_ref)
    print("window", cfg, flush=True)
    gy, gc, d = run_gpu(engine, fields, parities, inits, cfg, pal)
    for k in d_ref:
        assert abs(d[k] - d_ref[k]) <= 1e-4 * max(1.0, abs(d_ref[k])), (k, d[k], d_ref[k])
    ey = max(np.abs(a - b).max() for a, b in zip(gy, ry))
    ec = max(np.abs(a - b).max() for a, b in zip(gc, rc))
    assert ey <= TOL_IRE and ec <= TOL_IRE, (ey, ec)
    return gy, gc, d


@pytest.mark.parametrize("name", sorted(CASES))
def test_window_static(engine, orc32, name):
    fields, parities, inits, _ = window()
    _check(orc32, engine, fields, parities, inits, S3.WindowCfg(**CASES[name]))


@pytest.mark.parametrize("motion", [(0, 2), (1, 3)])
def test_window_moving(engine, orc32, motion):
    fields, parities, inits, _ = window(lines=64, width=160, motion=motion, seed=11)
    _check(orc32, engine, fields, parities, inits, S3.WindowCfg(cg_iterations=30, temporal_strength=1.0))


def test_window_pal(engine, orc32):
    fields, parities, inits, _ = window(lines=48, width=128, nf=10, pal=True, seed=13)
    _check(orc32, engine, fields, parities, inits, S3.WindowCfg(cg_iterations=30, temporal_strength=1.0), pal=True)


def test_window_3d_beats_2d_and_lossless(engine):
    # sequence_test.cpp:139-160 on the device
    fields, parities, inits, truth = window()
    _, c2, _ = run_gpu(engine, fields, parities, inits, S3.WindowCfg(cg_iterations=40, temporal_strength=-1.0))
    y3, c3, _ = run_gpu(engine, fields, parities, inits, S3.WindowCfg(cg_iterations=40, temporal_strength=2.0))

    def psnr(ch):
        return 10 * np.log10(1e4 / np.mean([np.mean(np.abs(a - t) ** 2) for a, t in zip(ch, truth)]))
    assert psnr(c3) > psnr(c2)
    for (s, car), y, c in zip(fields, y3, c3):
        assert np.abs(s - (y + (c * car).real)).max() < 1e-3


def test_verify_motion_bit_exact(engine, orc32):
    rng = np.random.default_rng(3)
    ya = rng.normal(50, 10, (96, 160)).astype(np.float32)
    yb = np.roll(ya, (1, 3), (0, 1)).astype(np.float32)
    pdy = rng.uniform(-2, 2, (3, 5)).astype(np.float32)
    pdx = rng.uniform(-4, 4, (3, 5)).astype(np.float32)
    pdy[1, 2], pdx[1, 2] = 1.2, 2.9   # near the true shift: the prediction wins
    dy, dx, cf = engine.verify_motion([dev(ya)], [dev(yb)], [dev(pdy)], [dev(pdx)], tile=32)
    wdy, wdx, wcf = orc32.verify_motion(ya, yb, pdy, pdx, 32)
    assert np.array_equal(dy[0].cpu().numpy(), wdy) and np.array_equal(dx[0].cpu().numpy(), wdx)
    assert np.array_equal(cf[0].cpu().numpy(), wcf)
