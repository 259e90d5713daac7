"""GPU parity of the solve's adjoint (hvd_gpu_refine_adjoint, FP32) against
the FP64 oracle adjoint (itself pinned to central differences and unrolled
autograd by tests/test_adjoint_cpu.py), and the torch.autograd binding."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import autograd as HA  # noqa: E402
from hvd_amd import gpu  # noqa: E402
import oracle as O  # noqa: E402
from adjoint_cases import TRAIN_CFG, crops, loss_grads  # noqa: E402
from hvd_amd import synth  # noqa: E402

DEV = "cuda:0"


def c2(z):
    return np.stack([z.real, z.imag], -1).astype(np.float32)


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _oracle(orc, c):
    cfg = O.Cfg.make(cg_iterations=4 * 2000, cg_tol=1e-14, **TRAIN_CFG)
    y, chi, st, _, _ = orc.variational_refine_stages(c["s"], c["car"], c["chi0"], cfg, hh_floor=0.0, gg_floor=1e-26)
    gy, gchi = loss_grads(y, chi, c["yt"], c["chit"])
    gs, gchi0, glam, _ = orc.refine_adjoint(c["s"], c["car"], cfg, st, chi, gchi, gy, adj_iters=4000, adj_tol=1e-13)
    return gs, gchi0, glam


def test_gpu_adjoint_matches_fp64_oracle(engine, orc64):
    """16 crops of 128 x 64 in one batch: FP32 forward + adjoint on the GPU vs
    the FP64 implicit gradient (dL/dS, dL/dchi0, dL/dlog lambda)."""
    cases = crops(orc64, n=16, h=64, w=128)
    params = gpu.SolverParams.make(cg_iterations=4 * 300, cg_tol=1e-7, **TRAIN_CFG)
    s = [t(c["s"].astype(np.float32)) for c in cases]
    car = [t(c2(c["car"])) for c in cases]
    chi = [t(c2(c["chi0"])) for c in cases]
    y = [torch.empty_like(x) for x in s]
    planes = [gpu.Engine.plane(s[i], chi[i], y=y[i], carrier=car[i]) for i in range(len(cases))]
    stages = engine.variational_refine_stages(planes, params)
    gys, gchis = [], []
    for i, c in enumerate(cases):
        ch = chi[i].cpu().numpy().astype(np.float64)
        gy, gchi = loss_grads(y[i].cpu().numpy().astype(np.float64), ch[..., 0] + 1j * ch[..., 1], c["yt"], c["chit"])
        gys.append(t(gy.astype(np.float32)))
        gchis.append(t(c2(gchi)))
    gs, gchi0, glam = engine.refine_adjoint(planes, params, stages, gchis, gys, adj_iterations=400, adj_tol=1e-7)
    worst = dict(s=0.0, chi0=0.0, lam=0.0)
    for i, c in enumerate(cases):
        rs, rc, rl = _oracle(orc64, c)
        got_c = gchi0[i].cpu().numpy().astype(np.float64)
        worst["s"] = max(worst["s"], np.abs(gs[i].cpu().numpy() - rs).max() / np.abs(rs).max())
        worst["chi0"] = max(worst["chi0"], max(np.abs(got_c[..., 0] - rc.real).max(),
                                               np.abs(got_c[..., 1] - rc.imag).max()) / np.abs(rc).max())
        worst["lam"] = max(worst["lam"], float(np.abs(glam[i] - rl).max() / np.abs(rl).max()))
    print("worst relative errors", worst, flush=True)
    # measured 2.7e-5, 2.7e-5 and 6.5e-6; the bars are about ten times that
    assert worst["s"] <= 3e-4 and worst["chi0"] <= 3e-4 and worst["lam"] <= 1e-4, worst


def test_autograd_refine_backward_and_adam(engine, orc64):
    """RefineFunction's backward returns the adjoint's gradients, and a few
    Adam steps on (log lambda_Y, log lambda_C) lower the loss."""
    cases = crops(orc64, n=8, h=64, w=128, seed0=500)
    s = t(np.stack([c["s"] for c in cases]).astype(np.float32))
    car = t(np.stack([c2(c["car"]) for c in cases]))
    chi0 = t(np.stack([c2(c["chi0"]) for c in cases]))
    yt = t(np.stack([c["yt"] for c in cases]).astype(np.float32))
    cht = t(np.stack([c2(c["chit"]) for c in cases]))
    ly = torch.tensor(0.5, device=DEV, requires_grad=True)
    lc = torch.tensor(-0.5, device=DEV, requires_grad=True)
    opt = torch.optim.Adam([ly, lc], lr=0.05)
    losses = []
    for _ in range(15):
        opt.zero_grad()
        y, chi = HA.refine(s, car, chi0, ly, lc, engine, cg_iterations=4 * 60, cg_tol=1e-6, adj_iterations=80)
        n = y[0].numel()
        L = ((y - yt) ** 2).sum() / n + ((chi - cht) ** 2).sum() / n
        L.backward()
        assert torch.isfinite(ly.grad) and torch.isfinite(lc.grad)
        losses.append(float(L.detach()))
        opt.step()
    print("losses", [round(x, 4) for x in losses], flush=True)
    assert losses[-1] < losses[0]


def _full_fields(orc64, seeds):
    """Full NTSC fields (263 x 758 active): the oracle's composite, carrier
    and FP64 holographic init, with the generator's truth."""
    g = synth.NTSC
    out = []
    for sd in seeds:
        raw, sc, _ = synth.ntsc_field(sd)
        s, car, _, _ = orc64.prepare_field(synth.ire_from_raw(raw), g)
        _, chi0 = orc64.holographic_init(s, car, g.sample_rate)
        out.append(dict(s=s, car=car, chi0=chi0, yt=sc.y, chit=sc.chi))
    return out


def test_gpu_adjoint_full_fields(engine, orc64):
    """Two full NTSC fields: the FP32 forward + adjoint on the GPU against the
    FP64 implicit gradient, at converged settings and at the training
    benchmark's (IRLS 4 x CG 25, adjoint 25 iterations per stage)."""
    import concurrent.futures as cf
    cases = _full_fields(orc64, (7000, 7001))
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        refs = list(ex.map(lambda c: _oracle(orc64, c), cases))

    def run(params, adj_iterations, adj_tol):
        s = [t(c["s"].astype(np.float32)) for c in cases]
        car = [t(c2(c["car"])) for c in cases]
        chi = [t(c2(c["chi0"])) for c in cases]
        y = [torch.empty_like(x) for x in s]
        planes = [gpu.Engine.plane(s[i], chi[i], y=y[i], carrier=car[i]) for i in range(len(cases))]
        stages = engine.variational_refine_stages(planes, params)
        gys, gchis = [], []
        for i, c in enumerate(cases):
            ch = chi[i].cpu().numpy().astype(np.float64)
            gy, gchi = loss_grads(y[i].cpu().numpy().astype(np.float64), ch[..., 0] + 1j * ch[..., 1], c["yt"],
                                  c["chit"])
            gys.append(t(gy.astype(np.float32)))
            gchis.append(t(c2(gchi)))
        gs, gchi0, glam = engine.refine_adjoint(planes, params, stages, gchis, gys, adj_iterations=adj_iterations,
                                                adj_tol=adj_tol)
        worst = dict(s=0.0, chi0=0.0, lam=0.0)
        for i, (rs, rc, rl) in enumerate(refs):
            got_c = gchi0[i].cpu().numpy().astype(np.float64)
            worst["s"] = max(worst["s"], np.abs(gs[i].cpu().numpy() - rs).max() / np.abs(rs).max())
            worst["chi0"] = max(worst["chi0"], max(np.abs(got_c[..., 0] - rc.real).max(),
                                                   np.abs(got_c[..., 1] - rc.imag).max()) / np.abs(rc).max())
            worst["lam"] = max(worst["lam"], float(np.abs(glam[i] - rl).max() / np.abs(rl).max()))
        return worst

    conv = run(gpu.SolverParams.make(cg_iterations=4 * 300, cg_tol=1e-7, **TRAIN_CFG), 400, 1e-7)
    print("full fields, converged: worst relative errors", conv, flush=True)
    for it, tol, ait, atol in ((25, 1e-12, 25, 1e-6), (60, 1e-6, 60, 1e-6), (100, 1e-7, 100, 1e-7)):
        e = run(gpu.SolverParams.make(cg_iterations=4 * it, cg_tol=tol, **TRAIN_CFG), ait, atol)
        print(f"full fields, IRLS 4 x CG <= {it} (tol {tol:g}), adjoint <= {ait} (tol {atol:g}) per stage: "
              f"worst relative errors", e, flush=True)
    # measured 1.3e-5, 1.0e-5 and 6.4e-6
    assert conv["s"] <= 2e-4 and conv["chi0"] <= 2e-4 and conv["lam"] <= 1e-4, conv
