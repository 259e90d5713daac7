"""The adjoint of the IRLS + CG solve (oracle/hvd_oracle.c refine_adjoint)
against FP64 central finite differences and against an unrolled FP64 torch
autograd gradient of an independently written forward.

The adjoint treats every IRLS stage as solved exactly, so the forward here
solves each stage to round-off (relative tolerance 1e-14, absolute floor
|g|^2 < 1e-26 N instead of hvd-core's 1e-10 N).  Directional derivatives
are taken along the adjoint's own gradient (largest derivative per unit
step, so the difference quotient's round-off is smallest) and along a random
direction (whose smaller derivative carries more of the solves' round-off).
Gradients: d L / d log lambda_Y, d L / d log lambda_C, d L / d S and
d L / d chi0, with L the MSE of Y, U, V against the generator's truth."""
import numpy as np
import pytest

import oracle as O
from adjoint_cases import TRAIN_CFG, crops, loss, loss_grads

FD_H = 1e-6
EXACT = dict(cg_iterations=4 * 2000, cg_tol=1e-14)
FLOORS = dict(hh_floor=0.0, gg_floor=1e-26)


def _forward(orc, c, lam_y=1.0, lam_c=1.0, s=None, chi0=None):
    cfg = O.Cfg.make(lambda_y=lam_y, lambda_c=lam_c, **TRAIN_CFG, **EXACT)
    y, chi, st, _, _ = orc.variational_refine_stages(c["s"] if s is None else s, c["car"],
                                                     c["chi0"] if chi0 is None else chi0, cfg, **FLOORS)
    return loss(y, chi, c["yt"], c["chit"]), y, chi, st, cfg


def _adjoint(orc, c):
    L, y, chi, st, cfg = _forward(orc, c)
    gy, gchi = loss_grads(y, chi, c["yt"], c["chit"])
    gs, gchi0, glam, _ = orc.refine_adjoint(c["s"], c["car"], cfg, st, chi, gchi, gy, adj_iters=4000, adj_tol=1e-13)
    return L, gs, gchi0, glam


@pytest.fixture(scope="module")
def cases(orc64):
    return crops(orc64, n=16, h=64, w=128)


def test_adjoint_matches_central_differences(orc64, cases):
    """16 crops of 128 x 64, FP64, central differences with h = 1e-6: the
    log-lambda derivatives and the derivatives along the gradients w.r.t. S
    and chi0 within 1e-5 relative; along random directions within 1e-4."""
    rng = np.random.default_rng(5)
    worst, worst_rand = 0.0, 0.0

    def fd(**kw_pm):
        (kp, vp), = kw_pm["p"].items()
        (km, vm), = kw_pm["m"].items()
        return (_forward(orc64, c, **{kp: vp})[0] - _forward(orc64, c, **{km: vm})[0]) / (2 * FD_H)

    for c in cases:
        _, gs, gchi0, glam = _adjoint(orc64, c)
        e = np.exp([FD_H, -FD_H])
        for k, name in enumerate(("lam_y", "lam_c")):
            d = fd(p={name: e[0]}, m={name: e[1]})
            worst = max(worst, abs(d - glam[k]) / abs(d))
        for dirs, acc in ((gs / np.abs(gs).max(), "grad"), (rng.normal(size=c["s"].shape), "rand")):
            d = fd(p={"s": c["s"] + FD_H * dirs}, m={"s": c["s"] - FD_H * dirs})
            r = abs(d - np.sum(gs * dirs)) / abs(d)
            worst, worst_rand = (max(worst, r), worst_rand) if acc == "grad" else (worst, max(worst_rand, r))
        rdir = rng.normal(size=c["s"].shape) + 1j * rng.normal(size=c["s"].shape)
        for dirs, acc in ((gchi0 / np.abs(gchi0).max(), "grad"), (rdir, "rand")):
            d = fd(p={"chi0": c["chi0"] + FD_H * dirs}, m={"chi0": c["chi0"] - FD_H * dirs})
            r = abs(d - np.sum(gchi0.real * dirs.real + gchi0.imag * dirs.imag)) / abs(d)
            worst, worst_rand = (max(worst, r), worst_rand) if acc == "grad" else (worst, max(worst_rand, r))
    assert worst <= 1e-5, worst
    assert worst_rand <= 1e-4, worst_rand


def test_adjoint_scale_invariance(orc64, cases):
    """The weights do not depend on lambda, so scaling lambda_Y and lambda_C
    together leaves the solution unchanged: dL/dlog lam_Y + dL/dlog lam_C = 0."""
    for c in cases[:4]:
        _, _, _, glam = _adjoint(orc64, c)
        assert abs(glam[0] + glam[1]) <= 1e-9 * max(abs(glam[0]), 1e-12)


# ---------------------------------------------------------------- unrolled autograd
torch = pytest.importorskip("torch")


def _dx(a):
    return torch.cat([a[:, 1:] - a[:, :-1], torch.zeros_like(a[:, :1])], 1)


def _dy(a):
    return torch.cat([a[1:] - a[:-1], torch.zeros_like(a[:1])], 0)


def _dxt(v):
    v = torch.cat([v[:, :-1], torch.zeros_like(v[:, :1])], 1)   # the last column's difference is not there
    return torch.cat([torch.zeros_like(v[:, :1]), v[:, :-1]], 1) - v


def _dyt(v):
    v = torch.cat([v[:-1], torch.zeros_like(v[:1])], 0)
    return torch.cat([torch.zeros_like(v[:1]), v[:-1]], 0) - v


def _unrolled(s, cr, ci, xr0, xi0, lam_y, lam_c, aniso, iters, eps=0.5, eps_c=1.0, kc=0.25, outer=4):
    """The 2-D solve written from the energy (E = lam_Y sum w |D Y|^2 +
    lam_C sum w_c |D chi|^2, lagged Charbonnier weights), Fletcher-Reeves CG
    with the gradient recomputed each step, unrolled for autograd."""
    mu_h, mu_v = lam_c * aniso, lam_c

    def grad(xr, xi, wx, wy, wcx, wcy, data):
        yc = (s if data else 0.0) - (xr * cr - xi * ci)
        img = 2 * lam_y * (_dxt(wx * _dx(yc)) + _dyt(wy * _dy(yc)))
        pr = 2 * (mu_h * _dxt(wcx * _dx(xr)) + mu_v * _dyt(wcy * _dy(xr)))
        pi = 2 * (mu_h * _dxt(wcx * _dx(xi)) + mu_v * _dyt(wcy * _dy(xi)))
        return -img * cr + pr, img * ci + pi

    xr, xi = xr0, xi0
    for _ in range(outer):
        yc = s - (xr * cr - xi * ci)
        gx, gy = _dx(yc), _dy(yc)
        mx2 = _dx(xr) ** 2 + _dx(xi) ** 2
        my2 = _dy(xr) ** 2 + _dy(xi) ** 2
        wx = eps / torch.sqrt(gx * gx + eps * eps)
        wy = eps / torch.sqrt(gy * gy + eps * eps)
        wcx = eps_c / torch.sqrt(mx2 + kc * gx * gx + eps_c * eps_c)
        wcy = eps_c / torch.sqrt(my2 + kc * gy * gy + eps_c * eps_c)
        g_r, g_i = grad(xr, xi, wx, wy, wcx, wcy, True)
        d_r, d_i = -g_r, -g_i
        gg = (g_r * g_r + g_i * g_i).sum()
        for _ in range(iters):
            hr, hi = grad(d_r, d_i, wx, wy, wcx, wcy, False)   # Hessian . d
            hh = 0.5 * (d_r * hr + d_i * hi).sum()             # the quadratic form the reference calls curv
            gd = (g_r * d_r + g_i * d_i).sum()
            alpha = -0.5 * gd / hh
            xr, xi = xr + alpha * d_r, xi + alpha * d_i
            g_r, g_i = grad(xr, xi, wx, wy, wcx, wcy, True)
            ggn = (g_r * g_r + g_i * g_i).sum()
            beta = ggn / gg
            d_r, d_i = -g_r + beta * d_r, -g_i + beta * d_i
            gg = ggn
    y = s - (xr * cr - xi * ci)
    return y, xr, xi


def test_adjoint_matches_unrolled_autograd(orc64):
    """On small crops the unrolled FP64 torch forward converges, and autograd
    through it gives the implicit gradient the adjoint computes."""
    for c in crops(orc64, n=2, h=20, w=28, seed0=321):
        _, gs, gchi0, glam = _adjoint(orc64, c)
        t = lambda a: torch.tensor(a, dtype=torch.float64, requires_grad=True)  # noqa: E731
        s, xr0, xi0 = t(c["s"]), t(c["chi0"].real), t(c["chi0"].imag)
        ly, lc = t(0.0), t(0.0)   # log lambdas
        cr = torch.tensor(c["car"].real, dtype=torch.float64)
        ci = torch.tensor(c["car"].imag, dtype=torch.float64)
        y, xr, xi = _unrolled(s, cr, ci, xr0, xi0, torch.exp(ly), torch.exp(lc), 0.7, iters=150)
        yt = torch.tensor(c["yt"])
        ctr, cti = torch.tensor(c["chit"].real), torch.tensor(c["chit"].imag)
        n = y.numel()
        L = ((y - yt) ** 2).sum() / n + ((xr - ctr) ** 2 + (xi - cti) ** 2).sum() / n
        g_s, g_xr, g_xi, g_ly, g_lc = torch.autograd.grad(L, [s, xr0, xi0, ly, lc])
        rel = lambda a, b: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))  # noqa: E731
        assert rel(gs, g_s.numpy()) <= 1e-6
        assert rel(gchi0.real, g_xr.numpy()) <= 1e-6 and rel(gchi0.imag, g_xi.numpy()) <= 1e-6
        assert abs(glam[0] - float(g_ly)) <= 1e-6 * abs(float(g_ly))
        assert abs(glam[1] - float(g_lc)) <= 1e-6 * abs(float(g_lc))
