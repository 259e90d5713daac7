"""GPU: the benchmark workloads at their full sizes, checked through properties
that do not depend on size, plus sampled fields against the FP64 oracle.

* 64 NTSC 910x263 fields in one batched decode (IRLS 4 x CG 25, every
  iteration run), with the launch shape the autotuner picks for that batch;
* 32 NTSC + 32 PAL fields mixed in one call (relative residual 1e-6 or 40
  iterations per stage);
* the largest plane the solver accepts (992 x 2048).

Properties: every output finite; the lossless split Y + Re[chi c] = S within
1e-3 IRE (lossless_identity_test.cpp:73-74, with c the carrier the decode
uses); a field's result is the same bits whether it is decoded
in the batch or alone (same strip height); sampled fields within TOL_IRE of
the FP64 oracle; every field of the mixed batch against the FP64 oracle
(per-field max |dY|, |dU|, |dV| printed).
"""
import numpy as np
import pytest

from conftest import TOL_IRE

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import gpu, synth  # noqa: E402
import oracle as O  # noqa: E402

DEV = "cuda:0"
CFG0 = dict(cg_iterations=100, irls_outer=4, fast=0, cg_tol=1e-12)
CFG2 = dict(cg_iterations=160, irls_outer=4, fast=0, cg_tol=1e-6)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.fixture()
def eng():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    e = gpu.Engine(0)
    yield e
    e.close()


def lossless_err(orc32, raw, g, y, chi, theta):
    """max |Y + Re[chi c] - S| with c formed as the decode forms it: the row
    factor e^{i theta} of the decoded theta (rounded to complex64) times the
    exact quarter-turn column factor e^{i pi/2 (a0 + x)}, and c -> -conj(c) on
    V-switched lines (ntsc_geometry.cpp:75-93)."""
    field = synth.ire_from_raw(raw)
    a0, fal = g.active_video_start, g.first_active_field_line
    s = field[fal:fal + g.lines, a0:a0 + g.width].astype(np.float64)
    th = theta[fal:fal + g.lines].astype(np.float64)
    row = np.exp(1j * th).astype(np.complex64).astype(np.complex128)
    col = (1j) ** ((a0 + np.arange(g.width)) % 4)
    car = row[:, None] * col[None, :]
    if g.uses_vswitch:
        vs = orc32.burst_lockin_phase_pal(field, g)[1][fal:fal + g.lines]
        car[vs < 0] = -np.conj(car[vs < 0])
    z = chi[..., 0].astype(np.float64) + 1j * chi[..., 1]
    return float(np.abs(y + (z * car).real - s).max())


def test_batch64_ntsc_full(eng, orc64, orc32):
    g = synth.NTSC
    raws = [synth.ntsc_field(500 + k)[0] for k in range(64)]
    rows, gpx, _ = eng.autotune_solver(g.lines, g.width, False, 64)
    print(f"batch of 64: tuned strip rows {rows}, group {gpx} px", flush=True)
    p = gpu.SolverParams.make(**CFG0)
    ys, chis, ths = eng.decode_fields_u16([dev(r.view(np.int16)) for r in raws], [g] * 64, p, want_theta=True)
    Y = [y.cpu().numpy() for y in ys]
    X = [c.cpu().numpy() for c in chis]
    T = [t.cpu().numpy() for t in ths]
    assert all(np.isfinite(y).all() and np.isfinite(x).all() for y, x in zip(Y, X))
    worst = max(lossless_err(orc32, raws[k], g, Y[k], X[k], T[k]) for k in range(0, 64, 7))
    print(f"lossless split, every 7th field: max {worst:.2e} IRE", flush=True)
    assert worst <= 1e-3
    # alone, at the batch's strip height: the same bits
    eng.set_solver_tuning(rows, 0)
    for k in (0, 37, 63):
        y1, c1, _ = eng.decode_fields_u16([dev(raws[k].view(np.int16))], [g], p)
        assert np.array_equal(y1[0].cpu().numpy(), Y[k]) and np.array_equal(c1[0].cpu().numpy(), X[k])
    # one field against the FP64 oracle
    want = orc64.decode_field(synth.ire_from_raw(raws[37]), g, O.Cfg.make(**CFG0))
    ey = np.abs(Y[37] - want["y"]).max()
    ec = np.abs(X[37][..., 0] + 1j * X[37][..., 1] - want["chi"]).max()
    print(f"field 37 vs FP64 oracle: |dY| {ey:.2e}, |dchi| {ec:.2e} IRE", flush=True)
    assert ey <= TOL_IRE and ec <= TOL_IRE


def test_mixed_32_32_full(eng, orc64, orc32):
    raws, gs = [], []
    for k in range(64):
        pal = k % 2 == 1
        raws.append((synth.pal_field if pal else synth.ntsc_field)(700 + k)[0])
        gs.append(synth.PAL if pal else synth.NTSC)
    p = gpu.SolverParams.make(**CFG2)
    ys, chis, ths = eng.decode_fields_u16([dev(r.view(np.int16)) for r in raws], gs, p, want_theta=True)
    Y = [y.cpu().numpy() for y in ys]
    X = [c.cpu().numpy() for c in chis]
    T = [t.cpu().numpy() for t in ths]
    assert all(np.isfinite(y).all() and np.isfinite(x).all() for y, x in zip(Y, X))
    worst = max(lossless_err(orc32, raws[k], gs[k], Y[k], X[k], T[k]) for k in range(64))
    print(f"mixed batch lossless split: max {worst:.2e} IRE", flush=True)
    assert worst <= 1e-3
    # every field against the FP64 oracle: Y, U = -Im chi, V = Re chi
    import concurrent.futures as cf
    import os

    def oracle_field(k):
        cfg = O.Cfg.make(is_pal=int(gs[k].uses_vswitch), **CFG2)
        return orc64.decode_field(synth.ire_from_raw(raws[k]), gs[k], cfg)

    with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        wants = list(ex.map(oracle_field, range(64)))
    errs = []
    for k, want in enumerate(wants):
        ey = float(np.abs(Y[k] - want["y"]).max())
        eu = float(np.abs(X[k][..., 1] - want["chi"].imag).max())
        ev = float(np.abs(X[k][..., 0] - want["chi"].real).max())
        errs.append((ey, eu, ev))
        print(f"field {k:2d} ({'PAL ' if gs[k].uses_vswitch else 'NTSC'}): max |dY| {ey:.2e} |dU| {eu:.2e} "
              f"|dV| {ev:.2e} IRE", flush=True)
    E = np.array(errs)
    print(f"all 64 fields: worst |dY| {E[:, 0].max():.2e} |dU| {E[:, 1].max():.2e} |dV| {E[:, 2].max():.2e} IRE",
          flush=True)
    assert E.max() <= TOL_IRE, np.unravel_index(np.argmax(E), E.shape)


def test_largest_plane(eng, orc64):
    h, w = 1024 - 32, 2048
    r, c = np.mgrid[0:h, 0:w].astype(np.float64)
    car = np.exp(1j * (0.5 * np.pi * c + 0.41 * r)).astype(np.complex64).astype(np.complex128)
    chi_true = 12 * np.sin(0.05 * (c - 2 * r)) - 7j * np.cos(0.03 * (c + r))
    s = (30 + 15 * np.sin(0.011 * c + 0.02 * r) + (chi_true * car).real).astype(np.float32)
    chi0 = (chi_true + 2.0 * np.sin(0.7 * c) * np.cos(0.3 * r)).astype(np.complex64).astype(np.complex128)
    cfg = dict(cg_iterations=16, irls_outer=2, fast=0, cg_tol=1e-12, chroma_aniso=0.0)
    dchi = dev(np.stack([chi0.real, chi0.imag], -1).astype(np.float32))
    dy = torch.zeros((h, w), device=DEV)
    dcar = dev(np.stack([car.real, car.imag], -1).astype(np.float32))
    eng.variational_refine([gpu.Engine.plane(dev(s), dchi, dy, carrier=dcar)], gpu.SolverParams.make(**cfg))
    y = dy.cpu().numpy()
    x = dchi.cpu().numpy()
    assert np.isfinite(y).all() and np.isfinite(x).all()
    z = x[..., 0].astype(np.float64) + 1j * x[..., 1]
    assert np.abs(y + (z * car).real - s).max() <= 1e-3
    y_ref, chi_ref, _, _ = orc64.variational_refine(s, car, chi0, O.Cfg.make(**cfg))
    ey, ec = np.abs(y - y_ref).max(), np.abs(z - chi_ref).max()
    print(f"992x2048 plane vs FP64 oracle: |dY| {ey:.2e}, |dchi| {ec:.2e} IRE", flush=True)
    assert ey <= TOL_IRE and ec <= TOL_IRE
