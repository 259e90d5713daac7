"""GPU: the benchmark workloads' own shapes, checked against the oracle.

* Motion estimation on 8 consecutive NTSC fields (7 pairs) of the moving
  scene generator: integer SAD over 16 x 8 blocks with a +-16 x +-4 search,
  and EstimateMotion (32-sample tiles, +-16 search) on the same fields; both
  bit-exact with the oracle (motion.cpp:88-230 for EstimateMotion, the
  brute-force SAD restatement for the block matcher).
* Streaming windows as the streaming benchmark runs them: per window, the
  SAD block matching of its 7 consecutive pairs and one batched decode of
  its 8 fields (lock-in, holographic init, IRLS 4 x CG <= 40 to relative
  residual 1e-6).  One window per noise level sigma in {0.2, 1, 3} IRE,
  chroma saturation up to 75 IRE, fast motion (rectangles up to 12
  px/field), and a PAL window.  Every field: the lossless split
  Y + Re[chi c] = S (lossless_identity_test.cpp:73-74); two fields per
  window against the FP64 oracle within TOL_IRE.
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

from conftest import TOL_IRE

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import gpu, synth  # noqa: E402
import oracle as O  # noqa: E402
from test_gpu_fullsize import lossless_err  # noqa: E402

DEV = "cuda:0"
CFG_STREAM = dict(cg_iterations=160, irls_outer=4, fast=0, cg_tol=1e-6)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def pool():
    return cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1))


# ---------------------------------------------------------------- 8-field motion
def test_motion_8_fields_sad(engine, orc64):
    raws, _ = synth.motion_sequence(71, n_fields=8)
    assert len(raws) == 8
    refs, curs = raws[:-1], raws[1:]
    vy, vx, sad = engine.block_match_sad_u16([dev(r.view(np.int16)) for r in refs],
                                             [dev(c.view(np.int16)) for c in curs])
    with pool() as ex:
        want = list(ex.map(lambda k: orc64.block_match_sad_u16(refs[k], curs[k]), range(7)))
    moving = 0
    for k, (wy, wx, ws) in enumerate(want):
        assert np.array_equal(vy[k].cpu().numpy(), wy), k
        assert np.array_equal(vx[k].cpu().numpy(), wx), k
        assert np.array_equal(sad[k].cpu().numpy(), ws), k
        moving += int(np.count_nonzero(wy) + np.count_nonzero(wx))
    assert moving > 0  # the sequence moves: the vectors are not all zero


def test_motion_8_fields_estimate_motion(engine, orc32):
    seq, _ = synth.motion_sequence(83, n_fields=8)
    g = synth.NTSC
    fal, a0 = g.first_active_field_line, g.active_video_start
    ys = [np.ascontiguousarray(synth.ire_from_raw(r)[fal:fal + g.lines, a0:a0 + g.width]) for r in seq]
    dy, dx, cfd = engine.estimate_motion([dev(y) for y in ys[:-1]], [dev(y) for y in ys[1:]], tile=32, search=16)
    with pool() as ex:
        want = list(ex.map(lambda k: orc32.estimate_motion(ys[k], ys[k + 1], 32, 16), range(7)))
    for k, (wdy, wdx, wcf) in enumerate(want):
        assert np.array_equal(dy[k].cpu().numpy(), wdy), k
        assert np.array_equal(dx[k].cpu().numpy(), wdx), k
        assert np.array_equal(cfd[k].cpu().numpy(), wcf), k


# ---------------------------------------------------------------- streaming windows
WINDOWS = [  # (standard, noise sigma IRE, chroma saturation IRE, fast motion, seed)
    ("ntsc", 0.2, 75.0, False, 9100),
    ("ntsc", 1.0, 50.0, True, 9101),
    ("ntsc", 3.0, 75.0, False, 9102),
    ("pal", 1.0, 75.0, True, 9103),
]


@pytest.mark.parametrize("std,sigma,sat,fast,seed", WINDOWS,
                         ids=[f"{w[0]}-sigma{w[1]}-sat{int(w[2])}{'-fast' if w[3] else ''}" for w in WINDOWS])
def test_streaming_window(engine, orc64, orc32, std, sigma, sat, fast, seed):
    g = synth.PAL if std == "pal" else synth.NTSC
    raws, _ = synth.motion_sequence(seed, n_fields=8, g=g, noise_ire=sigma, max_rect_v=12 if fast else 8,
                                    saturation=sat)
    d = [dev(r.view(np.int16)) for r in raws]
    vy, vx, sad = engine.block_match_sad_u16(d[:-1], d[1:])
    p = gpu.SolverParams.make(**CFG_STREAM)
    ys, chis, ths = engine.decode_fields_u16(d, [g] * 8, p, want_theta=True)
    Y = [y.cpu().numpy() for y in ys]
    X = [c.cpu().numpy() for c in chis]
    T = [t.cpu().numpy() for t in ths]
    assert all(np.isfinite(y).all() and np.isfinite(x).all() for y, x in zip(Y, X))
    worst = max(lossless_err(orc32, raws[k], g, Y[k], X[k], T[k]) for k in range(8))
    print(f"{std} sigma {sigma} sat {sat}{' fast' if fast else ''}: lossless split max {worst:.2e} IRE", flush=True)
    assert worst <= 1e-3
    cfg = O.Cfg.make(is_pal=int(g.uses_vswitch), **CFG_STREAM)
    sampled = (1, 6)
    with pool() as ex:
        sads = ex.map(lambda k: orc64.block_match_sad_u16(raws[k], raws[k + 1]), range(7))
        wants = ex.map(lambda k: orc64.decode_field(synth.ire_from_raw(raws[k]), g, cfg), sampled)
        for k, (wy, wx, ws) in enumerate(sads):
            assert np.array_equal(vy[k].cpu().numpy(), wy) and np.array_equal(vx[k].cpu().numpy(), wx), k
            assert np.array_equal(sad[k].cpu().numpy(), ws), k
        for k, want in zip(sampled, wants):
            ey = np.abs(Y[k] - want["y"]).max()
            eu = np.abs(X[k][..., 1] - want["chi"].imag).max()
            ev = np.abs(X[k][..., 0] - want["chi"].real).max()
            print(f"  field {k} vs FP64 oracle: |dY| {ey:.2e} |dU| {eu:.2e} |dV| {ev:.2e} IRE", flush=True)
            assert max(ey, eu, ev) <= TOL_IRE, (k, ey, eu, ev)
