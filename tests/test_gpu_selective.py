"""GPU parity of the selective 3-D window (hvd-core frame_bridge.cpp:385-720).

* hvd_gpu_ambiguous_boxes against the oracle (oracle/selective.py, whose
  boxes are hvd-core's own: tests/test_oracle_selective.py): the tile scores
  bit for bit and the same boxes or decision, on the oracle's scenes;
* hvd_gpu_decode_window with selective_3d against the oracle's
  DecodeWindowSelective on a window with two ambiguous regions (two boxes)
  and with the area cap exceeded (the whole window in 3-D): the same
  decision, Y and chi held to the bar of the 3-D window tests (the crops'
  3-D passes take discrete decisions that can flip on sub-tolerance
  differences at a tile: at most 2e-4 of the samples beyond TOL_IRE, none
  beyond 0.15 IRE).
"""
import numpy as np
import pytest

from conftest import TOL_IRE
from selective_cases import SCENES, noise_scene, scene

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import gpu  # noqa: E402
import selective as SEL  # noqa: E402
import sequence3d as S3  # noqa: E402

DEV = "cuda:0"
SR = 4.0 * 315e6 / 88.0


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(DEV)


def cdev(z):
    return dev(np.stack([np.asarray(z).real, np.asarray(z).imag], -1))


def chost(t):
    a = t.cpu().numpy()
    return a[..., 0] + 1j * a[..., 1]


def _as_tuples(boxes):
    if boxes == [SEL.Box()]:
        return None
    return [(b.fy0, b.fy1, b.x0, b.x1) for b in boxes]


@pytest.mark.parametrize("name", sorted(SCENES) + ["noise0", "noise1", "noise2"])
def test_ambiguous_boxes_match_oracle(engine, name):
    if name.startswith("noise"):
        k = int(name[5:])
        Y, chi = noise_scene(*[(242, 760), (288, 922), (80, 160)][k], k)
    else:
        Y, chi = scene(name)
    print("ambiguous boxes", name, Y.shape, flush=True)
    for area in (0.45, 0.1):
        got, sc = engine.ambiguous_boxes(dev(Y), cdev(chi), area, with_scores=True)
        torch.cuda.synchronize()
        assert np.array_equal(sc, SEL.tile_scores(Y, chi))
        assert got == _as_tuples(SEL.ambiguous_boxes(Y, chi, area))


def _run_gpu(engine, fields, parities, inits, cfg, max_area, pal=False):
    planes, outs = [], []
    for (s, car), chi0 in zip(fields, inits):
        dchi, dy = cdev(chi0), torch.zeros(s.shape, device=DEV)
        planes.append(gpu.Engine.plane(dev(s), dchi, dy, carrier=cdev(car)))
        outs.append((dy, dchi))
    sp = gpu.SolverParams.make(cg_iterations=cfg.cg_iterations, irls_outer=cfg.irls_outer, cg_tol=cfg.cg_tol,
                               fast=int(cfg.fast))
    wp = gpu.WindowParams.make(temporal_strength=cfg.temporal_strength, selective_3d=1, selective_max_area=max_area)
    d = engine.decode_window(planes, parities, pal, sp, wp)
    torch.cuda.synchronize()
    return [y.cpu().numpy() for y, _ in outs], [chost(c) for _, c in outs], d


@pytest.mark.parametrize("pal,max_area", [(False, 0.75), (False, 0.45), (True, 0.75)])
def test_selective_window_matches_oracle(engine, orc32, pal, max_area):
    from sequence_cases import window
    fields, par, inits, _ = window(lines=128, width=384, nf=8 if pal else 6, amb_region=(150, 220, 220, 330), seed=3,
                                   pal=pal)
    cfg = S3.WindowCfg(cg_iterations=20, temporal_strength=1.0)
    ry, rc, boxes = SEL.decode_window_selective(orc32, fields, par, inits, None, cfg,
                                                lambda s, c: orc32.holographic_init(s, c, SR), is_pal=pal,
                                                max_area=max_area)
    print("selective window", "pal" if pal else "ntsc", max_area, boxes, flush=True)
    gy, gc, d = _run_gpu(engine, fields, par, inits, cfg, max_area, pal)
    want = -1 if boxes == [SEL.Box()] else len(boxes)
    assert int(d["selective_boxes"]) == want
    for a, b in list(zip(gy, ry)) + list(zip(gc, rc)):
        e = np.abs(a - b)
        assert (e > TOL_IRE).mean() <= 2e-4 and e.max() <= 0.15, ((e > TOL_IRE).mean(), e.max())
