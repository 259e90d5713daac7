"""Pins of the selective 3-D window's oracle (oracle/selective.py; hvd-core
frame_bridge.cpp:385-720).

* AmbiguousBoxes against hvd-core's own, compiled in place from the
  reference checkout (oracle/_ref/libhvdref_fb.so): the same boxes, or the
  same "none" / "everywhere" decision, on scenes built to reach every branch
  (one band, several, more than four merged, column runs split and merged by
  their margins, a one-tile stripe, the excluded border ring, the area cap,
  clean and globally ambiguous fields) and on random content;
* |chi| as numpy's float32 hypot (the libm hypotf that std::abs of
  std::complex<float> calls);
* the count-normalised box mean against a direct FP64 mean over the
  edge-truncated 8-sample windows;
* the feather blend by what fixes it: untouched outside the box and on its
  edge, the 3-D result in the box's core, linear ramps of min(16, bh / 3)
  rows and min(48, bw / 3) columns in between.
"""
import numpy as np
import pytest

import selective as SEL
from selective_cases import SCENES, noise_scene, scene

F = np.float32


@pytest.fixture(scope="module")
def ref():
    try:
        return SEL.RefBoxes()
    except FileNotFoundError:
        pytest.skip("oracle/_ref/libhvdref_fb.so not built (needs the reference checkout)")


def test_cabs_is_libm_hypotf():
    rng = np.random.default_rng(0)
    a = (rng.normal(0, 1, 200000) * 10.0 ** rng.uniform(-6, 6, 200000)).astype(F)
    b = (rng.normal(0, 1, 200000) * 10.0 ** rng.uniform(-6, 6, 200000)).astype(F)
    assert np.array_equal(SEL.cabs(a + 1j * b), np.hypot(a, b))


def test_box_mean_is_the_count_normalised_window_mean():
    rng = np.random.default_rng(1)
    a = rng.normal(0, 10, (37, 53)).astype(F)
    m = SEL.box_mean(a)
    h, w = a.shape
    tmp = np.zeros((h, w))
    for y in range(h):
        ys = [y + d for d in range(-4, 4) if 0 <= y + d < h]
        tmp[y] = a[ys].astype(np.float64).mean(axis=0)
    ref = np.zeros((h, w))
    for x in range(w):
        xs = [x + d for d in range(-4, 4) if 0 <= x + d < w]
        ref[:, x] = tmp[:, xs].mean(axis=1)
    assert np.abs(m - ref).max() < 1e-4
    # a constant plane stays constant, edges included (the count, not the window, normalises)
    assert np.array_equal(SEL.box_mean(np.full((20, 30), 7.0, F)), np.full((20, 30), 7.0, F))


@pytest.mark.parametrize("name", sorted(SCENES))
def test_boxes_match_hvdcore(ref, name):
    Y, chi = scene(name)
    mine = SEL.ambiguous_boxes(Y, chi, 0.45)
    theirs = ref.boxes(Y, chi, 0.45)
    assert mine == theirs, (mine, theirs)


def test_scenes_reach_every_branch():
    """The scenes above are meant to reach each decision; this states which."""
    got = {n: SEL.ambiguous_boxes(*scene(n), 0.45) for n in SCENES}
    assert got["clean"] == [] and got["border_only"] == [] and got["stripe"] == []
    assert got["global"] == [SEL.Box()] and got["overflow"] == [SEL.Box()]
    assert len(got["one_band"]) == 1 and got["one_band"][0].valid
    assert len(got["three_bands"]) == 3
    assert len({(b.fy0, b.fy1) for b in got["six_bands_merged"]}) == 4
    assert len(got["two_runs"]) == 2 and len(got["runs_merged_by_margin"]) == 1
    for b in got["one_band"] + got["three_bands"]:
        assert b.fy0 % 2 == 0 and b.x0 >= 0


@pytest.mark.parametrize("seed", range(12))
def test_boxes_match_hvdcore_on_noise(ref, seed):
    h, w = [(242, 760), (288, 922), (80, 160)][seed % 3]
    Y, chi = noise_scene(h, w, seed)
    for area in (0.45, 0.1):
        assert SEL.ambiguous_boxes(Y, chi, area) == ref.boxes(Y, chi, area)


@pytest.mark.parametrize("bh,bw", [(64, 176), (40, 97), (5, 7)])
def test_feather_blend(bh, bw):
    rng = np.random.default_rng(bh + bw)
    h, w = bh + 30, bw + 50
    b = SEL.Box(10, 10 + bh, 20, 20 + bw, True)
    Yd = rng.normal(0, 10, (h, w)).astype(F)
    Cd = (rng.normal(0, 5, (h, w)) + 1j * rng.normal(0, 5, (h, w))).astype(np.complex64)
    Y3 = rng.normal(0, 10, (bh, bw)).astype(F)
    C3 = (rng.normal(0, 5, (bh, bw)) + 1j * rng.normal(0, 5, (bh, bw))).astype(np.complex64)
    Y0, C0 = Yd.copy(), Cd.copy()
    SEL.feather_blend(Yd, Cd, Y3, C3, b)
    out = np.ones((h, w), bool)
    out[b.fy0:b.fy1, b.x0:b.x1] = False
    assert np.array_equal(Yd[out], Y0[out]) and np.array_equal(Cd[out], C0[out])
    hy, hx = min(16, bh // 3), min(48, bw // 3)
    by, bc = Yd[b.fy0:b.fy1, b.x0:b.x1], Cd[b.fy0:b.fy1, b.x0:b.x1]
    y0, c0 = Y0[b.fy0:b.fy1, b.x0:b.x1], C0[b.fy0:b.fy1, b.x0:b.x1]
    # the box's first row and column carry weight 0 where a ramp exists
    if hy > 0:
        assert np.array_equal(by[0], y0[0]) and np.array_equal(by[-1], y0[-1])
    if hx > 0:
        assert np.array_equal(by[:, 0], y0[:, 0]) and np.array_equal(bc[:, -1], c0[:, -1])
    # the core takes the 3-D result
    core = (slice(hy, bh - hy), slice(hx, bw - hx))
    assert np.array_equal(by[core], Y3[core]) and np.array_equal(bc[core], C3[core])
    # the ramps are linear in the distance from the edge
    ry = np.ones(bh)
    ry[:hy] = np.arange(hy) / max(hy, 1)
    if hy:
        ry[bh - hy:] = (bh - 1 - np.arange(bh - hy, bh)) / hy
    rx = np.ones(bw)
    rx[:hx] = np.arange(hx) / max(hx, 1)
    if hx:
        rx[bw - hx:] = (bw - 1 - np.arange(bw - hx, bw)) / hx
    m = ry[:, None] * rx[None, :]
    want = y0.astype(np.float64) * (1 - m) + Y3.astype(np.float64) * m
    assert np.abs(by - want).max() < 1e-4 * (1 + np.abs(want).max())


def _sel_window():
    from sequence_cases import window
    return window(lines=128, width=384, nf=6, amb_region=(150, 220, 220, 330), seed=3)


SR = 4.0 * 315e6 / 88.0


@pytest.mark.parametrize("max_area", [0.75, 0.45])
def test_selective_composition(orc32, max_area):
    """DecodeWindowSelective's composition on the oracle: with boxes, the 2-D
    decode outside them and the crops' 3-D decodes in their cores; with the
    area cap exceeded, the full 3-D window.  (A clean field keeps the 2-D
    decode: no box, the first return.)"""
    import sequence3d as S3
    fields, par, inits, _ = _sel_window()
    cfg = S3.WindowCfg(cg_iterations=20, temporal_strength=1.0)
    init = lambda s, c: orc32.holographic_init(s, c, SR)  # noqa: E731
    Y, C, boxes = SEL.decode_window_selective(orc32, fields, par, inits, None, cfg, init, max_area=max_area)
    if max_area < 0.5:
        assert boxes == [SEL.Box()]
        y3, c3 = S3.decode_window_with_inits(orc32, fields, par, inits, cfg)
        assert all(np.array_equal(a, b) for a, b in zip(Y, y3)) and all(np.array_equal(a, b) for a, b in zip(C, c3))
        return
    assert len(boxes) == 2 and all(b.valid for b in boxes)
    y2, c2 = S3.decode_window_with_inits(orc32, fields, par, inits, S3.WindowCfg(cg_iterations=20,
                                                                                temporal_strength=-1.0))
    out = np.ones(y2[0].shape, bool)
    for b in boxes:
        out[b.fy0:b.fy1, b.x0:b.x1] = False
    for j in range(len(Y)):
        assert np.array_equal(Y[j][out], y2[j][out]) and np.array_equal(C[j][out], c2[j][out])
    for b in boxes:
        crop = [(s[b.fy0:b.fy1, b.x0:b.x1], c[b.fy0:b.fy1, b.x0:b.x1]) for s, c in fields]
        ci = [init(s, c) for s, c in crop]
        y3, c3 = S3.decode_window_with_inits(orc32, crop, par, [c for _, c in ci], cfg, inits_y=[y for y, _ in ci])
        bh, bw = b.fy1 - b.fy0, b.x1 - b.x0
        hy, hx = min(16, bh // 3), min(48, bw // 3)
        core = (slice(b.fy0 + hy, b.fy1 - hy), slice(b.x0 + hx, b.x1 - hx))
        for j in range(len(Y)):
            assert np.array_equal(Y[j][core], y3[j][hy:bh - hy, hx:bw - hx])
            assert np.array_equal(C[j][core], c3[j][hy:bh - hy, hx:bw - hx])
