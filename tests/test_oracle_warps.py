"""The oracle's temporal-neighbour warps (temporal.cpp:19-287:
VectorsPerPixel, WarpByTiles[Complex], WarpBilinearTiles,
UpsampleConfidence[Fast], ComplexCoherence, MotionCompensatePrev), pinned
bit-exact (FP32 restatement) against hvd-core's own functions built from
source (oracle/_ref), and against what the mathematics fixes."""
import numpy as np
import pytest

from warp_cases import motion_field, planes

TILE = 32
SHAPES = [(263, 758, 9, 24), (40, 70, 2, 3), (96, 128, 3, 4), (17, 5, 1, 1)]


@pytest.mark.parametrize("h,w,th,tw", SHAPES)
def test_vectors_and_warps_match_hvd_core(orc32, orc64, ref, h, w, th, tw):
    mdy, mdx, _ = motion_field(th, tw, h + w)
    comp, car, _ = planes(h, w, 3)
    dy, dx = orc32.vectors_per_pixel(mdy, mdx, TILE, h, w)
    rdy, rdx = ref.vectors_per_pixel(mdy, mdx, TILE, h, w)
    assert np.array_equal(dy.astype(np.float32), rdy) and np.array_equal(dx.astype(np.float32), rdx)
    assert np.array_equal(orc32.warp_by_tiles(comp, dy, dx).astype(np.float32),
                          ref.warp_by_tiles(comp, mdy, mdx, TILE))
    assert np.array_equal(orc32.warp_by_tiles(car, dy, dx).astype(np.complex64),
                          ref.warp_by_tiles(car, mdy, mdx, TILE).astype(np.complex64))
    for off in (0.0, 0.5, -0.5):
        assert np.array_equal(orc32.warp_bilinear(comp, dy[:-1], dx[:-1], off).astype(np.float32),
                              ref.warp_bilinear(comp, mdy, mdx, TILE, off, h - 1, w)) or h < 2
    d6y, d6x = orc64.vectors_per_pixel(mdy, mdx, TILE, h, w)
    assert np.abs(d6y - rdy).max() < 1e-5 and np.abs(d6x - rdx).max() < 1e-5


@pytest.mark.parametrize("h,w,th,tw", SHAPES)
@pytest.mark.parametrize("fast", [False, True])
def test_confidence_and_compensation_match_hvd_core(orc32, ref, h, w, th, tw, fast):
    mdy, mdx, conf = motion_field(th, tw, 7 * h + w)
    comp, car, _ = planes(h, w, 4)
    assert np.array_equal(orc32.upsample_confidence(conf, TILE, h, w, fast).astype(np.float32),
                          ref.upsample_confidence(conf, TILE, h, w, fast))
    o = orc32.motion_compensate_prev(comp, car, mdy, mdx, conf, TILE, fast)
    r = ref.motion_compensate_prev(comp, car, mdy, mdx, conf, TILE, fast)
    assert np.array_equal(o[0].astype(np.float32), r[0])
    assert np.array_equal(o[1].astype(np.complex64), r[1].astype(np.complex64))
    assert np.array_equal(o[2].astype(np.float32), r[2])


@pytest.mark.parametrize("r", [1, 3, 8])
def test_coherence_matches_hvd_core(orc32, orc64, ref, r):
    _, _, z1 = planes(60, 90, 5)
    _, _, z2 = planes(60, 90, 6)
    g = orc32.complex_coherence(z1, z2, r)
    assert np.array_equal(g.astype(np.float32), ref.complex_coherence(z1, z2, r))
    assert np.abs(orc64.complex_coherence(z1, z2, r) - g).max() < 1e-5


def test_warp_identities(orc64):
    h, w = 70, 100
    comp, car, z = planes(h, w, 8)
    # zero motion: vectors 0, the warps are the identity
    zero = np.zeros((3, 4))
    dy, dx = orc64.vectors_per_pixel(zero, zero, TILE, h, w)
    assert not dy.any() and not dx.any()
    assert np.array_equal(orc64.warp_by_tiles(comp, dy, dx), comp)
    # (the last row and column sit at the hs - 1.001 clip, a 0.999 blend)
    assert np.array_equal(orc64.warp_bilinear(comp, dy, dx, 0.0)[:-1, :-1], comp[:-1, :-1])
    # uniform integer motion (2, -3): every vector is exactly that, and the
    # warp is a clamped shift out(y, x) = a(y - 2, x + 3)
    dy, dx = orc64.vectors_per_pixel(np.full((3, 4), 2.0), np.full((3, 4), -3.0), TILE, h, w)
    assert np.all(dy == 2.0) and np.all(dx == -3.0)
    yy = np.clip(np.arange(h) - 2, 0, h - 1)[:, None]
    xx = np.clip(np.arange(w) + 3, 0, w - 1)[None, :]
    assert np.array_equal(orc64.warp_by_tiles(comp, dy, dx), comp[yy, xx])
    assert np.array_equal(orc64.warp_by_tiles(car.astype(np.complex128), dy, dx), car.astype(np.complex128)[yy, xx])
    # one isolated outlier tile is snapped to its 3x3 median (0)
    m = np.zeros((5, 5))
    m[2, 2] = 10.0
    dy, _ = orc64.vectors_per_pixel(m, np.zeros((5, 5)), TILE, 160, 160)
    assert not dy.any()
    # tile centres carry the tile's own vector exactly
    m = np.arange(20.0).reshape(4, 5) * 0.1
    dy, _ = orc64.vectors_per_pixel(m, np.zeros((4, 5)), TILE, 128, 160)
    assert np.allclose(dy[16::32, 16::32], m, atol=1e-12)


def test_coherence_and_confidence_identities(orc64):
    _, _, z = planes(50, 80, 9)
    z = z.astype(np.complex128)
    # a constant phase rotation keeps the phasors coherent: gamma = 1 (up to the 1e-6 guard)
    g = orc64.complex_coherence(z, z * np.exp(0.7j), 3)
    assert np.all(g > 1 - 1e-5) and np.all(g <= 1.0)
    # independent noise decorrelates
    rng = np.random.default_rng(1)
    n = rng.normal(size=z.shape) + 1j * rng.normal(size=z.shape)
    assert np.median(orc64.complex_coherence(z, n, 3)) < 0.5
    # constant tile confidence c: c^2 in the interior (BoxBlur divides by 2r + 1)
    u = orc64.upsample_confidence(np.full((4, 5), 0.6), TILE, 128, 160)
    assert np.allclose(u[8:-8, 8:-8], 0.36, atol=1e-12)
    uf = orc64.upsample_confidence(np.full((4, 5), 0.6), TILE, 128, 160, fast=True)
    assert np.allclose(uf, 0.36, atol=1e-12)
