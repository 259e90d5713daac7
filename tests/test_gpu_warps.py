"""GPU parity of the temporal neighbours' warps (temporal.cpp:19-287) with
the oracle's FP32 restatement, which is bit-exact with hvd-core's own
functions (test_oracle_warps.py): the GPU results are held to the same bits
(every operation is written in hvd-core's float order, contraction off)."""
import numpy as np
import pytest

from warp_cases import motion_field, planes

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TILE = 32
SHAPES = [(263, 758, 9, 24), (40, 70, 2, 3), (96, 128, 3, 4), (17, 5, 1, 1), (313, 922, 10, 29)]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(DEV)


def cdev(z):
    return dev(np.stack([z.real, z.imag], -1))


def chost(t):
    a = t.cpu().numpy()
    return (a[..., 0] + 1j * a[..., 1]).astype(np.complex64)


@pytest.mark.parametrize("h,w,th,tw", SHAPES)
def test_vectors_and_warps(engine, orc32, h, w, th, tw):
    mdy, mdx, _ = motion_field(th, tw, h + w)
    comp, car, _ = planes(h, w, 3)
    dy, dx = engine.vectors_per_pixel(dev(mdy), dev(mdx), TILE, h, w)
    ody, odx = orc32.vectors_per_pixel(mdy, mdx, TILE, h, w)
    np.testing.assert_array_equal(dy.cpu().numpy(), ody.astype(np.float32))
    np.testing.assert_array_equal(dx.cpu().numpy(), odx.astype(np.float32))
    np.testing.assert_array_equal(engine.warp_by_tiles(dev(comp), dy, dx).cpu().numpy(),
                                  orc32.warp_by_tiles(comp, ody, odx).astype(np.float32))
    np.testing.assert_array_equal(chost(engine.warp_by_tiles(cdev(car), dy, dx)),
                                  orc32.warp_by_tiles(car, ody, odx).astype(np.complex64))
    if h >= 2:
        for off in (0.0, 0.5, -0.5):
            got = engine.warp_bilinear(dev(comp), dy[:-1].contiguous(), dx[:-1].contiguous(), off).cpu().numpy()
            np.testing.assert_array_equal(got, orc32.warp_bilinear(comp, ody[:-1], odx[:-1], off).astype(np.float32))


@pytest.mark.parametrize("h,w,th,tw", SHAPES)
@pytest.mark.parametrize("fast", [False, True])
def test_confidence_and_compensation(engine, orc32, h, w, th, tw, fast):
    mdy, mdx, conf = motion_field(th, tw, 7 * h + w)
    comp, car, _ = planes(h, w, 4)
    np.testing.assert_array_equal(engine.upsample_confidence(dev(conf), TILE, h, w, fast).cpu().numpy(),
                                  orc32.upsample_confidence(conf, TILE, h, w, fast).astype(np.float32))
    oc, ocar, ocf = engine.motion_compensate_prev(dev(comp), cdev(car), dev(mdy), dev(mdx), dev(conf), TILE, fast)
    rc, rcar, rcf = orc32.motion_compensate_prev(comp, car, mdy, mdx, conf, TILE, fast)
    np.testing.assert_array_equal(oc.cpu().numpy(), rc.astype(np.float32))
    np.testing.assert_array_equal(chost(ocar), rcar.astype(np.complex64))
    np.testing.assert_array_equal(ocf.cpu().numpy(), rcf.astype(np.float32))


@pytest.mark.parametrize("r", [0, 1, 3, 8])
def test_complex_coherence(engine, orc32, r):
    _, _, z1 = planes(60, 90, 5)
    _, _, z2 = planes(60, 90, 6)
    got = engine.complex_coherence(cdev(z1), cdev(z2), r).cpu().numpy()
    np.testing.assert_array_equal(got, orc32.complex_coherence(z1, z2, r).astype(np.float32))


def test_warp_argument_errors(engine):
    a = dev(np.zeros((4, 4)))
    from hvd_amd import gpu
    with pytest.raises(gpu.HvdGpuError, match="in place"):
        engine._check(engine.lib.hvd_gpu_warp_by_tiles(engine.ctx, a.data_ptr(), 4, 4, 1, a.data_ptr(), a.data_ptr(),
                                                       a.data_ptr(), None), "warp_by_tiles")
    with pytest.raises(gpu.HvdGpuError, match="motion"):
        engine.upsample_confidence(dev(np.zeros((0, 3))), TILE, 8, 8)
