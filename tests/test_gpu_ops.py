"""GPU parity of the component operators the C++ API exposes on their own:
gradients.h's differences and adjoints (hvd_gpu_difference), EstimateNoiseIre,
Fft2d's numpy-convention 2-D DFT, FitTrajectory / TrajectorySnap and
DrizzleFrame.

Pins: the differences are single float subtractions (or a weight times
one), so they must equal numpy's closed forms bit for bit, and each adjoint
must satisfy <D a, b> = <a, D^T b> (gradients_test.cpp:40-56) to FP64
rounding of the FP32 outputs; EstimateNoiseIre and the trajectory fit and
snap equal the FP32 oracle (bit-exact with hvd-core) bit for bit; the DFT
is held against numpy.fft (FP64) to FP32 rounding and round-trips;
DrizzleFrame equals hvd-core's own, compiled from source (oracle/_ref), bit
for bit."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import gpu  # noqa: E402
import oracle as O  # noqa: E402

DEV = "cuda:0"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(DEV)


def cdev(z):
    return dev(np.stack([z.real, z.imag], -1))


def chost(t):
    a = t.cpu().numpy()
    return a[..., 0] + 1j * a[..., 1]


F = np.float32


def closed_form(op, a):
    h, w = a.shape
    d = np.zeros_like(a)
    if op == "Dx":
        d[:, :-1] = a[:, 1:] - a[:, :-1]
    elif op == "Dy":
        d[:-1] = a[1:] - a[:-1]
    elif op == "DxT":
        if w >= 2:
            d[:, 0] = -a[:, 0]
            d[:, 1:-1] = a[:, :-2] - a[:, 1:-1]
            d[:, -1] = a[:, -2]
    elif op == "DyT":
        if h >= 2:
            d[0] = -a[0]
            d[1:-1] = a[:-2] - a[1:-1]
            d[-1] = a[-2]
    elif op == "D1":
        d[:-1, :-1] = a[1:, 1:] - a[:-1, :-1]
    elif op == "D2":
        d[:-1, 1:] = a[1:, :-1] - a[:-1, 1:]
    elif op == "D1T":
        for y in range(h):
            for x in range(w):
                v = a.dtype.type(0)
                if y + 1 < h and x + 1 < w:
                    v = v - a[y, x]
                if y >= 1 and x >= 1:
                    v = v + a[y - 1, x - 1]
                d[y, x] = v
    elif op == "D2T":
        for y in range(h):
            for x in range(w):
                v = a.dtype.type(0)
                if y + 1 < h and x >= 1:
                    v = v - a[y, x]
                if y >= 1 and x + 1 < w:
                    v = v + a[y - 1, x + 1]
                d[y, x] = v
    return d


PAIRS = [("Dx", "DxT"), ("Dy", "DyT"), ("D1", "D1T"), ("D2", "D2T")]


@pytest.mark.parametrize("hw", [(9, 13), (1, 7), (6, 1), (2, 2), (1, 1)])
@pytest.mark.parametrize("pair", PAIRS)
def test_differences_closed_form_and_adjoint(engine, hw, pair):
    rng = np.random.default_rng(3)
    a = rng.normal(0, 10, hw).astype(F)
    b = rng.normal(0, 10, hw).astype(F)
    op, adj = pair
    da = engine.difference(op, dev(a)).cpu().numpy()
    tb = engine.difference(adj, dev(b)).cpu().numpy()
    assert np.array_equal(da, closed_form(op, a))
    assert np.array_equal(tb, closed_form(adj, b))
    lhs = float(np.dot(da.astype(np.float64).ravel(), b.astype(np.float64).ravel()))
    rhs = float(np.dot(a.astype(np.float64).ravel(), tb.astype(np.float64).ravel()))
    assert abs(lhs - rhs) <= 1e-4 * (1.0 + abs(lhs))
    # complex planes: each component on its own
    z = (rng.normal(0, 5, hw) + 1j * rng.normal(0, 5, hw)).astype(np.complex64)
    dz = chost(engine.difference(op, cdev(z)))
    assert np.array_equal(dz.real.astype(F), closed_form(op, z.real.astype(F)))
    assert np.array_equal(dz.imag.astype(F), closed_form(op, z.imag.astype(F)))


@pytest.mark.parametrize("op", ["Dx", "Dy", "D1", "D2"])
def test_weighted_differences(engine, op):
    rng = np.random.default_rng(4)
    a = rng.normal(0, 10, (11, 17)).astype(F)
    wt = rng.uniform(0, 2, (11, 17)).astype(F)
    got = engine.difference(op, dev(a), dev(wt)).cpu().numpy()
    assert np.array_equal(got, (wt * closed_form(op, a)).astype(F))
    with pytest.raises(gpu.HvdGpuError, match="weight"):
        engine.difference("DxT", dev(a), dev(wt))


@pytest.mark.parametrize("hw", [(242, 760), (30, 9), (5, 8)])
def test_estimate_noise_ire(engine, orc32, hw):
    rng = np.random.default_rng(5)
    h, w = hw
    x = np.arange(w)
    s = (40 + 20 * np.cos(0.5 * np.pi * x)[None, :] + rng.normal(0, 0.7, (h, w))).astype(F)
    got = engine.estimate_noise_ire(dev(s))
    assert got == F(orc32.estimate_noise_ire(s))
    if w > 8:
        assert 0.5 < got < 0.9


@pytest.mark.parametrize("hw", [(274, 760), (320, 922), (17, 31), (1, 5), (64, 64),
                                # Rader stages: 137 and 461 (one per transform), 37 * 41 and
                                # 37 * 37 (two), 509 in columns; 4093's tables exceed a block's
                                # LDS, so it runs as the direct pair form
                                (922, 74), (1018, 137), (2, 1517), (3, 1369), (74, 111), (3, 4093)])
def test_fft2d(engine, hw):
    rng = np.random.default_rng(6)
    z = (rng.normal(0, 1, hw) + 1j * rng.normal(0, 1, hw)).astype(np.complex64)
    f = chost(engine.fft2d(cdev(z)))
    want = np.fft.fft2(z.astype(np.complex128))
    scale = np.abs(want).max()
    assert np.abs(f - want).max() <= 2e-6 * scale * np.log2(max(hw[0] * hw[1], 2))
    back = chost(engine.fft2d(cdev(f.astype(np.complex64)), inverse=True))
    assert np.abs(back - z).max() <= 1e-5 * np.log2(max(hw[0] * hw[1], 2))


def _measurements(rng, m, th=5, tw=7):
    offs = [o for o in (-3, -2, -1, 1, 2, 3)][:m]
    shifts = [0.5 * ((o & 1) * (1 if o > 0 else -1)) for o in offs]
    v = rng.uniform(-2, 2, (th, tw)).astype(F)
    u = rng.uniform(-3, 3, (th, tw)).astype(F)
    dy = np.stack([o * v + s + rng.normal(0, 0.3, v.shape) for o, s in zip(offs, shifts)]).astype(F)
    dx = np.stack([o * u + rng.normal(0, 0.3, u.shape) for o in offs]).astype(F)
    dy[:, 0, 0] += 9.0  # an outlier tile
    conf = rng.uniform(0.0, 1.0, dy.shape).astype(F)
    return offs, shifts, dy, dx, conf


@pytest.mark.parametrize("m", [1, 2, 4, 6])
def test_trajectory_fit_and_snap(engine, orc32, m):
    rng = np.random.default_rng(7 + m)
    offs, shifts, dy, dx, conf = _measurements(rng, m)
    vy, vx = engine.trajectory_fit(offs, shifts, [dev(a) for a in dy], [dev(a) for a in dx], [dev(a) for a in conf])
    wy, wx = orc32.fit_trajectory(offs, np.asarray(shifts, F), dy, dx, conf)
    assert np.array_equal(vy.cpu().numpy(), wy) and np.array_equal(vx.cpu().numpy(), wx)
    ddy, ddx = [dev(a) for a in dy], [dev(a) for a in dx]
    engine.trajectory_snap(offs, shifts, ddy, ddx, [dev(a) for a in conf])
    sy, sx = orc32.trajectory_snap(offs, np.asarray(shifts, F), dy, dx, conf)
    for k in range(m):
        assert np.array_equal(ddy[k].cpu().numpy(), sy[k]) and np.array_equal(ddx[k].cpu().numpy(), sx[k])


@pytest.mark.parametrize("scale,nr_radius,moving", [(2, 2, False), (2, 1, True), (3, 2, True)])
def test_drizzle_frame_matches_hvdcore(engine, ref, scale, nr_radius, moving):
    from sequence_cases import window
    fields, parities, inits, _ = window(lines=40, width=96, nf=8, motion=(0, 2) if moving else (0, 0), seed=21)
    ys = [(s - (c * car).real).astype(F) for (s, car), c in zip(fields, inits)]
    chis = [c.astype(np.complex64) for c in inits]
    wy, wc = ref.drizzle_frame(2, ys, chis, parities, nr_radius=nr_radius, scale=scale)
    gy, gc = engine.drizzle_frame(2, [dev(a) for a in ys], [cdev(c) for c in chis], parities, nr_radius=nr_radius,
                                  scale=scale)
    assert np.array_equal(gy.cpu().numpy(), wy)
    gca = gc.cpu().numpy()
    assert np.array_equal(gca[..., 0], wc.real.astype(F)) and np.array_equal(gca[..., 1], wc.imag.astype(F))
