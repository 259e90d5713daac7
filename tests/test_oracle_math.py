"""Pin the FP64 oracle by what the mathematics fixes, independently of the
reference: textbook definitions, library routines, closed forms and brute
force on small inputs -- chosen so a dropped term, a wrong sign or index or a
transposed operand fails one of them.  These run everywhere (no reference
build needed)."""
import math

import numpy as np
import pytest

from hvd_amd import synth
import oracle as O


def rng(seed=0):
    return np.random.default_rng(seed)


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("hw", [(1, 1), (2, 3), (5, 7), (8, 16), (274, 760), (320, 922), (19, 137)])
def test_fft2d_matches_numpy(orc64, hw):
    r = rng(hw[0] * 7 + hw[1])
    a = r.normal(size=hw) + 1j * r.normal(size=hw)
    f = orc64.fft2d(a, -1)
    ref = np.fft.fft2(a)
    assert np.abs(f - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max())
    back = orc64.fft2d(f, +1)  # inverse divides by h*w (fft2d.h:30-33)
    assert np.abs(back - a).max() <= 1e-12 * max(1.0, np.abs(a).max()) * hw[0] * hw[1]


@pytest.mark.parametrize("hw", [(2, 3), (3, 5), (4, 6)])
def test_fft2d_brute_force(orc64, hw):
    h, w = hw
    r = rng(3)
    a = r.normal(size=hw) + 1j * r.normal(size=hw)
    out = np.zeros_like(a)
    for ky in range(h):
        for kx in range(w):
            acc = 0
            for y in range(h):
                for x in range(w):
                    acc += a[y, x] * np.exp(-2j * np.pi * (ky * y / h + kx * x / w))
            out[ky, kx] = acc
    assert np.abs(orc64.fft2d(a, -1) - out).max() < 1e-12


# ---------------------------------------------------------------- gradients
@pytest.mark.parametrize("hw", [(1, 1), (1, 6), (6, 1), (3, 4), (64, 48)])
def test_adjoint_identities(orc64, hw):
    # <D a, b> = <a, D^T b> for every operator pair (gradients_test.cpp:40-56)
    r = rng(11)
    a, b = r.normal(size=hw), r.normal(size=hw)
    for op, opt in ((0, 2), (1, 3)):
        lhs = np.sum(orc64.real_op(op, a) * b)
        rhs = np.sum(a * orc64.real_op(opt, b))
        assert abs(lhs - rhs) < 1e-10 * (1 + abs(lhs))
    ac = a + 1j * r.normal(size=hw)
    bc = b + 1j * r.normal(size=hw)
    for op, opt in ((0, 2), (1, 3), (4, 5), (6, 7)):
        lhs = np.vdot(orc64.cplx_op(op, ac), bc)
        rhs = np.vdot(ac, orc64.cplx_op(opt, bc))
        assert abs(lhs - rhs) < 1e-10 * (1 + abs(lhs))


def test_forward_differences_closed_form(orc64):
    a = rng(2).normal(size=(5, 6))
    dx = np.zeros_like(a)
    dx[:, :-1] = a[:, 1:] - a[:, :-1]
    dy = np.zeros_like(a)
    dy[:-1] = a[1:] - a[:-1]
    assert np.array_equal(orc64.real_op(0, a), dx) and np.array_equal(orc64.real_op(1, a), dy)


# ---------------------------------------------------------------- lock-in
def test_tridiag_dense_solve(orc64):
    for n in (1, 2, 7, 263):
        r = rng(n)
        d, a = r.normal(size=n), r.uniform(0.1, 2.0, size=n)
        lam = 25.0
        L = np.zeros((n, n))
        for i in range(n - 1):
            L[i, i] += 1; L[i + 1, i + 1] += 1; L[i, i + 1] -= 1; L[i + 1, i] -= 1
        x = np.linalg.solve(np.diag(a) + lam * L, a * d) if n > 1 else d.copy()
        assert np.abs(orc64.tridiag_smooth(d, a, lam) - x).max() < 1e-9
    d = np.array([1.0, 2.0, 3.0, 4.0])
    assert np.abs(orc64.tridiag_smooth(d, np.ones(4), 1e-9) - d).max() < 1e-6  # lockin_test.cpp:62-67


def _burst_field(h, w, x0, x1, theta, amp=20.0):
    g = synth.Geometry(w, h, 0, w, x0, x1, 0, 0, synth.FS_NTSC, 0.0, 0)
    f = np.zeros((h, w))
    xs = np.arange(x0, x1)
    for y in range(h):
        f[y, x0:x1] = -amp * np.sin(theta[y] + 0.5 * np.pi * xs)
    return g, f


def test_lockin_recovers_known_phase(orc64):
    # lockin_test.cpp:28-60: 30 x 200, theta = 0.3 + pi y + 0.05 sin(0.2 y)
    h = 30
    theta = 0.3 + np.pi * np.arange(h) + 0.05 * np.sin(0.2 * np.arange(h))
    g, f = _burst_field(h, 200, 20, 60, theta)
    got = orc64.burst_lockin_phase(f, g)
    err = np.abs(np.angle(np.exp(1j * (got - theta))))
    assert err[2:-2].max() < 0.05
    assert abs(orc64.burst_amplitude_ire(f, g) - 20.0) < 1.0
    # the pure pi-per-line model is a fixed point of the smoother: exact
    theta2 = 0.7 + np.pi * np.arange(h)
    g, f = _burst_field(h, 200, 20, 60, theta2)
    got = orc64.burst_lockin_phase(f, g)
    assert np.abs(np.angle(np.exp(1j * (got - theta2)))).max() < 1e-9


def _pal_burst_field(theta0, destroy=()):
    g = synth.PAL
    f = np.zeros((g.field_height, g.field_width))
    s = synth.vswitch_pattern(g, 1)
    lines = np.arange(g.field_height)
    xs = np.arange(g.colour_burst_start, g.colour_burst_end)
    c = synth.carrier(g, theta0, lines, xs, s)
    f[:, g.colour_burst_start:g.colour_burst_end] = (21.43 * np.exp(0.25j * np.pi) * c).real
    for y in destroy:
        f[y, g.colour_burst_start:g.colour_burst_end] = 0.0
    return g, f, s, theta0 + g.line_advance() * lines


@pytest.mark.parametrize("theta0", [0.0, 1.1, -2.5])
def test_pal_lockin_parity_and_phase(orc64, theta0):
    # pal_lockin_test.cpp:75-114: exact parity, clean and with destroyed bursts
    for destroy in ((), (3, 50, 51, 120, 200, 201, 250, 300)):
        g, f, s, theta = _pal_burst_field(theta0, destroy)
        th, vs = orc64.burst_lockin_phase_pal(f, g)
        assert np.array_equal(vs, s)
        err = np.abs(np.angle(np.exp(1j * (th - theta))))
        assert err.max() < 1e-6


def test_carrier_closed_form(orc64):
    g = synth.PAL
    th = rng(5).uniform(-3, 3, size=10)
    vs = np.array([1, -1] * 5, np.int8)
    c = orc64.make_carrier(th, vs, g)
    x = g.active_video_start + np.arange(g.width)
    want = np.exp(1j * (th[:, None] + 0.5 * np.pi * x[None, :]))
    want[vs < 0] = -np.conj(want[vs < 0])  # ntsc_geometry.cpp:88-89
    assert np.abs(c - want).max() < 1e-12


# ---------------------------------------------------------------- holographic init
def _gauss_gain(fx, fy, h, variant):
    """GaussianLpf (holographic_init.cpp:33-56) response at (fx Hz, fy c/line),
    vertical cutoff against the unpadded height h."""
    mhz, cph = ((0.8, 120.0), (1.8, 30.0))[variant]
    return math.exp(-0.5 * (fx / (mhz * 1e6)) ** 2) * math.exp(-0.5 * (fy / (cph / (2.0 * h))) ** 2)


@pytest.mark.parametrize("hw", [(48, 160), (64, 200), (242, 760)])
def test_init_dc_luma_closed_form(orc64, hw):
    """A DC luma plane L demodulates to one 2-D tone at (f_sc, 1/2 c/line); each
    crop passes it with gain G_v, both residual lumas are flat, so the blend
    weights are equal and chi0 = L (G_1 + G_2) conj(c) exactly."""
    g = synth.NTSC
    h, w = hw
    car = synth.carrier(g, 0.4, np.arange(h), g.active_video_start + np.arange(w), None)
    L = 55.0
    s = np.full((h, w), L)
    y0, chi0 = orc64.holographic_init(s, car, g.sample_rate)
    fsc = g.sample_rate / 4.0
    gain = _gauss_gain(fsc, 0.5, h, 0) + _gauss_gain(fsc, 0.5, h, 1)
    assert np.abs(chi0 - L * gain * np.conj(car)).max() < 1e-9 * L
    assert np.abs(y0 + (chi0 * car).real - s).max() < 1e-12  # Y0 + Re[chi0 c] = S exactly


@pytest.mark.parametrize("hw", [(48, 160), (64, 200)])
def test_init_flat_chroma_bound(orc64, hw):
    """Constant chroma chi: the demodulated image term conj(chi) conj(c)^2 sits at
    (fs/2, 0), so each variant gives chi_v = chi + G_v conj(chi) conj(c)^2 and the
    convex blend is within |chi| max_v G_v of chi everywhere."""
    g = synth.NTSC
    h, w = hw
    car = synth.carrier(g, 0.4, np.arange(h), g.active_video_start + np.arange(w), None)
    chi = 12.0 - 7.0j
    s = (chi * car).real
    y0, chi0 = orc64.holographic_init(s, car, g.sample_rate)
    bound = abs(chi) * max(_gauss_gain(g.sample_rate / 2, 0.0, h, v) for v in (0, 1))
    assert np.abs(chi0 - chi).max() <= bound * (1 + 1e-9)
    assert np.abs(chi0 - chi).max() > 0.1 * bound  # the image term is really there


# ---------------------------------------------------------------- aniso
def test_aniso_forced_and_clamps(orc64):
    chi0 = np.zeros((40, 40), complex)
    assert orc64.resolve_chroma_aniso(chi0, False, forced=0.73) == pytest.approx(0.73)
    # all-zero gradients: r = 0 -> NTSC floor 0.5, PAL floor 0.55
    assert orc64.resolve_chroma_aniso(chi0, False) == pytest.approx(0.5)
    assert orc64.resolve_chroma_aniso(chi0, True) == pytest.approx(0.55)


def test_aniso_ratio_map(orc64):
    # vertical ramp in chi: p98 |Dy| = 2 * slope per pair-average, |Dx| = 0 (px -> 1e-6 floor): r huge -> 1.0
    h, w = 40, 60
    yy = np.arange(h)[:, None] * np.ones((1, w))
    assert orc64.resolve_chroma_aniso(0.5 * yy + 0j, False) == pytest.approx(1.0)
    # horizontal ramp only: r = 0 -> 0.5
    xx = np.ones((h, 1)) * np.arange(w)[None, :]
    assert orc64.resolve_chroma_aniso(0.5 * xx + 0j, False) == pytest.approx(0.5)
    # both: |Dx| = a, |Dy| = 2b (group of 2 lines) -> r = 2b / a; pick r = 1.5
    a, b = 1.0, 0.75
    chi = a * xx + b * yy + 0j
    want = np.clip(0.5 + 1.1 * (2 * b / a - 1.3), 0.5, 1.0)
    assert orc64.resolve_chroma_aniso(chi, False) == pytest.approx(want, rel=1e-9)
    # PAL: group of 4 lines -> |Dy| = 4b
    want_pal = np.clip(0.7 + 0.9 * (4 * b / a - 1.3), 0.55, 1.0)
    assert orc64.resolve_chroma_aniso(chi, True) == pytest.approx(want_pal, rel=1e-9)


def test_aniso_cancels_line_alternating_leak(orc64):
    # NTSC: a leak alternating sign line to line cancels in 2-line averages
    # (pal_aniso_group_test.cpp:32-72 for the PAL 4-line case)
    h, w = 40, 60
    leak = np.where(np.arange(h) % 2 == 0, 5.0, -5.0)[:, None] * np.ones((1, w))
    xx = np.ones((h, 1)) * np.arange(w)[None, :]
    assert orc64.resolve_chroma_aniso(xx + leak + 0j, False) == pytest.approx(0.5)


# ---------------------------------------------------------------- solve
def _scene(h, w):
    r, c = np.mgrid[0:h, 0:w].astype(np.float64)
    car = np.exp(1j * (0.5 * math.pi * c))
    y_true = 30 + 20 * np.sin(0.35 * (r + c))
    diag = np.sin(0.4 * (c - r))
    chi_true = 10 * diag - 6j * diag
    s = y_true + (chi_true * car).real
    e = 3 * np.sin(0.9 * c) * np.cos(0.7 * r)
    return s, car, chi_true + e - 1j * e, chi_true


def test_refine_lossless_fast_and_loose(orc64):
    # variational_v2_test.cpp:96-158
    s, car, chi0, _ = _scene(48, 56)
    base = O.Cfg.make(cg_iterations=60, fast=0)
    y, chi, _, _ = orc64.variational_refine(s, car, chi0, base)
    assert np.abs(y + (chi * car).real - s).max() < 1e-9
    _, chi_f, _, _ = orc64.variational_refine(s, car, chi0, O.Cfg.make(cg_iterations=60, fast=1))
    assert np.sqrt(np.mean(np.abs(chi_f - chi) ** 2)) < 1.0
    _, chi_l, _, _ = orc64.variational_refine(s, car, chi0, O.Cfg.make(cg_iterations=60, fast=0, cg_tol=0.2))
    assert np.sqrt(np.mean(np.abs(chi_l - chi) ** 2)) < 2.0


def test_refine_stationary_point(orc64):
    """With the IRLS weights frozen (one outer stage) and CG run to
    convergence, the solve must reach the minimiser of the quadratic: the
    directional derivative of the energy along any direction vanishes.  The
    energy is evaluated here independently with numpy."""
    s, car, chi0, _ = _scene(12, 16)
    eps, eps_c, k, lam_c, aniso = 0.5, 1.0, 0.25, 1.0, 0.5
    cfg = O.Cfg.make(cg_iterations=400, irls_outer=1, fast=0, cg_tol=1e-14, chroma_aniso=aniso)
    _, chi, _, _ = orc64.variational_refine(s, car, chi0, cfg)

    def dx(a):
        d = np.zeros_like(a); d[:, :-1] = a[:, 1:] - a[:, :-1]; return d

    def dy(a):
        d = np.zeros_like(a); d[:-1] = a[1:] - a[:-1]; return d

    yc0 = s - (chi0 * car).real
    gx, gy = dx(yc0), dy(yc0)
    wx, wy = eps / np.sqrt(gx ** 2 + eps ** 2), eps / np.sqrt(gy ** 2 + eps ** 2)
    wcx = eps_c / np.sqrt(np.abs(dx(chi0)) ** 2 + k * gx ** 2 + eps_c ** 2)
    wcy = eps_c / np.sqrt(np.abs(dy(chi0)) ** 2 + k * gy ** 2 + eps_c ** 2)

    def energy(c):
        yc = s - (c * car).real
        return (np.sum(wx * dx(yc) ** 2 + wy * dy(yc) ** 2)
                + lam_c * aniso * np.sum(wcx * np.abs(dx(c)) ** 2) + lam_c * np.sum(wcy * np.abs(dy(c)) ** 2))

    e0 = energy(chi)
    r = rng(9)
    for _ in range(5):
        v = r.normal(size=chi.shape) + 1j * r.normal(size=chi.shape)
        h = 1e-4
        deriv = (energy(chi + h * v) - energy(chi - h * v)) / (2 * h)
        assert abs(deriv) < 1e-6 * max(1.0, e0)
    assert energy(chi) < energy(chi0)


# ---------------------------------------------------------------- motion
def _sad_brute(ref, cur, bh, bw, sy, sx):
    h, w = ref.shape
    nby, nbx = -(-h // bh), -(-w // bw)
    vy = np.zeros((nby, nbx), np.int16)
    vx = np.zeros((nby, nbx), np.int16)
    sad = np.zeros((nby, nbx), np.int32)
    yy, xx = np.mgrid[0:h, 0:w]
    for by in range(nby):
        for bx in range(nbx):
            sl = (slice(by * bh, min(by * bh + bh, h)), slice(bx * bw, min(bx * bw + bw, w)))
            best = None
            for dyv in range(-sy, sy + 1):
                for dxv in range(-sx, sx + 1):
                    ry = np.clip(yy[sl] - dyv, 0, h - 1)
                    rx = np.clip(xx[sl] - dxv, 0, w - 1)
                    cost = int(np.abs(cur[sl].astype(np.int64) - ref[ry, rx].astype(np.int64)).sum())
                    key = (cost, abs(dyv) + abs(dxv))  # raster order: first wins on equal key
                    if best is None or key < best[0]:
                        best = (key, dyv, dxv)
            vy[by, bx], vx[by, bx], sad[by, bx] = best[1], best[2], best[0][0]
    return vy, vx, sad


@pytest.mark.parametrize("shape,blk,srch", [((17, 23), (4, 8), (2, 4)), ((16, 32), (8, 16), (3, 5)),
                                             ((9, 9), (8, 16), (1, 2))])
def test_sad_block_match_brute_force(orc64, shape, blk, srch):
    r = rng(13)
    ref = r.integers(0, 65536, size=shape, dtype=np.uint16)
    cur = np.roll(ref, (1, -2), axis=(0, 1))
    for a, b in ((ref, cur), (np.full(shape, 77, np.uint16), np.full(shape, 77, np.uint16))):
        got = orc64.block_match_sad_u16(a, b, bh=blk[0], bw=blk[1], sy=srch[0], sx=srch[1])
        want = _sad_brute(a, b, blk[0], blk[1], srch[0], srch[1])
        for x, y in zip(got, want):
            assert np.array_equal(x, y)


def test_sad_recovers_global_shift(orc64):
    r = rng(14)
    base = r.integers(0, 65536, size=(80, 160), dtype=np.uint16)
    ref = base[8:72, 16:144]
    cur = base[8 - 3:72 - 3, 16 + 5:144 + 5]  # cur(y, x) = ref(y - 3, x + 5)
    vy, vx, sad = orc64.block_match_sad_u16(ref, cur, bh=8, bw=16, sy=4, sx=16)
    inner = (slice(1, -1), slice(1, -1))
    assert np.all(vy[inner] == 3) and np.all(vx[inner] == -5) and np.all(sad[inner] == 0)


def _motion_texture(h, w):
    """Sinusoids plus a hash over 4x4 blocks (aperiodic at the coarse level),
    the scene motion_test.cpp:31-55 builds."""
    yy, xx = np.mgrid[0:h, 0:w]
    s = ((yy // 4).astype(np.uint32) * np.uint32(73856093)) ^ ((xx // 4).astype(np.uint32) * np.uint32(19349663))
    s ^= s >> np.uint32(13)
    s = (s * np.uint32(0x5BD1E995)).astype(np.uint32)
    s ^= s >> np.uint32(15)
    noise = ((s & np.uint32(0xFFFF)).astype(np.float64) / 65535.0 - 0.5) * 30.0
    fy, fx = yy.astype(np.float64), xx.astype(np.float64)
    return 40 * np.sin(fy * 0.31) * np.cos(fx * 0.27) + 15 * np.sin(fx * 0.23 + fy * 0.17) + noise


def test_estimate_motion_known_shift(orc64):
    # motion_test.cpp:78-128: 128 x 128, tile 32, cur(y, x) = ref(y - 3, x + 2)
    # (clamped); confident interior tiles carry the shift within 0.6 px and
    # most interior tiles are confident.
    ref = _motion_texture(128, 128)
    yy, xx = np.mgrid[0:128, 0:128]
    cur = ref[np.clip(yy - 3, 0, 127), np.clip(xx + 2, 0, 127)]
    dy, dx, conf = orc64.estimate_motion(ref, cur, 32, 8)
    inner = (slice(1, -1), slice(1, -1))
    good = conf[inner] > 0.15
    assert good.sum() * 2 > good.size
    assert np.abs(dy[inner][good] - 3).max() < 0.6 and np.abs(dx[inner][good] + 2).max() < 0.6


def test_box_blur_prefix_equals_naive(orc64):
    # motion_v2_test.cpp:81-91: prefix-sum blur == naive zero-padded convolution
    a = rng(16).normal(size=(23, 31))
    for r in (1, 2, 4):
        k = np.ones(2 * r + 1) / (2 * r + 1)
        naive = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), 0, a)
        naive = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), 1, naive)
        assert np.abs(orc64.box_blur(a, r) - naive).max() < 1e-12
