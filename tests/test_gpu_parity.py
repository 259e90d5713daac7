"""GPU parity: every hot-path entry point of the C ABI against the FP64 oracle.

Float results are held to TOL_IRE = 1e-4 of the 140 IRE code range (the
precision bar for Y, C = S - Y and chi = V - iU); integer results (V-switch
parity, SAD vectors) must be bit-exact.
"""
import math

import numpy as np
import pytest

from conftest import TOL_IRE

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import gpu, synth  # noqa: E402
import oracle as O  # noqa: E402

DEV = "cuda:0"


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(DEV)


def cplx_dev(z):
    return dev(np.stack([z.real, z.imag], -1).astype(np.float32))


def cplx_host(t):
    a = t.cpu().numpy()
    return a[..., 0].astype(np.float64) + 1j * a[..., 1]


def field_ire(raw):
    return synth.ire_from_raw(raw)


def theta_err_ok(got, want):
    """theta is stored as float32 (as hvd-core returns it); allow one float32
    ulp of the unwrapped phase plus 2e-6 rad of arithmetic difference."""
    err = np.abs(np.angle(np.exp(1j * (np.asarray(got, np.float64) - want))))
    bound = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64) + 2e-6
    return bool(np.all(err <= bound)), float(err.max())


# ---------------------------------------------------------------- lock-in
@pytest.mark.parametrize("seed", [1, 2])
def test_lockin_ntsc(engine, orc64, seed):
    raw, _, _ = synth.ntsc_field(seed)
    f = field_ire(raw)
    th, vs = engine.burst_lockin_phase([dev(f)], [synth.NTSC])
    got = th[0].cpu().numpy().astype(np.float64)
    want = orc64.burst_lockin_phase(f, synth.NTSC)
    assert vs[0] is None
    ok, worst = theta_err_ok(got, want)
    assert ok, worst


@pytest.mark.parametrize("seed", [3, 4])
def test_lockin_pal(engine, orc64, seed):
    raw, _, _ = synth.pal_field(seed)
    f = field_ire(raw)
    th, vs = engine.burst_lockin_phase([dev(f)], [synth.PAL])
    want_t, want_v = orc64.burst_lockin_phase_pal(f, synth.PAL)
    assert np.array_equal(vs[0].cpu().numpy(), want_v)
    got = th[0].cpu().numpy().astype(np.float64)
    ok, worst = theta_err_ok(got, want_t)
    assert ok, worst


def test_lockin_no_burst(engine, orc64):
    # a field with no burst anywhere keeps raw angles (NTSC) / model + alternating parity (PAL)
    for g in (synth.NTSC, synth.PAL):
        f = np.zeros((g.field_height, g.field_width), np.float32)
        th, vs = engine.burst_lockin_phase([dev(f)], [g])
        if g.uses_vswitch:
            wt, wv = orc64.burst_lockin_phase_pal(f, g)
            assert np.array_equal(vs[0].cpu().numpy(), wv)
        else:
            wt = orc64.burst_lockin_phase(f, g)
        ok, worst = theta_err_ok(th[0].cpu().numpy(), wt)
        assert ok, worst


def test_burst_amplitude(engine, orc64):
    raw_n, _, _ = synth.ntsc_field(5)
    raw_p, _, _ = synth.pal_field(6)
    fs = [field_ire(raw_n), field_ire(raw_p)]
    amp = engine.burst_amplitude_ire([dev(f) for f in fs], [synth.NTSC, synth.PAL]).cpu().numpy()
    want = [orc64.burst_amplitude_ire(fs[0], synth.NTSC), orc64.burst_amplitude_ire(fs[1], synth.PAL)]
    assert np.allclose(amp, want, atol=1e-4)
    assert abs(amp[0] - 20.0) < 0.3 and abs(amp[1] - 21.43) < 0.3


# ---------------------------------------------------------------- init
def _prepared(orc64, g, raw):
    f = field_ire(raw)
    s, car, th, vs = orc64.prepare_field(f, g)
    fal = g.first_active_field_line
    th_act = th[fal:fal + g.lines].astype(np.float32)
    vs_act = vs[fal:fal + g.lines] if g.uses_vswitch else None
    # the oracle's carrier from the same float theta the GPU receives
    car = orc64.make_carrier(th_act, vs_act, g)
    return s.astype(np.float32), car, th_act, vs_act


def _plane(g, s, chi, y, th, vs):
    return gpu.Engine.plane(s, chi, y, theta=th, vswitch=vs, a0=g.active_video_start,
                            w0=gpu.geometry_phase_per_sample(g), sample_rate=g.sample_rate)


@pytest.mark.parametrize("std", ["ntsc", "pal"])
def test_holographic_init(engine, orc64, std):
    g = synth.NTSC if std == "ntsc" else synth.PAL
    raw = (synth.ntsc_field if std == "ntsc" else synth.pal_field)(11)[0]
    s, car, th, vs = _prepared(orc64, g, raw)
    y0_ref, chi0_ref = orc64.holographic_init(s, car, g.sample_rate)
    ds, dchi, dy = dev(s), torch.zeros((g.lines, g.width, 2), device=DEV), torch.zeros((g.lines, g.width), device=DEV)
    dth = dev(th)
    dvs = None if vs is None else dev(vs)
    engine.holographic_init([_plane(g, ds, dchi, dy, dth, dvs)])
    chi0 = cplx_host(dchi)
    y0 = dy.cpu().numpy()
    assert np.abs(chi0 - chi0_ref).max() <= TOL_IRE
    assert np.abs(y0 - y0_ref).max() <= TOL_IRE
    # lossless split Y0 + Re[chi0 c] = S (lossless_identity_test.cpp:73-74)
    assert np.abs(y0 + (chi0 * car).real - s).max() < 1e-3


@pytest.mark.parametrize("hw", [(53, 41), (21, 137), (242, 922), (7, 74), (288, 760)])
def test_holographic_init_any_size(engine, orc64, hw):
    """Planes of other sizes than the two standards take the run-time plans:
    Rader stages (padded heights 85 = 5 * 17, 53 = 53; widths 41, 137, 922),
    pair-form primes and mixed radices, in one batch with a standard plane."""
    rng = np.random.default_rng(17)
    g = synth.NTSC
    planes, refs, outs = [], [], []
    for h, w in [hw, (g.lines, g.width)]:
        x = np.arange(w)
        ph = 0.5 * np.pi * x[None, :] + np.pi * np.arange(h)[:, None]
        s = (40 + 10 * rng.standard_normal((h, w)) + 20 * np.cos(ph + 0.3)).astype(np.float32)
        car = np.exp(1j * ph).astype(np.complex64).astype(np.complex128)
        refs.append((s, car) + orc64.holographic_init(s, car, g.sample_rate))
        dchi = torch.zeros((h, w, 2), device=DEV)
        dy = torch.zeros((h, w), device=DEV)
        outs.append((dchi, dy))
        planes.append(gpu.Engine.plane(dev(s), dchi, dy, carrier=cplx_dev(car.astype(np.complex64)),
                                       sample_rate=g.sample_rate))
    engine.holographic_init(planes)
    for (s, car, y0_ref, chi0_ref), (dchi, dy) in zip(refs, outs):
        assert np.abs(cplx_host(dchi) - chi0_ref).max() <= TOL_IRE
        assert np.abs(dy.cpu().numpy() - y0_ref).max() <= TOL_IRE


def test_holographic_init_batch_invariant(engine, orc64):
    """A field's init has the same bits alone and in a batch with fields of
    other sizes and transform classes."""
    rng = np.random.default_rng(23)
    fields = []
    for h, w in [(288, 922), (242, 760), (53, 41), (288, 922), (21, 137)]:
        x = np.arange(w)
        ph = 0.5 * np.pi * x[None, :] + np.pi * np.arange(h)[:, None]
        s = (40 + 10 * rng.standard_normal((h, w)) + 20 * np.cos(ph + 0.3)).astype(np.float32)
        fields.append((s, np.exp(1j * ph).astype(np.complex64)))

    def run(idx):
        outs, planes = [], []
        for i in idx:
            s, car = fields[i]
            dchi = torch.zeros(s.shape + (2,), device=DEV)
            dy = torch.zeros(s.shape, device=DEV)
            outs.append((dchi, dy))
            planes.append(gpu.Engine.plane(dev(s), dchi, dy, carrier=cplx_dev(car), sample_rate=synth.PAL.sample_rate))
        engine.holographic_init(planes)
        return [(c.cpu().numpy(), y.cpu().numpy()) for c, y in outs]

    batch = run(range(len(fields)))
    for i in range(len(fields)):
        (c1, y1), = run([i])
        assert np.array_equal(c1, batch[i][0]) and np.array_equal(y1, batch[i][1]), i


def test_holographic_init_symmetry_variant(engine, orc64):
    g = synth.NTSC
    s, car, th, vs = _prepared(orc64, g, synth.ntsc_field(12)[0])
    y0_ref, chi0_ref = orc64.holographic_init(s, car, g.sample_rate, symmetry_variant=True)
    dchi = torch.zeros((g.lines, g.width, 2), device=DEV)
    dy = torch.zeros((g.lines, g.width), device=DEV)
    engine.holographic_init([_plane(g, dev(s), dchi, dy, dev(th), None)], symmetry_variant=True)
    assert np.abs(cplx_host(dchi) - chi0_ref).max() <= TOL_IRE
    assert np.abs(dy.cpu().numpy() - y0_ref).max() <= TOL_IRE


# ---------------------------------------------------------------- solve
CFG0 = dict(cg_iterations=100, irls_outer=4, fast=0, cg_tol=1e-12)


def _refine_case(engine, orc64, g, raw, **kw):
    s, car, th, vs = _prepared(orc64, g, raw)
    _, chi0 = orc64.holographic_init(s, car, g.sample_rate)
    chi0 = chi0.astype(np.complex64).astype(np.complex128)
    cfg = O.Cfg.make(is_pal=int(g.uses_vswitch), **kw)
    y_ref, chi_ref, an_ref, it_ref = orc64.variational_refine(s, car, chi0, cfg)
    dchi = cplx_dev(chi0)
    dy = torch.zeros((g.lines, g.width), device=DEV)
    p = gpu.SolverParams.make(is_pal=int(g.uses_vswitch), **kw)
    iters, aniso = engine.variational_refine([_plane(g, dev(s), dchi, dy, dev(th), None if vs is None else dev(vs))],
                                             p, want_stats=True)
    chi = cplx_host(dchi)
    y = dy.cpu().numpy()
    return dict(s=s, car=car, y=y, chi=chi, y_ref=y_ref, chi_ref=chi_ref, aniso=aniso[0], aniso_ref=an_ref,
                iters=iters[0], iters_ref=it_ref)


@pytest.mark.parametrize("std", ["ntsc", "pal"])
def test_refine_full_field(engine, orc64, std):
    g = synth.NTSC if std == "ntsc" else synth.PAL
    raw = (synth.ntsc_field if std == "ntsc" else synth.pal_field)(21)[0]
    r = _refine_case(engine, orc64, g, raw, **CFG0)
    assert abs(r["aniso"] - r["aniso_ref"]) < 1e-4
    assert np.abs(r["y"] - r["y_ref"]).max() <= TOL_IRE
    assert np.abs(r["chi"] - r["chi_ref"]).max() <= TOL_IRE
    c, c_ref = r["s"] - r["y"], r["s"] - r["y_ref"]
    assert np.abs(c - c_ref).max() <= TOL_IRE
    # lossless split (variational_v2_test.cpp:112)
    assert np.abs(r["y"] + (r["chi"] * r["car"]).real - r["s"]).max() < 1e-3


@pytest.mark.parametrize("kw", [dict(cg_iterations=2, irls_outer=4, fast=1),          # hvd-core defaults
                                dict(cg_iterations=60, irls_outer=4, fast=1),
                                dict(cg_iterations=30, irls_outer=3, fast=0, chroma_aniso=0.5),
                                dict(cg_iterations=40, irls_outer=4, fast=0, cg_tol=0.2, lambda_c=2.0)])
def test_refine_configs(engine, orc64, kw):
    r = _refine_case(engine, orc64, synth.NTSC, synth.ntsc_field(31)[0], **kw)
    assert np.array_equal(r["iters"], r["iters_ref"])
    assert np.abs(r["y"] - r["y_ref"]).max() <= TOL_IRE
    assert np.abs(r["chi"] - r["chi_ref"]).max() <= TOL_IRE


def _scene(h, w):
    # variational_v2_test.cpp:64-95 style solver-only scene with a corrupted chi0
    r, c = np.mgrid[0:h, 0:w].astype(np.float64)
    car = np.exp(1j * (0.5 * math.pi * c))
    y_true = 30 + 20 * np.sin(0.35 * (r + c))
    diag = np.sin(0.4 * (c - r))
    chi_true = 10 * diag - 6j * diag
    s = (y_true + (chi_true * car).real).astype(np.float32)
    e = 3 * np.sin(0.9 * c) * np.cos(0.7 * r)
    chi0 = (chi_true + e - 1j * e).astype(np.complex64).astype(np.complex128)
    return s, car.astype(np.complex64).astype(np.complex128), chi0


@pytest.mark.parametrize("hw", [(48, 56), (1, 1), (1, 9), (9, 1), (2, 2), (5, 130), (67, 3)])
def test_refine_explicit_carrier_shapes(engine, orc64, hw):
    h, w = hw
    s, car, chi0 = _scene(h, w)
    cfg = dict(cg_iterations=60, irls_outer=4, fast=0, chroma_aniso=0.5)
    y_ref, chi_ref, _, it_ref = orc64.variational_refine(s, car, chi0, O.Cfg.make(**cfg))
    dchi = cplx_dev(chi0)
    dy = torch.zeros((h, w), device=DEV)
    pl = gpu.Engine.plane(dev(s), dchi, dy, carrier=cplx_dev(car))
    iters, _ = engine.variational_refine([pl], gpu.SolverParams.make(**cfg), want_stats=True)
    assert np.abs(dy.cpu().numpy() - y_ref).max() <= TOL_IRE
    assert np.abs(cplx_host(dchi) - chi_ref).max() <= TOL_IRE


def test_refine_batch_matches_single(engine, orc64):
    # a batch of uneven planes gives the same result as one launch per plane
    cases = [_scene(48, 56), _scene(30, 200), _scene(7, 7)]
    cfg = gpu.SolverParams.make(cg_iterations=40, irls_outer=4, fast=0, chroma_aniso=0.7)
    single = []
    for s, car, chi0 in cases:
        dchi, dy = cplx_dev(chi0), torch.zeros(s.shape, device=DEV)
        engine.variational_refine([gpu.Engine.plane(dev(s), dchi, dy, carrier=cplx_dev(car))], cfg)
        single.append((dy.cpu().numpy(), dchi.cpu().numpy()))
    planes, outs = [], []
    for s, car, chi0 in cases:
        dchi, dy = cplx_dev(chi0), torch.zeros(s.shape, device=DEV)
        planes.append(gpu.Engine.plane(dev(s), dchi, dy, carrier=cplx_dev(car)))
        outs.append((dy, dchi))
    engine.variational_refine(planes, cfg)
    for (y1, c1), (dy, dchi) in zip(single, outs):
        assert np.array_equal(y1, dy.cpu().numpy()) and np.array_equal(c1, dchi.cpu().numpy())


# ---------------------------------------------------------------- field driver
@pytest.mark.parametrize("std", ["ntsc", "pal"])
def test_decode_field_u16(engine, orc64, std):
    g = synth.NTSC if std == "ntsc" else synth.PAL
    raw = (synth.ntsc_field if std == "ntsc" else synth.pal_field)(41)[0]
    cfg = dict(cg_iterations=40, irls_outer=4, fast=0, cg_tol=1e-6)
    want = orc64.decode_field(field_ire(raw), g, O.Cfg.make(**cfg))
    ys, chis, ths = engine.decode_fields_u16([dev(raw.view(np.int16))], [g], gpu.SolverParams.make(**cfg),
                                             want_theta=True)
    ok, worst = theta_err_ok(ths[0].cpu().numpy(), want["theta"])
    assert ok, worst
    assert np.abs(ys[0].cpu().numpy() - want["y"]).max() <= TOL_IRE
    assert np.abs(cplx_host(chis[0]) - want["chi"]).max() <= TOL_IRE


def test_decode_mixed_batch(engine, orc64):
    # 2 NTSC + 2 PAL fields of uneven size in one call equal per-field calls
    raws = [synth.ntsc_field(51)[0], synth.pal_field(52)[0], synth.ntsc_field(53)[0], synth.pal_field(54)[0]]
    gs = [synth.NTSC, synth.PAL, synth.NTSC, synth.PAL]
    p = gpu.SolverParams.make(cg_iterations=40, irls_outer=4, fast=0, cg_tol=1e-6)
    ys, chis, _ = engine.decode_fields_u16([dev(r.view(np.int16)) for r in raws], gs, p)
    for r, g, y, c in zip(raws, gs, ys, chis):
        y1, c1, _ = engine.decode_fields_u16([dev(r.view(np.int16))], [g], p)
        assert np.array_equal(y.cpu().numpy(), y1[0].cpu().numpy())
        assert np.array_equal(c.cpu().numpy(), c1[0].cpu().numpy())


def test_decode_chroma_phase(engine, orc64):
    g = synth.PAL
    raw = synth.pal_field(61)[0]
    cfg = dict(cg_iterations=20, irls_outer=2, fast=0)
    want = orc64.decode_field(field_ire(raw), g, O.Cfg.make(**cfg), chroma_phase_deg=12.5)
    ys, chis, _ = engine.decode_fields_u16([dev(raw.view(np.int16))], [g], gpu.SolverParams.make(**cfg),
                                           chroma_phase_deg=12.5)
    assert np.abs(ys[0].cpu().numpy() - want["y"]).max() <= TOL_IRE
    assert np.abs(cplx_host(chis[0]) - want["chi"]).max() <= TOL_IRE


# ---------------------------------------------------------------- motion
def test_sad_block_match_sequence(engine, orc64):
    raws, _ = synth.motion_sequence(71, n_fields=4)
    refs, curs = raws[:-1], raws[1:]
    vy, vx, sad = engine.block_match_sad_u16([dev(r.view(np.int16)) for r in refs],
                                             [dev(c.view(np.int16)) for c in curs])
    for k in range(len(refs)):
        wy, wx, ws = orc64.block_match_sad_u16(refs[k], curs[k])
        assert np.array_equal(vy[k].cpu().numpy(), wy)
        assert np.array_equal(vx[k].cpu().numpy(), wx)
        assert np.array_equal(sad[k].cpu().numpy(), ws)


@pytest.mark.parametrize("shape,blk,srch", [((17, 23), (8, 16), (4, 16)), ((8, 16), (8, 16), (0, 0)),
                                             ((40, 70), (16, 32), (16, 32)), ((33, 5), (4, 4), (2, 3)),
                                             ((29, 45), (8, 5), (3, 6)), ((24, 96), (8, 2), (1, 1))])
def test_sad_block_match_edges(engine, orc64, shape, blk, srch):
    rng = np.random.default_rng(5)
    ref_p = rng.integers(0, 65536, size=shape, dtype=np.uint16)
    cur_p = np.roll(ref_p, (1, -2), axis=(0, 1))
    flat = np.full(shape, 1234, np.uint16)  # all-tie field: zero vector must win
    for a, b in ((ref_p, cur_p), (flat, flat)):
        vy, vx, sad = engine.block_match_sad_u16([dev(a.view(np.int16))], [dev(b.view(np.int16))],
                                                 bh=blk[0], bw=blk[1], sy=srch[0], sx=srch[1])
        wy, wx, ws = orc64.block_match_sad_u16(a, b, bh=blk[0], bw=blk[1], sy=srch[0], sx=srch[1])
        assert np.array_equal(vy[0].cpu().numpy(), wy) and np.array_equal(vx[0].cpu().numpy(), wx)
        assert np.array_equal(sad[0].cpu().numpy(), ws)


def test_sample_to_ire(engine):
    raw = np.arange(0, 65536, 7, dtype=np.uint16)
    got = engine.sample_to_ire_u16(dev(raw.view(np.int16)), 20480.0, 61440.0).cpu().numpy()
    assert np.array_equal(got, synth.ire_from_raw(raw))


# ---------------------------------------------------------------- EstimateMotion (float SSD pyramid)
def _motion_pairs():
    from test_oracle_math import _motion_texture
    ref = _motion_texture(128, 128).astype(np.float32)
    yy, xx = np.mgrid[0:128, 0:128]
    cur = ref[np.clip(yy - 3, 0, 127), np.clip(xx + 2, 0, 127)]
    pairs = [(ref, cur, 32, 8)]
    # consecutive luma planes of a moving NTSC scene at active size
    sc = synth.make_scene(synth.NTSC, np.random.default_rng(81), canvas_pad=24)
    Y = sc.y.astype(np.float32)
    a = Y[24:24 + 242, 24:24 + 760]
    b = Y[24 - 1:24 - 1 + 242, 24 + 5:24 + 5 + 760]
    pairs.append((a, b, 32, 16))
    r = np.random.default_rng(82)
    pairs.append((r.normal(size=(37, 53)).astype(np.float32) * 20, r.normal(size=(37, 53)).astype(np.float32) * 20,
                  16, 8))
    return pairs


@pytest.mark.parametrize("k", [0, 1, 2])
def test_estimate_motion_bit_exact(engine, orc32, k):
    ref_p, cur_p, tile, search = _motion_pairs()[k]
    dy, dx, cf = engine.estimate_motion([dev(ref_p)], [dev(cur_p)], tile=tile, search=search)
    wdy, wdx, wcf = orc32.estimate_motion(ref_p, cur_p, tile, search)
    assert np.array_equal(dy[0].cpu().numpy(), wdy)
    assert np.array_equal(dx[0].cpu().numpy(), wdx)
    assert np.array_equal(cf[0].cpu().numpy(), wcf)


def test_estimate_motion_batch(engine, orc32):
    seq, _ = synth.motion_sequence(83, n_fields=4)
    g = synth.NTSC
    ys = [synth.ire_from_raw(r)[g.first_active_field_line:, g.active_video_start:g.active_video_end] for r in seq]
    ys = [np.ascontiguousarray(y) for y in ys]
    dy, dx, cf = engine.estimate_motion([dev(y) for y in ys[:-1]], [dev(y) for y in ys[1:]], tile=32, search=16)
    for k in range(3):
        wdy, wdx, wcf = orc32.estimate_motion(ys[k], ys[k + 1], 32, 16)
        assert np.array_equal(dy[k].cpu().numpy(), wdy) and np.array_equal(dx[k].cpu().numpy(), wdx)
        assert np.array_equal(cf[k].cpu().numpy(), wcf)


# ---------------------------------------------------------------- error paths
def test_abi_rejects_bad_arguments(engine):
    g = synth.NTSC
    raw = dev(synth.ntsc_field(1)[0].view(np.int16))
    with pytest.raises(gpu.HvdGpuError, match="diag_prior"):
        engine.decode_fields_u16([raw], [g], gpu.SolverParams.make(diag_prior=-0.5))
    bad = synth.Geometry(910, 700, 134, 894, 76, 112, 21, 0, g.sample_rate, 0.0, 0)
    with pytest.raises(gpu.HvdGpuError, match="lines"):
        engine.burst_lockin_phase([dev(np.zeros((700, 910), np.float32))], [bad])
    bad = synth.Geometry(910, 263, 134, 2000, 76, 112, 21, 0, g.sample_rate, 0.0, 0)
    with pytest.raises(gpu.HvdGpuError, match="active"):
        engine.decode_fields_u16([raw], [bad], gpu.SolverParams.make())
    with pytest.raises(gpu.HvdGpuError, match="search"):
        engine.block_match_sad_u16([raw], [raw], sy=40)
    # the context stays usable after a rejected call
    ys, _, _ = engine.decode_fields_u16([raw], [g], gpu.SolverParams.make())
    assert torch.isfinite(ys[0]).all()


# ---------------------------------------------------------------- frame bridge
def _bridge_case(engine, orc64, fpd, nframes, core, **bkw):
    import bridge as OB
    frames = synth.frames_s16(fpd, nframes, seed=91)
    fp_o = OB.FrameParams(**fpd)
    fp_g = gpu.FrameParams(**{k: v for k, v in fpd.items() if k != "extra_sample_lines"})
    ex = fpd["extra_sample_lines"]
    for i, e in enumerate(ex):
        fp_g.extra_sample_lines[i] = e
    fp_g.extra_sample_line_count = len(ex)
    cfg = dict(cg_iterations=40, irls_outer=4, fast=0, cg_tol=1e-6)
    want, wgain, f1b = OB.decode_frame_sequence_window(orc64, frames, core[0], core[1], fp_o,
                                                       O.Cfg.make(**cfg), synth.Geometry, **bkw)
    got, gain = engine.decode_frame_window_s16([dev(f) for f in frames], core[0], core[1], fp_g,
                                               gpu.SolverParams.make(**cfg), gpu.BridgeParams.make(**bkw),
                                               want_gain=True)
    return frames, want, wgain, got, gain, f1b


def _woven_composite(frame, fpd, f1_bottom):
    """The raw composite codes rearranged the way the parity weave places the
    decoded fields (frame_bridge.cpp:899-918), flat frame_width x frame_height."""
    fw, fh, f1 = fpd["frame_width"], fpd["frame_height"], fpd["field1_lines"]
    flat = np.concatenate([frame[synth.frame_line_offset(fpd, L):synth.frame_line_offset(fpd, L) + fw]
                           for L in range(fh)]).astype(np.int32).reshape(fh, fw)
    if not f1_bottom:
        return flat.reshape(-1)
    g = synth.PAL if fpd["standard"] == 1 else synth.NTSC
    fal, lines = fpd["first_active_frame_line"] // 2, g.lines
    out = flat.copy()
    for fl in range(lines):
        top, bot = fal + fl, f1 + fal + fl
        if bot < fh:
            out[top], out[bot] = flat[bot], flat[top]
    return out.reshape(-1)


@pytest.mark.parametrize("std,nframes,core", [("ntsc", 2, (0, 2)), ("ntsc", 3, (1, 2)), ("pal", 1, (0, 1))])
def test_frame_window_bridge(engine, orc64, std, nframes, core):
    fpd = synth.FRAME_PAL if std == "pal" else synth.FRAME_NTSC
    frames, want, wgain, got, gain, f1b = _bridge_case(engine, orc64, fpd, nframes, core)
    assert abs(gain - wgain) < 1e-6
    fw, fh = fpd["frame_width"], fpd["frame_height"]
    scale = (fpd["white_level"] - fpd["blanking_level"]) / 100.0
    for (wl, wc, wu, wv), (gl, gc, gu, gv), t in zip(want, got, range(core[0], core[1])):
        gl, gc, gu, gv = (x.cpu().numpy() for x in (gl, gc, gu, gv))
        dl = np.abs(gl.astype(np.int32) - wl)
        dc = np.abs(gc.astype(np.int32) - wc)
        assert dl.max() <= 1 and dc.max() <= 1           # requantised codes within one code
        assert (dl == 0).mean() > 0.999 and (dc == 0).mean() > 0.999
        assert np.abs(gu - wu).max() <= TOL_IRE * scale * max(gain, 1.0)
        assert np.abs(gv - wv).max() <= TOL_IRE * scale * max(gain, 1.0)
        # lossless split: luma + chroma reproduces the composite within one code
        # (frame_bridge_test.cpp:90-96), checked in the active picture
        comp = _woven_composite(frames[t], fpd, f1b)
        active = gc != 0
        assert np.abs(gl.astype(np.int32)[active] + gc.astype(np.int32)[active] - comp[active]).max() <= 1
        # outside the active picture: composite passthrough, zero chroma
        assert np.array_equal(gl[~active & (wc == 0)], wl[~active & (wc == 0)])


def test_frame_window_options(engine, orc64):
    # forced field order, monochrome, chroma gain and phase trim, ACC off
    fpd = synth.FRAME_NTSC
    _, want, wgain, got, gain, f1b = _bridge_case(engine, orc64, fpd, 1, (0, 1), field_order=2, monochrome=1,
                                                  chroma_gain=1.3, acc=0, chroma_phase_deg=20.0)
    assert f1b
    assert abs(gain - 1.3) < 1e-6 and abs(wgain - 1.3) < 1e-6
    gl, gc, gu, gv = (x.cpu().numpy() for x in got[0])
    assert not gc.any() and not gu.any() and not gv.any()
    assert np.abs(gl.astype(np.int32) - want[0][0]).max() <= 1


def _fp_gpu(fpd):
    fp_g = gpu.FrameParams(**{k: v for k, v in fpd.items() if k != "extra_sample_lines"})
    for i, e in enumerate(fpd["extra_sample_lines"]):
        fp_g.extra_sample_lines[i] = e
    fp_g.extra_sample_line_count = len(fpd["extra_sample_lines"])
    return fp_g


def gpu_carrier(engine):
    """The library's MakeCarrier (hvd_gpu_make_carrier, pinned against the
    oracle in test_gpu_components) as a carrier function for the oracle."""
    def make(th, vs, g):
        t = torch.from_numpy(np.ascontiguousarray(th, np.float32)).to(DEV)
        v = None if vs is None else torch.from_numpy(np.ascontiguousarray(vs, np.int8)).to(DEV)
        a = engine.make_carrier(t, v, g).cpu().numpy()
        return (a[..., 0] + 1j * a[..., 1]).astype(np.complex64)
    return make


def gpu_init(engine):
    """The library's HolographicInit (pinned against the oracle's in
    test_holographic_init) as an init function for the oracle."""
    def init(s, car, sample_rate):
        dchi = torch.zeros(s.shape + (2,), device=DEV)
        dy = torch.zeros(s.shape, device=DEV)
        engine.holographic_init([gpu.Engine.plane(dev(s), dchi, dy, carrier=dev(np.stack([car.real, car.imag], -1)),
                                                  sample_rate=sample_rate)])
        c = dchi.cpu().numpy()
        return dy.cpu().numpy(), (c[..., 0] + 1j * c[..., 1]).astype(np.complex64)
    return init


@pytest.mark.parametrize("std,nframes,core,strength,field_order,phase", [
    ("ntsc", 3, (1, 2), 1.0, 0, 0.0), ("ntsc", 3, (0, 3), 0.0, 2, 0.0), ("pal", 3, (1, 2), 1.0, 0, 0.0),
    ("pal", 3, (0, 2), 0.8, 0, 15.0)])
def test_frame_window_3d(engine, orc64, orc32, std, nframes, core, strength, field_order, phase):
    """DecodeFrameSequenceWindow with the window driver engaged: the device's
    split, ACC, field order, lock-in, chroma-phase trim, float carriers,
    inits, 3-D passes and weave against the oracle's restatement (FP64
    lock-in and inits, the FP32 window driver bit-exact with hvd-core's).

    The carriers are hvd-core's float MakeCarrier planes on both sides.  The
    3-D passes turn differences far below TOL_IRE into rare decision flips:
    the device's sinf/cosf round differently from the C library's at about a
    quarter of the samples (by under an ulp), the FP32 init differs from the
    FP64 one by ~1e-5 IRE, and either moves hvd-core's own output by up to
    0.06 IRE at a few samples.  So the strict bar holds against the oracle
    fed the library's carriers and inits (each pinned against the oracle on
    its own: test_gpu_components, test_holographic_init), and against the
    oracle's own carriers and FP64 inits all but 2e-4 of the chroma samples
    stay within TOL_IRE and none beyond 0.15 IRE."""
    import bridge as OB
    import sequence3d as S3
    fpd = synth.FRAME_PAL if std == "pal" else synth.FRAME_NTSC
    frames = synth.frames_s16(fpd, nframes, seed=95)
    fp_o = OB.FrameParams(**fpd)
    cfg = dict(cg_iterations=20, irls_outer=4, fast=1, cg_tol=0.0)
    d_ref, d = {}, {}
    kw = dict(field_order=field_order, chroma_phase_deg=phase)
    want, wgain, f1b = OB.decode_frame_sequence_window_3d(
        orc64, orc32, frames, core[0], core[1], fp_o, S3.WindowCfg(temporal_strength=strength, **cfg),
        synth.Geometry, diag=d_ref, make_carrier=gpu_carrier(engine), holographic_init=gpu_init(engine), **kw)
    want_c, _, _ = OB.decode_frame_sequence_window_3d(
        orc64, orc32, frames, core[0], core[1], fp_o, S3.WindowCfg(temporal_strength=strength, **cfg),
        synth.Geometry, **kw)
    print("frame window 3-D", std, nframes, core, strength, flush=True)
    got, gain = engine.decode_frame_window_s16([dev(f) for f in frames], core[0], core[1], _fp_gpu(fpd),
                                               gpu.SolverParams.make(**cfg),
                                               gpu.BridgeParams.make(temporal_strength=strength, **kw),
                                               want_gain=True, wparams=gpu.WindowParams.make(), diag=d)
    if strength > 0:
        assert d_ref["resolved_strength"] > 0
    for k in d_ref:
        assert abs(d[k] - d_ref[k]) <= 1e-4 * max(1.0, abs(d_ref[k])), (k, d[k], d_ref[k])
    assert abs(gain - wgain) < 1e-6
    scale = (fpd["white_level"] - fpd["blanking_level"]) / 100.0
    for w_, wc_, g_ in zip(want, want_c, got):
        close_3d(g_, w_, scale * max(gain, 1.0))
        g_ = [x.cpu().numpy() for x in g_]
        for a, b in ((g_[2], wc_[2]), (g_[3], wc_[3])):
            e = np.abs(a - b) / (scale * max(gain, 1.0))
            assert (e > TOL_IRE).mean() <= 2e-4 and e.max() <= 0.15


def close_3d(got, want, uv_scale):
    """Requantised codes within one code (almost all exact), u and v within
    TOL_IRE."""
    gl, gc, gu, gv = (x.cpu().numpy() if hasattr(x, "cpu") else x for x in got)
    wl, wc, wu, wv = want
    dl = np.abs(gl.astype(np.int32) - wl)
    dc = np.abs(gc.astype(np.int32) - wc)
    assert dl.max() <= 1 and dc.max() <= 1
    assert (dl == 0).mean() > 0.999 and (dc == 0).mean() > 0.999
    for a, b in ((gu, wu), (gv, wv)):
        assert np.abs(a - b).max() <= TOL_IRE * uv_scale


def test_frame_window_3d_default_entry(engine):
    # the plain entry point takes HvdConfig's default window knobs: the same
    # frames as the _ex call with null knobs
    fpd = synth.FRAME_NTSC
    frames = [dev(f) for f in synth.frames_s16(fpd, 3, seed=96)]
    sp, bp = gpu.SolverParams.make(cg_iterations=10), gpu.BridgeParams.make(temporal_strength=0.0)
    a, _ = engine.decode_frame_window_s16(frames, 0, 3, _fp_gpu(fpd), sp, bp)
    d = {}
    b, _ = engine.decode_frame_window_s16(frames, 0, 3, _fp_gpu(fpd), sp, bp, wparams=gpu.WindowParams.make(),
                                          diag=d)
    for x, y in zip(a, b):
        for p, q in zip(x, y):
            assert torch.equal(p, q)
    assert d["sigma_ire"] > 0


def test_frame_window_rejects_bad_window(engine):
    fpd = synth.FRAME_NTSC
    frames = synth.frames_s16(fpd, 2, seed=92)
    fp_g = gpu.FrameParams(**{k: v for k, v in fpd.items() if k != "extra_sample_lines"})
    with pytest.raises(gpu.HvdGpuError, match="window"):
        engine.decode_frame_window_s16([dev(f) for f in frames], 1, 1, fp_g, gpu.SolverParams.make(),
                                       gpu.BridgeParams.make())
