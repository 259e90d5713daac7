"""GPU parity of the component entry points the hvd-core C++ API is built
on: frame split and field-order vote, explicit carrier planes, TridiagSmooth,
BoxBlur, the FrameYc chroma plane and the woven-frame re-weave."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import gpu, synth  # noqa: E402
import bridge as OB  # noqa: E402

DEV = "cuda:0"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def fparams(fpd):
    fp = gpu.FrameParams(**{k: v for k, v in fpd.items() if k != "extra_sample_lines"})
    for i, e in enumerate(fpd["extra_sample_lines"]):
        fp.extra_sample_lines[i] = e
    fp.extra_sample_line_count = len(fpd["extra_sample_lines"])
    return fp


@pytest.mark.parametrize("std", ["ntsc", "pal"])
def test_split_frame_fields_and_vote(engine, std):
    fpd = synth.FRAME_PAL if std == "pal" else synth.FRAME_NTSC
    frame = synth.frames_s16(fpd, 1, seed=5)[0]
    fp_o = OB.FrameParams(**fpd)
    want1, want2 = OB.split_frame_fields(frame, fp_o)
    b1, b2 = engine.split_frame_fields(dev(frame), fparams(fpd), fpd["blanking_level"])
    np.testing.assert_array_equal(b1.cpu().numpy(), want1)
    np.testing.assert_array_equal(b2.cpu().numpy(), want2)
    # the vote on the frame and on its blocks swapped (field order inverted)
    v = engine.field_parity_vote(b1, b2, fparams(fpd))
    assert v == OB.detect_field_parity(want1, want2, fp_o)
    assert v == 1
    # chroma-channel split: re-centred on chroma_dc, missing rows read 0
    c1, c2 = engine.split_frame_fields(dev(frame), fparams(fpd), fpd["chroma_dc"])
    cpi = np.float32((fpd["white_level"] - fpd["blanking_level"]) / 100.0)
    raw1 = ((want1 * cpi).astype(np.float32) + np.float32(fpd["blanking_level"]))  # codes back
    exp1 = ((np.round(raw1) - np.float32(fpd["chroma_dc"])) / cpi).astype(np.float32)
    np.testing.assert_allclose(c1.cpu().numpy(), exp1, rtol=0, atol=1e-4)
    f2 = fpd["frame_height"] - fpd["field1_lines"]
    if f2 < fpd["field1_lines"]:
        assert float(c2[f2:].abs().max()) == 0.0


def test_field_parity_vote_inverted(engine):
    fpd = synth.FRAME_NTSC
    frame = synth.frames_s16(fpd, 1, seed=6)[0]
    b1, b2 = engine.split_frame_fields(dev(frame), fparams(fpd), fpd["blanking_level"])
    fp_o = OB.FrameParams(**fpd)
    want = OB.detect_field_parity(b2.cpu().numpy(), b1.cpu().numpy(), fp_o)
    assert engine.field_parity_vote(b2, b1, fparams(fpd)) == want


@pytest.mark.parametrize("std", ["ntsc", "pal"])
def test_make_carrier(engine, orc32, std):
    g = synth.PAL if std == "pal" else synth.NTSC
    rng = np.random.default_rng(3)
    lines = 37
    theta = (rng.uniform(-50, 50, lines)).astype(np.float32)
    vs = np.where(np.arange(lines) % 2 == 0, 1, -1).astype(np.int8) if std == "pal" else None
    got = engine.make_carrier(dev(theta), dev(vs) if vs is not None else None, g).cpu().numpy()
    want = orc32.make_carrier(theta, vs, g)
    want = np.stack([want.real, want.imag], -1)
    # same float phase; sin/cos differ by at most a couple of float ulps
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-6)
    assert np.allclose(np.hypot(got[..., 0], got[..., 1]), 1.0, atol=1e-6)


def test_tridiag_smooth(engine, orc64):
    rng = np.random.default_rng(9)
    for n, lam in [(263, 25.0), (4, 1e-6), (1, 25.0), (2, 3.0)]:
        d = rng.normal(size=n).astype(np.float32)
        a = rng.uniform(0, 2, size=n).astype(np.float32)
        got = engine.tridiag_smooth(dev(d), dev(a), lam).cpu().numpy()
        want = orc64.tridiag_smooth(d.astype(np.float64), a.astype(np.float64), lam)
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("hw,r", [((242, 760), 2), ((7, 5), 3), ((1, 9), 1), ((3, 2100), 4), ((5, 1500), 8)])
def test_box_blur(engine, orc32, hw, r):
    rng = np.random.default_rng(11)
    a = rng.normal(50, 20, size=hw).astype(np.float32)
    got = engine.box_blur(dev(a), r).cpu().numpy()
    np.testing.assert_array_equal(got, orc32.box_blur(a, r))


def test_box_blur_rows_that_round(engine, orc32):
    """Rows whose running sums round (the along-x pass's left-to-right
    fallback): tiny samples beside large ones, huge cancelling samples, a
    zero row, denormals; every row must still match hvd-core's order bit for
    bit."""
    rng = np.random.default_rng(12)
    h, w = 8, 700
    a = rng.normal(50, 20, size=(h, w)).astype(np.float32)
    a[0, ::7] = 1e-9                      # below 2^-12 of the row's largest
    a[1, :] = 0.0                         # all zero
    a[2, 10], a[2, 11], a[2, 300] = 1e30, 1.0, -1e30  # cancellation after rounding
    a[3, ::3] = np.float32(1e-40)          # denormals
    a[4, :] = rng.normal(0, 1e-6, w).astype(np.float32)  # a tiny row (still exact)
    a[5, ::2] *= -1
    for r in (1, 2, 5):
        got = engine.box_blur(dev(a), r).cpu().numpy()
        np.testing.assert_array_equal(got, orc32.box_blur(a, r))


def test_box_blur_columns_that_round(engine, orc32):
    """Columns whose running sums round take the along-y pass's top-to-bottom
    fallback, the others its row segments: the same cases down the columns
    of a tall plane (several segments), beside ordinary columns; every output
    must match hvd-core's order bit for bit."""
    rng = np.random.default_rng(14)
    a = rng.normal(50, 20, size=(300, 70)).astype(np.float32)
    a[::7, 0] = 1e-9
    a[:, 1] = 0.0
    a[10, 2], a[11, 2], a[200, 2] = 1e30, 1.0, -1e30
    a[::3, 3] = np.float32(1e-40)
    a[:, 4] = rng.normal(0, 1e-6, 300).astype(np.float32)
    a[::2, 5] *= -1
    a[:, 6] = 2.0 ** np.arange(300) % 7  # small integers: exact in any order
    for r in (1, 2, 8, 31):
        got = engine.box_blur(dev(a), r).cpu().numpy()
        np.testing.assert_array_equal(got, orc32.box_blur(a, r))


def test_chroma_plane(engine):
    rng = np.random.default_rng(13)
    s = rng.normal(size=(9, 31)).astype(np.float32)
    y = rng.normal(size=(9, 31)).astype(np.float32)
    chi = rng.normal(size=(9, 31, 2)).astype(np.float32)
    car = rng.normal(size=(9, 31, 2)).astype(np.float32)
    np.testing.assert_array_equal(engine.chroma_plane(0, s=dev(s), y=dev(y)).cpu().numpy(), s - y)
    want = (chi[..., 0] * car[..., 0]).astype(np.float32) - (chi[..., 1] * car[..., 1]).astype(np.float32)
    np.testing.assert_array_equal(engine.chroma_plane(1, chi=dev(chi), carrier=dev(car)).cpu().numpy(), want)


def _lround(x):
    x = np.asarray(x, np.float64)
    return (np.sign(x) * np.floor(np.abs(x) + 0.5)).astype(np.int64)


@pytest.mark.parametrize("with_y", [True, False])
def test_weave_frame_s16(engine, with_y):
    """Woven row r -> flat block r % 2, field line first_active + r / 2
    (frame_bridge.cpp:203-253); everything else passes the source through."""
    fpd = synth.FRAME_PAL
    fp = fparams(fpd)
    rng = np.random.default_rng(17)
    src = rng.integers(0, 1024, size=synth.frame_line_offset(fpd, fpd["frame_height"])).astype(np.int16)
    g = synth.PAL
    rows, width = 2 * g.lines, g.active_video_end - g.active_video_start
    y = rng.uniform(-10, 110, size=(rows, width)).astype(np.float32)
    ch = rng.uniform(-60, 60, size=(rows, width)).astype(np.float32)
    chi = rng.uniform(-40, 40, size=(rows, width, 2)).astype(np.float32)
    gain = 1.25
    lum, chr_, u, v = [t.cpu().numpy() for t in engine.weave_frame_s16(dev(src), fp, dev(y) if with_y else None,
                                                                      dev(ch), dev(chi), gain)]
    fw, fh, f1 = fpd["frame_width"], fpd["frame_height"], fpd["field1_lines"]
    scale = np.float32((fpd["white_level"] - fpd["blanking_level"]) / 100.0)
    e_lum = np.concatenate([src[synth.frame_line_offset(fpd, L):synth.frame_line_offset(fpd, L) + fw]
                            for L in range(fh)]).astype(np.int16)
    e_chr = np.zeros(fw * fh, np.int16)
    e_u = np.zeros(fw * fh)
    e_v = np.zeros(fw * fh)
    fal, a0 = fpd["first_active_frame_line"] // 2, fpd["active_video_start"]
    for r in range(rows):
        line = (fal + r // 2) if r % 2 == 0 else f1 + fal + r // 2
        if line >= fh:
            continue
        idx = line * fw + a0 + np.arange(width)
        if with_y:
            e_lum[idx] = np.clip(_lround(np.float32(fpd["blanking_level"]) + (y[r] * scale).astype(np.float32)),
                                 0, 1023)
        e_chr[idx] = np.clip(_lround((ch[r] * scale).astype(np.float32)), -1023, 1023)
        e_v[idx] = np.float32(gain) * chi[r, :, 0].astype(np.float64) * float(scale)
        e_u[idx] = np.float32(gain) * (-chi[r, :, 1]).astype(np.float64) * float(scale)
    np.testing.assert_array_equal(lum, e_lum)
    np.testing.assert_array_equal(chr_, e_chr)
    np.testing.assert_array_equal(u, e_u)
    np.testing.assert_array_equal(v, e_v)
