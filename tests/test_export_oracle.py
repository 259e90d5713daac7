"""Pins of the oracle's stage-side export (orc-plugin_hvd/src/
hvd_chroma_decoder_stage.cpp: ReorderToWoven :316-413, YuvToRgb8 :79-88,
the luma-excursion normalisation of :500-638) by what fixes them: the
woven row order written out for a small frame, and the YUV->RGB matrix as
the inverse of the textbook BT.601 encode (Y' = 0.299 R + 0.587 G +
0.114 B, U = 0.492 (B - Y'), V = 0.877 (R - Y'), Poynton eq. 28.1/28.6,
the equations the stage's comment names)."""
import numpy as np
import pytest

import bridge as B


def _fp(fw=8, fh=10, f1=5, a0=2, a1=6, y0=2, y1=8, black=280.0, white=800.0, blank=240.0):
    return B.FrameParams(standard=0, frame_width=fw, frame_height=fh, field1_lines=f1, active_video_start=a0,
                         active_video_end=a1, colour_burst_start=0, colour_burst_end=0,
                         first_active_frame_line=y0, last_active_frame_line=y1, black_level=black,
                         white_level=white, blanking_level=blank, chroma_dc=0.0, sample_rate=4.0 * 3579545.0)


def _tagged(fp):
    # luma[flat_line, col] = 100 flat_line + col: every output names its source
    fl, col = np.mgrid[0:fp.frame_height, 0:fp.frame_width]
    return (100 * fl + col).astype(np.int16), (0.5 * fl).astype(np.float64), (0.25 * col).astype(np.float64)


def test_woven_rows_alternate_fields():
    fp = _fp()
    y, u, v = B.reorder_to_woven(*_tagged(fp), fp)
    # frame lines 2..7 -> flat lines 1, 6, 2, 7, 3, 8 (field 2 starts at flat 5)
    assert y.shape == (6, 4)
    assert list(y[:, 0] // 100) == [1, 6, 2, 7, 3, 8]
    assert list(y[0, :] % 100) == [2, 3, 4, 5]
    assert np.array_equal(u[:, 0], 0.5 * np.array([1, 6, 2, 7, 3, 8]))


def test_woven_single_field_and_full_raster():
    fp = _fp()
    y1, _, _ = B.reorder_to_woven(*_tagged(fp), fp, field=1)
    assert list(y1[:, 0] // 100) == [6, 7, 8]  # field lines 1..3 of field 2 (frame lines 3, 5, 7)
    y0, _, _ = B.reorder_to_woven(*_tagged(fp), fp, field=0)
    assert list(y0[:, 0] // 100) == [1, 2, 3]
    yf, _, _ = B.reorder_to_woven(*_tagged(fp), fp, full_raster=True)
    assert yf.shape == (10, 8)
    assert list(yf[:, 0] // 100) == [0, 5, 1, 6, 2, 7, 3, 8, 4, 9]


def _encode(rgb):
    """Textbook BT.601 R'G'B' -> Y'UV (Poynton eq. 28.1 / 28.6)."""
    r, g, b = rgb[..., 0], rgb[..., 1], rgb[..., 2]
    y = 0.299 * r + 0.587 * g + 0.114 * b
    return y, 0.492 * (b - y), 0.877 * (r - y)


@pytest.mark.parametrize("black,white", [(280.0, 800.0), (256.0, 852.0)])
def test_rgb_inverts_the_bt601_encode(black, white):
    fp = _fp(black=black, white=white)
    rgb = np.random.default_rng(3).uniform(0, 1, (500, 3))
    ny, nu, nv = _encode(rgb)
    rng = white - black
    got = B.woven_rgb24(black + ny * rng, nu * rng, nv * rng, fp).astype(int)
    # the 5-digit matrix inverts the encode to about 1e-3 of full scale
    assert np.abs(got - np.round(rgb * 255)).max() <= 1


def test_rgb_black_white_grey_and_clamps():
    fp = _fp()
    rng = fp.white_level - fp.black_level
    y = np.array([fp.black_level, fp.white_level, fp.black_level + 0.5 * rng, fp.white_level + 100.0, 0.0])
    z = np.zeros_like(y)
    got = B.woven_rgb24(y, z, z, fp)
    assert got[0].tolist() == [0, 0, 0] and got[1].tolist() == [255, 255, 255]
    assert got[2].tolist() == [128, 128, 128]  # 0.5 * 255 + 0.5 rounds up
    assert got[3].tolist() == [255, 255, 255] and got[4].tolist() == [0, 0, 0]
