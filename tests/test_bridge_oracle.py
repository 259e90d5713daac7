"""The frame-bridge oracle's plumbing against closed forms (no GPU)."""
import numpy as np

import bridge as OB
from hvd_amd import synth


def test_pal_line_offsets():
    # 625-line PAL: 709 379 samples, the 4 extra on flat lines 312 and 624
    # (frame_bridge.h:44-112); field 2 starts 2 samples late.
    fp = OB.FrameParams(**synth.FRAME_PAL)
    assert fp.buffer_len() == 709379
    assert OB.frame_line_offset(fp, 312) == 312 * 1135
    assert OB.frame_line_offset(fp, 313) == 313 * 1135 + 2
    assert OB.frame_line_offset(fp, 624) == 624 * 1135 + 2
    assert OB.frame_line_offset(fp, 625) == 625 * 1135 + 4


def test_split_roundtrip_and_levels():
    fpd = synth.FRAME_PAL
    fp = OB.FrameParams(**fpd)
    frame = synth.frames_s16(fpd, 1, seed=3)[0]
    top, bot = OB.split_frame_fields(frame, fp)
    cpi = (fp.white_level - fp.blanking_level) / 100.0
    # codes = blanking + ire * cpi exactly (SampleToIre inverse)
    o = OB.frame_line_offset(fp, 313 + 5)
    assert np.allclose(fp.blanking_level + bot[5] * cpi, frame[o:o + fp.frame_width], atol=1e-3)
    # field 2 has one line fewer: its last row reads blanking (0 IRE)
    assert np.all(bot[fp.frame_height - fp.field1_lines:] == 0.0)


def test_field_parity_vote():
    # static scene, field 2 sampled half a frame line below field 1: field 1 is top
    fpd = synth.FRAME_NTSC
    fp = OB.FrameParams(**fpd)
    frame = synth.frames_s16(fpd, 1, seed=4, noise_ire=0.0)[0]
    top, bot = OB.split_frame_fields(frame, fp)
    assert OB.detect_field_parity(top, bot, fp) == 1
    assert OB.detect_field_parity(bot, top, fp) == -1
    flat = np.zeros_like(top)
    assert OB.detect_field_parity(flat, flat, fp) == 0


def test_lround_half_away():
    assert list(OB.lround([0.5, 1.5, -0.5, -1.5, 2.49, -2.51])) == [1, 2, -1, -2, 2, -3]
