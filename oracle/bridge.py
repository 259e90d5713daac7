"""Oracle for hvd-core's frame bridge, 2-D sequence window.

TEST INFRASTRUCTURE ONLY.  Restates frame_bridge.cpp (hvd-core/src):
FrameLineOffset (frame_bridge.h:106-112), SplitFrameFields (:61-99),
DetectFieldParity (:102-149), the window ACC and field-order vote of
DecodeFrameSequenceWindow (:776-795, :759-765) and its parity weave with
ClampU10 / ClampS10 requantisation (:867-945, :25-39).  The per-field decode
is the FP64 oracle's decode_field (sequence.cpp:229-281 with the temporal
terms off).  Integer / byte plumbing is numpy.
"""
from __future__ import annotations

import dataclasses

import math

import numpy as np


@dataclasses.dataclass
class FrameParams:
    standard: int
    frame_width: int
    frame_height: int
    field1_lines: int
    active_video_start: int
    active_video_end: int
    colour_burst_start: int
    colour_burst_end: int
    first_active_frame_line: int
    last_active_frame_line: int
    black_level: float
    white_level: float
    blanking_level: float
    chroma_dc: float
    sample_rate: float
    subcarrier_hz: float = 0.0
    extra_sample_lines: tuple = ()

    def buffer_len(self) -> int:
        return frame_line_offset(self, self.frame_height)


def frame_line_offset(fp: FrameParams, line: int) -> int:
    return line * fp.frame_width + sum(1 for e in fp.extra_sample_lines if e < line)


def field_geometry(fp: FrameParams, Geometry):
    """FieldGeometryFromParams (frame_bridge.cpp:43-58)."""
    return Geometry(fp.frame_width, fp.field1_lines, fp.active_video_start, fp.active_video_end,
                    fp.colour_burst_start, fp.colour_burst_end, fp.first_active_frame_line // 2,
                    fp.last_active_frame_line // 2 if fp.last_active_frame_line else 0, fp.sample_rate,
                    fp.subcarrier_hz, fp.standard)


def split_frame_fields(frame: np.ndarray, fp: FrameParams):
    fw, fh, f1 = fp.frame_width, fp.frame_height, fp.field1_lines
    f2 = fh - f1
    cpi = np.float32((fp.white_level - fp.blanking_level) / 100.0)
    top = np.full((f1, fw), fp.blanking_level, np.float32)
    bot = np.full((f1, fw), fp.blanking_level, np.float32)
    for y in range(f1):
        o = frame_line_offset(fp, y)
        top[y] = frame[o:o + fw].astype(np.float32)
        if y < f2:
            o = frame_line_offset(fp, f1 + y)
            bot[y] = frame[o:o + fw].astype(np.float32)
    to_ire = lambda a: ((a - np.float32(fp.blanking_level)) / cpi).astype(np.float32)  # noqa: E731
    return to_ire(top), to_ire(bot)


def detect_field_parity(b1, b2, fp: FrameParams) -> int:
    f1 = fp.field1_lines
    fal = fp.first_active_frame_line // 2
    lal = fp.last_active_frame_line // 2 if fp.last_active_frame_line else f1
    a0, a1 = fp.active_video_start, fp.active_video_end
    lines, width = lal - fal, a1 - a0
    if lines < 8 or width < 32:
        return 0
    ystep = max(1, lines // 120)
    ep = em = 0.0
    n = 0
    q = np.float32(0.25)
    h = np.float32(0.5)
    for r in range(fal + 1, lal - 1, ystep):
        for x in range(a0, a1 - 3, 16):
            lp = lambda p, row: q * (((p[row, x] + p[row, x + 1]) + p[row, x + 2]) + p[row, x + 3])  # noqa: E731
            b, ac, au, ad = lp(b2, r), lp(b1, r), lp(b1, r - 1), lp(b1, r + 1)
            ep += abs(float(np.float32(b - h * np.float32(ac + ad))))
            em += abs(float(np.float32(b - h * np.float32(ac + au))))
            n += 1
    if n == 0:
        return 0
    ep /= n
    em /= n
    lo, hi = min(ep, em), max(ep, em)
    if hi <= 1e-9 or (hi - lo) / hi < 0.02:
        return 0
    return 1 if ep < em else -1


def lround(x):
    """std::lround: half away from zero."""
    x = np.asarray(x, np.float64)
    return (np.sign(x) * np.floor(np.abs(x) + 0.5)).astype(np.int64)


def decode_frame_sequence_window(orc, frames, core_begin, core_end, fp: FrameParams, cfg, Geometry, *,
                                 chroma_gain=1.0, acc=True, field_order=0, monochrome=False, chroma_phase_deg=0.0):
    """Returns (list of (luma, chroma, u, v) flat planes per core frame, gain,
    field 1 is bottom)."""
    g = field_geometry(fp, Geometry)
    pairs, f1_bottom, gain = _window_decisions(orc, frames, fp, g, chroma_gain, acc, field_order)
    dec = []
    for top, bot in pairs:
        for fld in (top, bot):
            r = orc.decode_field(fld, g, cfg, chroma_phase_deg)
            fal = g.first_active_field_line
            vs = r["vswitch"][fal:fal + g.lines] if g.uses_vswitch else None
            car = orc.make_carrier(r["theta"][fal:fal + g.lines].astype(np.float32), vs, g)
            dec.append((r["y"], r["chi"], car))
    return _weave(frames, dec, core_begin, core_end, fp, g, gain, f1_bottom, monochrome), float(gain), bool(f1_bottom)


def _window_decisions(orc, frames, fp, g, chroma_gain, acc, field_order):
    pairs = [split_frame_fields(fr, fp) for fr in frames]
    amps = []
    for top, bot in pairs:
        amps += [orc.burst_amplitude_ire(top, g), orc.burst_amplitude_ire(bot, g)]
    if field_order == 0:
        vote = sum(detect_field_parity(t, b, fp) for t, b in pairs)
        f1_bottom = vote < 0
    else:
        f1_bottom = field_order == 2
    gain = 1.0
    if acc:
        med = sorted(amps)[len(amps) // 2]
        gain = orc.acc_gain(med, 21.43 if g.standard == 1 else 20.0)
    return pairs, bool(f1_bottom), np.float32(np.float32(chroma_gain) * np.float32(gain))


def decode_frame_sequence_window_3d(orc64, orc32, frames, core_begin, core_end, fp: FrameParams, wcfg, Geometry, *,
                                    chroma_gain=1.0, acc=True, field_order=0, monochrome=False,
                                    chroma_phase_deg=0.0, diag=None, make_carrier=None, holographic_init=None):
    """DecodeFrameSequenceWindow (frame_bridge.cpp:722-947) with the window
    driver: PrepareFieldObs per field (sequence.cpp:229-268; lock-in in
    FP64, hvd-core's float MakeCarrier), parities from the measured field
    order (:759-774), HolographicInit (FP64, stored as complex64) and the
    FP32 window driver restatement (oracle/sequence3d.py, bit-exact with
    hvd-core's own DecodeFieldWindowWithInits).  wcfg: a
    sequence3d.WindowCfg.  make_carrier(theta, vswitch, g) replaces
    hvd-core's MakeCarrier when given (a checker that feeds the oracle the
    carrier planes of the implementation under test: the 3-D passes turn
    sub-ulp differences of sinf/cosf into rare decision flips);
    holographic_init(s, carrier, sample_rate) -> (Y0, chi0) likewise
    replaces the oracle's FP64 init.  Returns (frames, gain, field 1 is
    bottom)."""
    import sequence3d as S3
    g = field_geometry(fp, Geometry)
    pairs, f1_bottom, gain = _window_decisions(orc64, frames, fp, g, chroma_gain, acc, field_order)
    fal = g.first_active_field_line
    fields, parities, inits, inits_y, cars = [], [], [], [], []
    p0 = 1 if f1_bottom else 0
    for top, bot in pairs:
        for k, fld in enumerate((top, bot)):
            s, _, th, vs = orc64.prepare_field(fld, g)  # untrimmed theta: the trim below is hvd-core's float one
            vs_act = vs[fal:fal + g.lines] if g.uses_vswitch else None
            th_act = th[fal:fal + g.lines].astype(np.float32)
            if chroma_phase_deg != 0.0:
                # ApplyChromaPhase (ntsc_geometry.h:265-276), float adds
                off = np.float32(-np.float32(chroma_phase_deg) * np.float32(math.pi) / np.float32(180.0))
                sgn = np.where((vs_act if vs_act is not None else np.ones_like(th_act)) < 0, -off, off)
                th_act = (th_act + sgn.astype(np.float32)).astype(np.float32)
            car = (make_carrier or orc32.make_carrier)(th_act, vs_act, g).astype(np.complex64)
            if holographic_init is not None:
                y0, chi0 = holographic_init(s.astype(np.float32), car, g.sample_rate)
            else:
                y0, chi0 = orc64.holographic_init(s, car.astype(np.complex128), g.sample_rate)
            fields.append((s.astype(np.float32), car))
            inits.append(chi0.astype(np.complex64))
            inits_y.append(y0.astype(np.float32))
            parities.append(p0 if k == 0 else 1 - p0)
            cars.append(car)
    ys, chis = S3.decode_window_with_inits(orc32, fields, parities, inits, wcfg, is_pal=g.uses_vswitch,
                                          inits_y=inits_y, diag=diag)
    dec = [(y, c.astype(np.complex128), car.astype(np.complex128)) for y, c, car in zip(ys, chis, cars)]
    return _weave(frames, dec, core_begin, core_end, fp, g, gain, f1_bottom, monochrome), float(gain), f1_bottom


def _weave(frames, dec, core_begin, core_end, fp, g, gain, f1_bottom, monochrome):
    """The core frames' weave and requantisation (frame_bridge.cpp:867-945)."""
    fw, fh, f1 = fp.frame_width, fp.frame_height, fp.field1_lines
    scale = np.float32((fp.white_level - fp.blanking_level) / 100.0)
    p0 = 1 if f1_bottom else 0
    outs = []
    fal = g.first_active_field_line
    a0 = fp.active_video_start
    for t in range(core_begin, core_end):
        luma = np.zeros(fw * fh, np.int16)
        for line in range(fh):
            o = frame_line_offset(fp, line)
            luma[line * fw:(line + 1) * fw] = frames[t][o:o + fw]
        chroma = np.zeros(fw * fh, np.int16)
        u = np.zeros(fw * fh)
        v = np.zeros(fw * fh)
        lines = dec[2 * t][0].shape[0]
        aw = dec[2 * t][0].shape[1]
        for r in range(2 * lines):
            use_first = (r % 2) == p0
            y, chi, car = dec[2 * t] if use_first else dec[2 * t + 1]
            fl = r // 2
            flat = fal + fl if r % 2 == 0 else f1 + fal + fl
            if flat >= fh:
                continue
            idx = flat * fw + a0 + np.arange(aw)
            yi = y[fl].astype(np.float32)
            ch = np.zeros(aw, complex) if monochrome else chi[fl]
            lum = np.float32(fp.blanking_level) + yi * scale
            luma[idx] = np.clip(lround(lum), 0, 1023)
            re = (ch * car[fl]).real.astype(np.float32) * scale
            chroma[idx] = np.clip(lround(re), -1023, 1023)
            v[idx] = float(gain) * ch.real * float(scale)
            u[idx] = float(gain) * (-ch.imag) * float(scale)
        outs.append((luma, chroma, u, v))
    return outs


# ---- the stage's woven picture and RGB24 export ---------------------------
# orc-plugin_hvd/src/hvd_chroma_decoder_stage.cpp: ReorderToWoven (:316-413),
# YuvToRgb8 (:79-88), the luma-excursion normalisation of
# WovenToPreviewImage / WriteWovenAsRgb24 (:500-638).

def reorder_to_woven(luma, u, v, fp: FrameParams, full_raster=False, field=-1):
    """Field-sequential YcFrameS16 planes (frame_height x frame_width) ->
    woven (y, u, v) float64 planes, raster order or one field's lines."""
    fw, fh, f1 = fp.frame_width, fp.frame_height, fp.field1_lines
    a0 = 0 if full_raster else max(0, fp.active_video_start)
    a1 = fw if full_raster else (fp.active_video_end if fp.active_video_end > a0 else fw)
    y0 = 0 if full_raster else max(0, fp.first_active_frame_line)
    y1 = fh if full_raster else (fp.last_active_frame_line if fp.last_active_frame_line > y0 else fh)
    width = max(0, a1 - a0)
    planes = [np.asarray(p).reshape(fh, fw) for p in (luma, u, v)]
    if field in (0, 1):
        nfl = f1 if field == 0 else fh - f1
        rows = [(fl if field == 0 else f1 + fl) for fl in range(nfl) if y0 <= 2 * fl + field < y1]
    else:
        rows = []
        for frame_line in range(y0, y1):
            rows.append(frame_line // 2 if frame_line % 2 == 0 else f1 + frame_line // 2)
    height = len(rows)
    if width == 0 or height == 0:
        return (np.zeros((0, 0)),) * 3
    cols = np.arange(a0, a0 + width)
    ok_c = (cols >= 0) & (cols < fw)
    out = [np.zeros((height, width)) for _ in range(3)]
    for row, fl in enumerate(rows):
        if fl < 0 or fl >= fh:  # a raster row past the frame stays zero
            continue
        for o, p in zip(out, planes):
            o[row, ok_c] = p[fl, cols[ok_c]].astype(np.float64)
    return tuple(out)


def yuv_to_rgb8(ny, nu, nv):
    """YuvToRgb8 on normalised Y, U, V (float64 arrays) -> uint8 [..., 3]."""
    chans = (ny + 1.13983 * nv, ny - 0.39465 * nu - 0.58060 * nv, ny + 2.03211 * nu)
    return np.stack([(np.clip(c, 0.0, 1.0) * 255.0 + 0.5).astype(np.uint8) for c in chans], -1)


def woven_rgb24(oy, ou, ov, fp: FrameParams):
    rng = float(np.float32(fp.white_level) - np.float32(fp.black_level))  # float difference
    black = float(np.float32(fp.black_level))
    return yuv_to_rgb8((oy - black) / rng, ou / rng, ov / rng)
