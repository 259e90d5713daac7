"""Seeded synthetic composite fields (NTSC and PAL, 4fsc, uint16 codes).

Signal model (the composite convention of hvd-core,
engine/ntsc_geometry.h:84-88 and lockin.cpp:270-291):

    S(line, x) = Y + Re[chi * c],   chi = V - iU,
    c = exp(i phi)                      NTSC, and PAL lines with s = +1
    c = -exp(-i phi)                    PAL lines with s = -1 (V-switch)
    phi = theta0 + line_advance * line + (pi/2) * x    (x = absolute sample)

NTSC burst is -20 sin(phi) (chi_burst = +20i, the -U axis); PAL burst is
Re[B e^{i pi/4} c] (the swinging +/-135 deg burst).  IRE maps to codes as
code = round(4096 + 409.6 * (ire + 40)), i.e. -40..+100 IRE -> 0x1000..0xF000,
blanking (0 IRE) = 20480, white (100 IRE) = 61440.

Everything is numpy with an explicit seed, so tests and the benchmark see the
same fields on every machine.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

CODE_BLANKING = 20480.0
CODE_WHITE = 61440.0
CODES_PER_IRE = (CODE_WHITE - CODE_BLANKING) / 100.0  # 409.6


@dataclasses.dataclass(frozen=True)
class Geometry:
    """Field geometry (mirrors hvd::FieldGeometry, ntsc_geometry.h:142-267)."""

    field_width: int
    field_height: int
    active_video_start: int
    active_video_end: int
    colour_burst_start: int
    colour_burst_end: int
    first_active_field_line: int
    last_active_field_line: int
    sample_rate: float
    subcarrier_hz: float
    standard: int  # 0 NTSC, 1 PAL, 2 PAL-M

    @property
    def lines(self) -> int:
        last = self.last_active_field_line or self.field_height
        return last - self.first_active_field_line

    @property
    def width(self) -> int:
        return self.active_video_end - self.active_video_start

    @property
    def uses_vswitch(self) -> bool:
        return self.standard in (1, 2)

    def line_advance(self) -> float:
        if self.subcarrier_hz <= 0:
            return {1: 2 * math.pi * 0.7516, 2: 0.5 * math.pi}.get(self.standard, math.pi)
        cyc = self.subcarrier_hz * self.field_width / self.sample_rate
        return 2 * math.pi * (cyc - math.floor(cyc))


FS_NTSC = 4.0 * 315.0e6 / 88.0
FS_PAL = 4.0 * 4433618.75

# NTSC 910x263: active 134..894 x lines 21..263, burst 76..112 (36 samples).
NTSC = Geometry(910, 263, 134, 894, 76, 112, 21, 0, FS_NTSC, 0.0, 0)
# PAL 1135x313: active 185..1107 x lines 22..310, burst 98..138 (40 samples).
PAL = Geometry(1135, 313, 185, 1107, 98, 138, 22, 310, FS_PAL, 0.0, 1)


def _bilinear(grid: np.ndarray, h: int, w: int) -> np.ndarray:
    gh, gw = grid.shape
    ys = np.linspace(0.0, gh - 1.0, h)
    xs = np.linspace(0.0, gw - 1.0, w)
    y0 = np.minimum(np.floor(ys).astype(int), gh - 2)
    x0 = np.minimum(np.floor(xs).astype(int), gw - 2)
    fy = (ys - y0)[:, None]
    fx = (xs - x0)[None, :]
    g00 = grid[y0][:, x0]
    g01 = grid[y0][:, x0 + 1]
    g10 = grid[y0 + 1][:, x0]
    g11 = grid[y0 + 1][:, x0 + 1]
    return (1 - fy) * ((1 - fx) * g00 + fx * g01) + fy * ((1 - fx) * g10 + fx * g11)


@dataclasses.dataclass
class Scene:
    """Ground-truth active picture: luma Y and chroma chi = V - iU (IRE)."""

    y: np.ndarray  # (lines, width) float64
    chi: np.ndarray  # (lines, width) complex128


def make_scene(g: Geometry, rng: np.random.Generator, canvas_pad: int = 0,
               saturation: float = 40.0, line_factor: int = 1) -> Scene:
    """Luma = horizontal ramp + 8 Gaussian blobs clipped to 0..100 IRE;
    chroma = bilinear-upsampled 16x8 grid with |C| <= `saturation` IRE.
    `canvas_pad` enlarges the canvas on every side (for motion crops);
    `line_factor` = 2 renders at frame resolution (both fields' lines)."""
    h = g.lines * line_factor + 2 * canvas_pad
    w = g.width + 2 * canvas_pad
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    luma = 10.0 + 70.0 * xx / max(w - 1, 1)
    for _ in range(8):
        cy, cx = rng.uniform(0, h), rng.uniform(0, w)
        sy, sx = rng.uniform(6, 40), rng.uniform(15, 120)
        amp = rng.uniform(-45, 45)
        luma += amp * np.exp(-0.5 * (((yy - cy) / sy) ** 2 + ((xx - cx) / sx) ** 2))
    luma = np.clip(luma, 0.0, 100.0)
    grid = rng.uniform(-1, 1, size=(8, 16)) + 1j * rng.uniform(-1, 1, size=(8, 16))
    grid *= saturation / np.abs(grid).max()
    chi = _bilinear(grid.real, h, w) + 1j * _bilinear(grid.imag, h, w)
    return Scene(luma, chi)


def carrier(g: Geometry, theta0: float, lines: np.ndarray, xs: np.ndarray,
            vswitch: np.ndarray | None) -> np.ndarray:
    """Effective carrier c over (lines x xs), absolute line/sample indices."""
    phi = theta0 + g.line_advance() * lines[:, None] + 0.5 * math.pi * xs[None, :]
    c = np.exp(1j * phi)
    if vswitch is not None:
        sw = vswitch[lines] < 0
        c[sw] = -np.conj(c[sw])
    return c


def vswitch_pattern(g: Geometry, first: int = 1) -> np.ndarray | None:
    if not g.uses_vswitch:
        return None
    s = np.where(np.arange(g.field_height) % 2 == 0, first, -first)
    return s.astype(np.int8)


def render_field(g: Geometry, scene: Scene, rng: np.random.Generator, *,
                 theta0: float = 0.0, noise_ire: float = 0.5,
                 vswitch_first: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """Compose one raw field (uint16, field_height x field_width) from an
    active-area scene.  Returns (raw, ire_clean) where ire_clean is the
    noise-free composite in IRE."""
    fh, fw = g.field_height, g.field_width
    ire = np.zeros((fh, fw), dtype=np.float64)
    sync_end = 60 if g.standard != 1 else 84
    ire[:, :sync_end] = -40.0
    vs = vswitch_pattern(g, vswitch_first)
    fal = g.first_active_field_line
    lal = g.last_active_field_line or fh
    lines = np.arange(fal, lal)
    # burst on the active lines only
    bx = np.arange(g.colour_burst_start, g.colour_burst_end)
    cb = carrier(g, theta0, lines, bx, vs)
    if g.uses_vswitch:
        burst = (21.43 * np.exp(0.25j * math.pi) * cb).real
    else:
        burst = (20j * cb).real  # = -20 sin(phi)
    ire[fal:lal, g.colour_burst_start:g.colour_burst_end] = burst
    ax = np.arange(g.active_video_start, g.active_video_end)
    ca = carrier(g, theta0, lines, ax, vs)
    ire[fal:lal, g.active_video_start:g.active_video_end] = scene.y + (scene.chi * ca).real
    noisy = ire + rng.normal(0.0, noise_ire, size=ire.shape) if noise_ire > 0 else ire
    raw = np.clip(np.round(4096.0 + 409.6 * (noisy + 40.0)), 0, 65535).astype(np.uint16)
    return raw, ire


def ire_from_raw(raw: np.ndarray) -> np.ndarray:
    """SampleToIre (ntsc_geometry.h:301-304) for the uint16 code mapping."""
    return ((raw.astype(np.float32) - np.float32(CODE_BLANKING)) / np.float32(CODES_PER_IRE)).astype(np.float32)


def ntsc_field(seed: int, noise_ire: float = 0.5, theta0: float = 0.0):
    """One NTSC 910x263 field: (raw uint16, Scene, clean IRE)."""
    rng = np.random.default_rng(seed)
    sc = make_scene(NTSC, rng)
    raw, ire = render_field(NTSC, sc, rng, theta0=theta0, noise_ire=noise_ire)
    return raw, sc, ire


def pal_field(seed: int, noise_ire: float = 0.5, theta0: float = 0.0):
    """One PAL 1135x313 field: (raw uint16, Scene, clean IRE)."""
    rng = np.random.default_rng(seed)
    sc = make_scene(PAL, rng)
    raw, ire = render_field(PAL, sc, rng, theta0=theta0, noise_ire=noise_ire)
    return raw, sc, ire


def ntsc_theta0(field_index: int) -> float:
    """Start phase of consecutive NTSC fields (4-field sequence)."""
    return (field_index % 4) * 1.5 * math.pi


def motion_sequence(seed: int, n_fields: int = 8, g: Geometry = NTSC, noise_ire: float = 0.5,
                    max_global=(1, 3), max_rect_v: int = 8, saturation: float = 40.0):
    """Consecutive fields of one scene under seeded integer motion: a global
    translation (|dy| <= max_global[0], |dx| <= max_global[1] per field) plus
    four 32x16 rectangles with their own integer velocities (<= max_rect_v
    px/field); chroma up to `saturation` IRE.  Returns (raws [n, fh, fw]
    uint16, truth dict)."""
    rng = np.random.default_rng(seed)
    pad = 64
    sc = make_scene(g, rng, canvas_pad=pad, saturation=saturation)
    gdy = int(rng.integers(-max_global[0], max_global[0] + 1))
    gdx = int(rng.integers(-max_global[1], max_global[1] + 1))
    rects = []
    for _ in range(4):
        ry = int(rng.integers(20, g.lines - 40))
        rx = int(rng.integers(40, g.width - 80))
        vy = int(rng.integers(-min(max_rect_v, 4), min(max_rect_v, 4) + 1))
        vx = int(rng.integers(-max_rect_v, max_rect_v + 1))
        lum = float(rng.uniform(5, 95))
        ch = complex(rng.uniform(-25, 25), rng.uniform(-25, 25))
        rects.append((ry, rx, vy, vx, lum, ch))
    raws = []
    for t in range(n_fields):
        oy = pad - gdy * t
        ox = pad - gdx * t
        oy = int(np.clip(oy, 0, 2 * pad))
        ox = int(np.clip(ox, 0, 2 * pad))
        y = sc.y[oy:oy + g.lines, ox:ox + g.width].copy()
        chi = sc.chi[oy:oy + g.lines, ox:ox + g.width].copy()
        for (ry, rx, vy, vx, lum, ch) in rects:
            y0, x0 = ry + vy * t, rx + vx * t
            ya, yb = max(y0, 0), min(y0 + 16, g.lines)
            xa, xb = max(x0, 0), min(x0 + 32, g.width)
            if ya < yb and xa < xb:
                y[ya:yb, xa:xb] = lum
                chi[ya:yb, xa:xb] = ch
        raw, _ = render_field(g, Scene(y, chi), rng, theta0=ntsc_theta0(t), noise_ire=noise_ire)
        raws.append(raw)
    truth = {"global": (gdy, gdx), "rects": rects}
    return np.stack(raws), truth


# ---------------------------------------------------------------------------
# Field-sequential 10-bit frames (the int16 layout of hvd::FrameParams,
# frame_bridge.h:40-112): field 1 on flat lines [0, field1_lines), field 2
# after it; PAL frames carry two extra samples on flat lines 312 and 624.

FRAME_NTSC = dict(standard=0, frame_width=910, frame_height=525, field1_lines=263, active_video_start=134,
                  active_video_end=894, colour_burst_start=76, colour_burst_end=112, first_active_frame_line=42,
                  last_active_frame_line=0, black_level=282.0, white_level=800.0, blanking_level=240.0,
                  chroma_dc=512.0, sample_rate=FS_NTSC, subcarrier_hz=0.0, extra_sample_lines=())
FRAME_PAL = dict(standard=1, frame_width=1135, frame_height=625, field1_lines=313, active_video_start=185,
                 active_video_end=1107, colour_burst_start=98, colour_burst_end=138, first_active_frame_line=44,
                 last_active_frame_line=620, black_level=256.0, white_level=844.0, blanking_level=256.0,
                 chroma_dc=512.0, sample_rate=FS_PAL, subcarrier_hz=0.0, extra_sample_lines=(312, 312, 624, 624))


def frame_line_offset(fp: dict, line: int) -> int:
    return line * fp["frame_width"] + sum(1 for e in fp["extra_sample_lines"] if e < line)


def frames_s16(fp: dict, n: int, seed: int, noise_ire: float = 0.5) -> list:
    """n field-sequential int16 frames of one static scene (field 2 samples
    the scene half a frame line below field 1)."""
    g = PAL if fp["standard"] == 1 else NTSC
    rng = np.random.default_rng(seed)
    sc = make_scene(g, rng, line_factor=2)
    cpi = (fp["white_level"] - fp["blanking_level"]) / 100.0
    fw, fh, f1 = fp["frame_width"], fp["frame_height"], fp["field1_lines"]
    out = []
    for t in range(n):
        buf = np.full(frame_line_offset(fp, fh), int(fp["blanking_level"]), np.int16)
        for blk in range(2):
            field_scene = Scene(sc.y[blk::2], sc.chi[blk::2])
            theta0 = ntsc_theta0(2 * t + blk) if fp["standard"] != 1 else 0.7 * (2 * t + blk)
            _, ire = render_field(g, field_scene, rng, theta0=theta0, noise_ire=0.0,
                                  vswitch_first=1 if blk == 0 else -1)
            if noise_ire > 0:
                ire = ire + rng.normal(0.0, noise_ire, size=ire.shape)
            codes = np.clip(np.round(fp["blanking_level"] + cpi * ire), 0, 1023).astype(np.int16)
            nlines = f1 if blk == 0 else fh - f1
            for y in range(nlines):
                o = frame_line_offset(fp, y + blk * f1)
                buf[o:o + fw] = codes[y]
        out.append(buf)
    return out
