"""Oracle of hvd-core's selective 3-D window (frame_bridge.cpp:385-720):
AmbiguousBoxes (:399-621), CropObs (:623-636) and DecodeWindowSelective
(:638-720) with its feather blend.

TEST INFRASTRUCTURE ONLY.  Restated in numpy in hvd-core's order of
operations: the count-normalised box means add their window terms in order
in float and divide by the count; the tile energies add their pixels in
raster order in double; |chi| is std::abs of std::complex<float>, which is
glibc's hypotf: the double square root of the double sum of squares,
rounded to float (the equality with numpy's float32 hypot, which calls the
same libm, is checked in the tests).  The windows themselves are decoded by
the window driver's oracle (sequence3d.decode_window_with_inits).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import pathlib

import numpy as np

import sequence3d as S3

F = np.float32
TILE, HALO_X, HALO_Y, MAX_BANDS = 16, 48, 16, 4
REFLIB_FB = pathlib.Path(__file__).resolve().parent / "_ref" / "libhvdref_fb.so"


@dataclasses.dataclass
class Box:
    """SelBox (:389-392): field coordinates [fy0, fy1) x [x0, x1)."""
    fy0: int = 0
    fy1: int = 0
    x0: int = 0
    x1: int = 0
    valid: bool = False


def cabs(chi):
    """std::abs(std::complex<float>) = hypotf: sqrt in double, rounded to float."""
    re = chi.real.astype(F).astype(np.float64)
    im = chi.imag.astype(F).astype(np.float64)
    return np.sqrt(re * re + im * im).astype(F)


def _count_mean(a, axis):
    """One pass of box2d (:415-444): window d in [-k/2, k - k/2) with k =
    TILE / 2, edge-truncated; float sum in window order / the count."""
    k = TILE // 2
    n_ax = a.shape[axis]
    acc = np.zeros_like(a, F)
    cnt = np.zeros(a.shape, np.int32)
    idx = np.arange(n_ax)
    for d in range(-(k // 2), k - k // 2):
        src = idx + d
        ok = (src >= 0) & (src < n_ax)
        take = np.take(a, np.clip(src, 0, n_ax - 1), axis=axis)
        shape = [1, 1]
        shape[axis] = n_ax
        okb = np.broadcast_to(ok.reshape(shape), a.shape)
        acc = np.where(okb, (acc + take).astype(F), acc)
        cnt = cnt + okb
    return (acc / np.maximum(cnt, 1).astype(F)).astype(F)


def box_mean(a):
    """box2d: along y (each column) first, then along x."""
    return _count_mean(_count_mean(np.asarray(a, F), 0), 1)


def tile_scores(Y, chi):
    """Per 16 x 16 tile sqrt(mean (Y - box(Y))^2 * mean (|chi| - box|chi|)^2)
    (:446-478): the tile's pixels in raster order, summed in double."""
    Y = np.asarray(Y, F)
    h, w = Y.shape
    ya, ca = Y, cabs(np.asarray(chi))
    dy = (ya - box_mean(ya)).astype(F).astype(np.float64)
    dc = (ca - box_mean(ca)).astype(F).astype(np.float64)
    th, tw = (h + TILE - 1) // TILE, (w + TILE - 1) // TILE
    texY = np.zeros((th, tw))
    texC = np.zeros((th, tw))
    cnt = np.zeros((th, tw), np.int64)
    for r in range(TILE):
        for c in range(TILE):
            ys, xs = np.arange(th) * TILE + r, np.arange(tw) * TILE + c
            ok = (ys[:, None] < h) & (xs[None, :] < w)
            yy, xx = np.minimum(ys, h - 1)[:, None], np.minimum(xs, w - 1)[None, :]
            texY = np.where(ok, texY + dy[yy, xx] * dy[yy, xx], texY)
            texC = np.where(ok, texC + dc[yy, xx] * dc[yy, xx], texC)
            cnt = cnt + ok
    n = np.maximum(cnt, 1).astype(np.float64)
    return np.sqrt((texY / n) * (texC / n)).astype(F)


def boxes_from_scores(score, h, w, max_area):
    """The decision part of AmbiguousBoxes (:479-620) on the tile scores."""
    th, tw = score.shape
    flat = score.ravel()
    med = F(np.sort(flat)[flat.size // 2])  # nth_element at size / 2
    thr = F(4.0) * med
    flagged = score >= thr
    dense = np.zeros_like(flagged)
    for ty in range(th):
        for tx in range(tw):
            if not flagged[ty, tx]:
                continue
            nb = 0
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if dy == 0 and dx == 0:
                        continue
                    yy, xx = ty + dy, tx + dx
                    if 0 <= yy < th and 0 <= xx < tw:
                        nb += int(flagged[yy, xx])
            border = ty == 0 or ty == th - 1 or tx == 0 or tx == tw - 1
            dense[ty, tx] = (not border) and nb >= 2
    bands, start, prev = [], -1, -1
    for ty in range(th):
        if int(dense[ty].sum()) >= 2:
            if start < 0:
                start = ty
            elif ty - prev > 1:
                bands.append([start, prev])
                start = ty
            prev = ty
    if start >= 0:
        bands.append([start, prev])
    if not bands:
        return [Box()] if med > F(3.0) else []
    while len(bands) > MAX_BANDS:
        bi, best = 0, None
        for i in range(len(bands) - 1):
            gap = bands[i + 1][0] - bands[i][1]
            if best is None or gap < best:
                best, bi = gap, i
        bands[bi][1] = bands[bi + 1][1]
        del bands[bi + 1]
    boxes = []
    for tr0, tr1 in bands:
        colf = dense[tr0:tr1 + 1].any(axis=0)
        y0raw = max(0, tr0 * TILE - HALO_Y)
        y0 = y0raw - y0raw % 2
        y1 = min(h, (tr1 + 1) * TILE + HALO_Y)
        runs, r0, rp = [], -1, -1
        for tx in range(tw):
            if not colf[tx]:
                continue
            if r0 < 0:
                r0 = rp = tx
                continue
            if tx - rp <= 2:
                rp = tx
                continue
            runs.append((r0, rp))
            r0 = rp = tx
        if r0 >= 0:
            runs.append((r0, rp))
        first = len(boxes)
        for c0, c1 in runs:
            if c1 - c0 < 1:
                continue
            x0 = max(0, c0 * TILE - HALO_X)
            x1 = min(w, (c1 + 1) * TILE + HALO_X)
            if len(boxes) > first and x0 <= boxes[-1].x1:
                boxes[-1].x1 = x1
            else:
                boxes.append(Box(y0, y1, x0, x1, True))
    total = sum(float(b.fy1 - b.fy0) * (b.x1 - b.x0) for b in boxes)
    if total > float(F(max_area)) * h * w:
        return [Box()]
    return boxes


def ambiguous_boxes(Y, chi, max_area):
    """AmbiguousBoxes (:399-621): [] = clean (stay 2-D), [Box()] (invalid) =
    ambiguity everywhere (full 3-D), else the valid boxes."""
    Y = np.asarray(Y, F)
    h, w = Y.shape
    return boxes_from_scores(tile_scores(Y, chi), h, w, max_area)


def feather_blend(Yd, Cd, Y3, C3, b: Box):
    """DecodeWindowSelective's blend of one box (:693-718), in place on the
    2-D field (Yd float32, Cd complex64) from the box's 3-D decode."""
    bh, bw = b.fy1 - b.fy0, b.x1 - b.x0
    hy, hx = min(16, bh // 3), min(48, bw // 3)
    y = np.arange(bh)
    x = np.arange(bw)
    ry = np.ones(bh, F)
    if hy > 0:
        lo, hi = y < hy, (y >= bh - hy) & ~(y < hy)
        ry[lo] = (y[lo].astype(F) / F(hy)).astype(F)
        ry[hi] = ((bh - 1 - y[hi]).astype(F) / F(hy)).astype(F)
    rx = np.ones(bw, F)
    if hx > 0:
        lo, hi = x < hx, (x >= bw - hx) & ~(x < hx)
        rx[lo] = (x[lo].astype(F) / F(hx)).astype(F)
        rx[hi] = ((bw - 1 - x[hi]).astype(F) / F(hx)).astype(F)
    m = (ry[:, None] * rx[None, :]).astype(F)
    om = (F(1.0) - m).astype(F)
    sl = (slice(b.fy0, b.fy1), slice(b.x0, b.x1))
    Yd[sl] = ((Yd[sl] * om).astype(F) + (np.asarray(Y3, F) * m).astype(F)).astype(F)
    cr = ((Cd[sl].real.astype(F) * om).astype(F) + (C3.real.astype(F) * m).astype(F)).astype(F)
    ci = ((Cd[sl].imag.astype(F) * om).astype(F) + (C3.imag.astype(F) * m).astype(F)).astype(F)
    Cd[sl] = cr + 1j * ci


def decode_window_selective(orc, fields, parities, inits_chi, inits_y, cfg, init_fn, is_pal=False,
                            max_area=0.45):
    """DecodeWindowSelective (:638-720) over fields [(S, carrier)] with their
    inits; init_fn(s, carrier) -> (y0, chi0) is HolographicInit for the
    crops.  Returns (Ys, chis, boxes)."""
    c2d = dataclasses.replace(cfg, temporal_strength=-1.0)
    Ys, chis = S3.decode_window_with_inits(orc, fields, parities, inits_chi, c2d, is_pal, inits_y=inits_y)
    Ys = [np.array(y, F) for y in Ys]
    chis = [np.array(c, np.complex64) for c in chis]
    boxes = ambiguous_boxes(Ys[0], chis[0], max_area)
    if not boxes:
        return Ys, chis, boxes
    if not boxes[0].valid:
        Ys, chis = S3.decode_window_with_inits(orc, fields, parities, inits_chi, cfg, is_pal, inits_y=inits_y)
        return list(Ys), list(chis), boxes
    dec3 = []
    for b in boxes:
        crop = [(np.ascontiguousarray(s[b.fy0:b.fy1, b.x0:b.x1]), np.ascontiguousarray(car[b.fy0:b.fy1, b.x0:b.x1]))
                for s, car in fields]
        ci = [init_fn(s, car) for s, car in crop]
        dec3.append(S3.decode_window_with_inits(orc, crop, parities, [c for _, c in ci], cfg, is_pal,
                                                inits_y=[y for y, _ in ci]))
    for b, (Y3s, C3s) in zip(boxes, dec3):
        for j in range(len(Ys)):
            feather_blend(Ys[j], chis[j], Y3s[j], C3s[j], b)
    return Ys, chis, boxes


class RefBoxes:
    """hvd-core's own AmbiguousBoxes (oracle/_ref/libhvdref_fb.so)."""

    def __init__(self):
        if not REFLIB_FB.exists():
            raise FileNotFoundError(str(REFLIB_FB))
        # lazy binding: the unit's engine and FFT callees are not in the
        # library and AmbiguousBoxes reaches none of them
        self.lib = C.CDLL(str(REFLIB_FB), mode=os.RTLD_LAZY)

    def boxes(self, Y, chi, max_area):
        Y = np.ascontiguousarray(Y, F)
        h, w = Y.shape
        cb = np.zeros((h, 2 * w), F)
        cb[:, 0::2], cb[:, 1::2] = np.asarray(chi).real, np.asarray(chi).imag
        out = np.zeros(4 * 256, np.int32)
        fp = C.POINTER(C.c_float)
        n = self.lib.ref_ambiguous_boxes(Y.ctypes.data_as(fp), cb.ctypes.data_as(fp), h, w, C.c_float(max_area),
                                         out.ctypes.data_as(C.POINTER(C.c_int)), 256)
        if n == 0:
            return []
        if n < 0:
            return [Box()]
        return [Box(*map(int, out[4 * k:4 * k + 4]), True) for k in range(n)]
