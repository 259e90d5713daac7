"""Shared windows for the 3-D window driver (DecodeFieldWindowWithInits,
hvd-core engine/sequence.cpp:371-1114): interlaced fields of one frame-grid
scene under the NTSC 4-field carrier sequence (a half turn per line, three
quarter turns per field) or a PAL-like 0.7516-turn line advance with a
V-switch, with an optional per-field translation and an init chi0 that is
the truth plus a structured error (the style of hvd-core's sequence_test)."""
import math

import numpy as np

F = np.float32


def window(lines=48, width=128, nf=8, noise=0.8, seed=7, ambiguous=True, motion=(0, 0), pal=False,
           amb_region=None):
    """amb_region: (y0, y1, x0, x1) of the frame grid that alone carries the
    ambiguous luma texture (default: all of it)."""
    rng = np.random.default_rng(seed)
    fl = 2 * lines
    pad = 24
    y, x = np.mgrid[0:fl + 2 * pad, 0:width + 2 * pad].astype(F)
    yf = 30 + 22 * np.sin(F(0.06) * (y + 2 * x))
    if ambiguous:
        tex = 14 * np.cos(F(math.pi / 2) * x + F(0.4) * y) * np.sin(F(0.15) * y)
        if amb_region is not None:
            y0, y1, x0, x1 = amb_region
            tex = tex * ((y >= y0 + pad) & (y < y1 + pad) & (x >= x0 + pad) & (x < x1 + pad))
        yf = yf + tex
    inside = (x > 40 + pad) & (x < 90 + pad) & (y > 24 + pad) & (y < 72 + pad)
    cf = np.where(inside, F(12) - 8j, F(3) + 2j).astype(np.complex64)
    fields, parities, inits, truth = [], [], [], []
    for f in range(nf):
        p = f % 2
        r, xx = np.mgrid[0:lines, 0:width]
        sy = np.clip(2 * r + p + pad - motion[0] * f, 0, fl + 2 * pad - 1)
        sx = np.clip(xx + pad - motion[1] * f, 0, width + 2 * pad - 1)
        if pal:
            phi = F(math.pi / 2) * xx + F(2 * math.pi * 0.7516) * r + F(0.7 * f)
            sw = np.where((r + f) % 2 == 0, 1.0, -1.0)
        else:
            phi = F(math.pi / 2) * xx + F(math.pi) * r + F(f * 3 * math.pi / 2)
            sw = np.ones_like(r, dtype=np.float64)
        phi = phi.astype(F)
        c = (np.cos(phi) + 1j * np.sin(phi)).astype(np.complex64)
        c = np.where(sw < 0, -np.conj(c), c).astype(np.complex64)
        ct = cf[sy, sx]
        s = (yf[sy, sx] + (ct * c).real + rng.normal(0, noise, (lines, width))).astype(F)
        e = (4 * np.sin(F(0.8) * xx) * np.cos(F(0.6) * r)).astype(F)
        inits.append((ct + e - 1j * e).astype(np.complex64))
        fields.append((s, c))
        parities.append(p)
        truth.append(ct)
    return fields, parities, inits, truth
