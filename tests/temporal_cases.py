"""Shared scenes for the 3-D (temporal-neighbour) solve: a field and
neighbours that see the same luma and chroma against carriers a quarter and
a half turn away, with noise and per-pixel confidence."""
import math

import numpy as np


def temporal_scene(h, w, seed, n_nbr=2):
    rng = np.random.default_rng(seed)
    r, c = np.mgrid[0:h, 0:w].astype(np.float64)
    car = np.exp(1j * (0.5 * math.pi * c + 0.3 * r)).astype(np.complex64).astype(np.complex128)
    y_true = 30 + 20 * np.sin(0.35 * (r + c))
    chi_true = 10 * np.sin(0.4 * (c - r)) - 6j * np.cos(0.3 * (c + r))
    s = (y_true + (chi_true * car).real).astype(np.float32)
    chi0 = (chi_true + 3 * np.sin(0.9 * c) * (1 - 1j)).astype(np.complex64).astype(np.complex128)
    nbrs = []
    for j in range(n_nbr):
        phase = 0.5 * math.pi * (j + 1)
        cw = np.exp(1j * (0.5 * math.pi * c + 0.3 * r + phase)).astype(np.complex64).astype(np.complex128)
        sw = (y_true + (chi_true * cw).real + rng.normal(0, 0.5, (h, w))).astype(np.float32)
        conf = rng.uniform(0.3, 1.0, (h, w)).astype(np.float32)
        nbrs.append((sw, cw, conf))
    return s, car, chi0, nbrs
