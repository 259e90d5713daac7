"""Shared inputs for the temporal neighbours' warps (temporal.cpp): a tile
motion field with outliers and exact half-pel ties, and planes to warp."""
import numpy as np


def motion_field(th, tw, seed, spread=6.0):
    rng = np.random.default_rng(seed)
    ty, tx = np.mgrid[0:th, 0:tw].astype(np.float64)
    mdy = np.round(2 * np.sin(0.7 * ty + 0.3 * tx) + rng.normal(0, 0.6, (th, tw)), 2)
    mdx = np.round(spread * np.cos(0.4 * tx - 0.2 * ty) + rng.normal(0, 0.6, (th, tw)), 2)
    k = rng.choice(th * tw, size=max(1, th * tw // 12), replace=False)
    mdy.flat[k] += rng.choice([-12.0, 12.0], size=k.size)  # outliers for the median snap
    mdx.flat[k[: k.size // 2]] = 2.5                          # half-pel ties for lround
    conf = rng.uniform(0.0, 1.0, (th, tw))
    return mdy.astype(np.float32), mdx.astype(np.float32), conf.astype(np.float32)


def planes(h, w, seed):
    rng = np.random.default_rng(seed)
    r, c = np.mgrid[0:h, 0:w].astype(np.float64)
    comp = (40 + 20 * np.sin(0.3 * r + 0.11 * c) + rng.normal(0, 1, (h, w))).astype(np.float32)
    car = np.exp(1j * (0.5 * np.pi * c + 0.2 * r)).astype(np.complex64)
    z = ((5 + 3 * np.sin(0.05 * c)) * np.exp(1j * 0.02 * (r + c)) + rng.normal(0, 0.5, (h, w))
         + 1j * rng.normal(0, 0.5, (h, w))).astype(np.complex64)
    return comp, car, z
