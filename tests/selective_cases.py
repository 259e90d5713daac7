"""Decoded-field scenes for the selective 3-D window's box detection
(hvd-core frame_bridge.cpp:399-621): smooth luma and flat chroma with a
little noise, plus textured patches (luma near the subcarrier and chroma
magnitude ripple) on chosen 16 x 16 tiles, so the tiles' ambiguity scores
put chosen tiles above 4 x the median."""
import numpy as np

F = np.float32

# name -> (lines, width, patches [(ty0, ty1, tx0, tx1) inclusive tile ranges], texture gain)
SCENES = {
    "clean": (242, 760, [], 0.0),
    "global": (242, 760, [(0, 15, 0, 47)], 6.0),
    "one_band": (242, 760, [(5, 6, 10, 20)], 4.0),
    "three_bands": (242, 760, [(1, 2, 5, 12), (7, 8, 20, 30), (13, 14, 3, 9)], 4.0),
    "six_bands_merged": (242, 760, [(1, 1, 4, 8), (3, 3, 10, 14), (5, 5, 4, 8), (8, 8, 20, 26), (11, 11, 30, 36),
                                    (13, 13, 4, 9)], 4.0),
    "two_runs": (242, 760, [(6, 7, 4, 9), (6, 7, 20, 28)], 4.0),
    "runs_merged_by_margin": (242, 760, [(6, 7, 4, 9), (6, 7, 13, 18)], 4.0),
    "stripe": (242, 760, [(2, 12, 20, 20)], 4.0),
    "border_only": (242, 760, [(0, 0, 0, 47), (15, 15, 0, 47)], 4.0),
    "overflow": (242, 760, [(1, 13, 1, 40)], 4.0),
    "pal": (288, 922, [(4, 6, 30, 44)], 4.0),
    "small": (96, 192, [(2, 3, 3, 8)], 4.0),
}


def scene(name, seed=5):
    h, w, patches, gain = SCENES[name]
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    Y = 40 + 15 * np.sin(0.02 * x + 0.03 * y) + rng.normal(0, 0.3, (h, w))
    chi = (6 + 3j) + 0.3 * (rng.normal(0, 1, (h, w)) + 1j * rng.normal(0, 1, (h, w)))
    for ty0, ty1, tx0, tx1 in patches:
        sl = (slice(16 * ty0, min(h, 16 * (ty1 + 1))), slice(16 * tx0, min(w, 16 * (tx1 + 1))))
        Y[sl] += gain * 4 * np.cos(0.5 * np.pi * x[sl] + 0.7 * y[sl])
        chi[sl] *= 1 + 0.15 * gain * np.cos(1.3 * x[sl] - 0.9 * y[sl])
    return Y.astype(F), chi.astype(np.complex64)


def noise_scene(h, w, seed):
    """Decisions on random content: tiles near the threshold by chance."""
    rng = np.random.default_rng(seed)
    Y = rng.normal(50, 5, (h, w))
    chi = rng.normal(0, 3, (h, w)) + 1j * rng.normal(0, 3, (h, w))
    k = rng.integers(0, 4)
    for _ in range(k):
        ty, tx = rng.integers(1, h // 16 - 1), rng.integers(1, w // 16 - 2)
        Y[16 * ty:16 * ty + 32, 16 * tx:16 * tx + 48] *= 3
    return Y.astype(F), chi.astype(np.complex64)
