"""Shared cases for the solve's adjoint tests: crops of synthetic NTSC fields
with their generator ground truth, the oracle's carrier and FP64 holographic
init, and the training loss (MSE of Y, U, V against the truth)."""
import numpy as np

from hvd_amd import synth

# forced aniso: the AUTO aniso is an order statistic of chi0 whose (continuous)
# dependence on chi0 the adjoint does not differentiate
TRAIN_CFG = dict(irls_outer=4, fast=0, chroma_aniso=0.7)


def crops(orc64, n=16, h=64, w=128, seed0=300):
    """n (s, carrier, chi0, y_true, chi_true) crops of h x w from NTSC fields
    (four crops per field)."""
    out = []
    g = synth.NTSC
    per = 4
    for f in range((n + per - 1) // per):
        raw, sc, _ = synth.ntsc_field(seed0 + f)
        field = synth.ire_from_raw(raw)
        s, car, _, _ = orc64.prepare_field(field, g)
        _, chi0 = orc64.holographic_init(s, car, g.sample_rate)
        rng = np.random.default_rng(seed0 + 1000 + f)
        for _ in range(per):
            if len(out) == n:
                break
            r0 = int(rng.integers(0, g.lines - h))
            c0 = int(rng.integers(0, g.width - w))
            sl = (slice(r0, r0 + h), slice(c0, c0 + w))
            out.append(dict(s=np.ascontiguousarray(s[sl]), car=np.ascontiguousarray(car[sl]),
                            chi0=np.ascontiguousarray(chi0[sl]), yt=sc.y[sl], chit=sc.chi[sl]))
    return out


def loss(y, chi, yt, chit):
    """MSE of Y and of (U, V) = (-Im chi, Re chi) against the truth."""
    n = y.size
    return float(np.sum((y - yt) ** 2) / n + np.sum(np.abs(chi - chit) ** 2) / n)


def loss_grads(y, chi, yt, chit):
    """dL/dY and dL/dchi (as a complex array of real-pair derivatives)."""
    n = y.size
    return 2.0 * (y - yt) / n, 2.0 * (chi - chit) / n
