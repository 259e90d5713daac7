"""The condition under which the along-x box-blur pass adds a row in
parallel (motion.hip, k_me_blur_rows): every sample is a multiple of 2^e
(e the exponent of the smallest ulp among the row's nonzero floats) and
w max|a| < 2^(53 + e).  Checked here against exact rational arithmetic: on
rows that meet it, hvd-core's left-to-right FP64 running sums (motion.cpp:
38-55) never round, so any order of the adds gives the same bits; rows that
fail it include ones whose sums do round."""
from fractions import Fraction

import numpy as np


def certificate(row):
    nz = row[row != 0]
    if nz.size == 0:
        return True
    bits = nz.view(np.uint32)
    emin = int((np.maximum((bits >> 23) & 255, 1).astype(np.int64) - 150).min())
    return float(len(row)) * float(np.abs(row).max()) < 2.0 ** (53 + emin)


def left_to_right(row):
    c = [0.0]
    for v in row:
        c.append(c[-1] + float(v))
    return c


def exact(row):
    c = [Fraction(0)]
    for v in row:
        c.append(c[-1] + Fraction(float(v)))
    return c


def test_certified_rows_sum_exactly():
    rng = np.random.default_rng(5)
    checked = 0
    for k in range(300):
        w = int(rng.integers(1, 900))
        scale = 10.0 ** rng.uniform(-8, 6)
        row = (rng.normal(0, 1, w) * scale).astype(np.float32)
        if k % 3 == 0:
            row[rng.integers(0, w, max(1, w // 10))] = 0.0
        if certificate(row):
            checked += 1
            assert [Fraction(x) for x in left_to_right(row)] == exact(row)
    assert checked > 200


def test_failing_rows_include_rounding_ones():
    # the condition is not vacuous: rows it rejects can round left to right
    row = np.array([1e30, 1.0, -1e30, 3.0], np.float32)
    assert not certificate(row)
    assert [Fraction(x) for x in left_to_right(row)] != exact(row)
    tiny = np.full(600, 100.0, np.float32)
    tiny[::5] = 1e-9
    assert not certificate(tiny)


def test_certificate_planted_mistake_is_caught():
    # a certificate one bit too generous (54 instead of 53) would admit a row
    # whose left-to-right sums round: 2^29 + 2^29 + (1 + 2^-23) needs 54 bits
    row = np.array([2.0 ** 29, 2.0 ** 29, 1.0 + 2.0 ** -23], np.float32)
    bits = row[row != 0].view(np.uint32)
    emin = int((np.maximum((bits >> 23) & 255, 1).astype(np.int64) - 150).min())
    loose = float(len(row)) * float(np.abs(row).max()) < 2.0 ** (54 + emin)
    rounds = [Fraction(x) for x in left_to_right(row)] != exact(row)
    assert not certificate(row)
    assert loose and rounds
