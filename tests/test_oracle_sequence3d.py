"""Pins of the oracle's window driver (oracle/sequence3d.py) against
hvd-core's own DecodeFieldWindowWithInits compiled from source
(oracle/_ref): the FP32 restatement must give hvd-core's Y, chi and
decisions (resolved strength, ambiguity, sigma, gates) bit for bit, through
the 2-D path, fixed and adaptive 3-D strength, the anchored second pass
(SynthReference + VariationalRefineJoint), fast mode's predicted+verified
motion and trajectory snap, slow Gauss-Seidel mode, the PSI init, moving
content and the PAL offsets.  hvd-core's own sequence_test properties
(3-D beats 2-D on static ambiguous content, the lossless split) are checked
on the oracle too."""
import numpy as np
import pytest

import oracle as O
import sequence3d as S3
from sequence_cases import window

CASES = {
    "2d": dict(cg_iterations=40, temporal_strength=-1.0),
    "3d_fixed": dict(cg_iterations=40, temporal_strength=2.0),
    "auto": dict(cg_iterations=40),
    "slow": dict(cg_iterations=40, temporal_strength=2.0, fast=False),
    "psi": dict(cg_iterations=40, temporal_strength=1.0, psi_init=True),
    "sequential": dict(cg_iterations=40, temporal_strength=1.0, parallel_across_fields=False, output_fidelity=False),
    "defaults": dict(),
    "one_pass_no_snap": dict(cg_iterations=30, temporal_strength=0.8, passes=1, trajectory_fit=False),
}


@pytest.fixture(scope="module")
def orc32():
    return O.Oracle(fp32=True)


def _compare(orc32, ref, fields, parities, inits, cfg, pal=False):
    d = {}
    oy, oc = S3.decode_window_with_inits(orc32, fields, parities, inits, cfg, is_pal=pal, diag=d)
    ry, rc, rd = ref.decode_window_with_inits(fields, parities, inits, cfg.seq(pal), g=_pal_geom() if pal else None)
    for a, b in zip(oy, ry):
        assert np.array_equal(a, b)
    for a, b in zip(oc, rc):
        assert np.array_equal(a, b)
    for k in rd:
        assert np.float32(d[k]) == np.float32(rd[k]), k
    return oy, oc, d


def _pal_geom():
    from hvd_amd import synth
    return synth.PAL


@pytest.mark.parametrize("name", sorted(CASES))
def test_window_static_bit_exact(orc32, ref, name):
    fields, parities, inits, _ = window()
    _compare(orc32, ref, fields, parities, inits, S3.WindowCfg(**CASES[name]))


@pytest.mark.parametrize("motion", [(0, 2), (1, 3)])
def test_window_moving_bit_exact(orc32, ref, motion):
    fields, parities, inits, _ = window(lines=64, width=160, motion=motion, seed=11)
    _compare(orc32, ref, fields, parities, inits, S3.WindowCfg(cg_iterations=30, temporal_strength=1.0))


def test_window_pal_bit_exact(orc32, ref):
    fields, parities, inits, _ = window(lines=48, width=128, nf=10, pal=True, seed=13)
    _compare(orc32, ref, fields, parities, inits, S3.WindowCfg(cg_iterations=30, temporal_strength=1.0), pal=True)


def test_window_clean_scene_backs_off(orc32, ref):
    # no luma energy at the subcarrier: the adaptive scan resolves a weak strength
    fields, parities, inits, _ = window(ambiguous=False, seed=9)
    _, _, d = _compare(orc32, ref, fields, parities, inits, S3.WindowCfg(cg_iterations=40))
    assert d["resolved_strength"] < 0.3


def _psnr(chis, truth):
    m = np.mean([np.mean(np.abs(c - t) ** 2) for c, t in zip(chis, truth)])
    return 10 * np.log10(1e4 / m)


def test_oracle_3d_beats_2d_and_is_lossless(orc32):
    # sequence_test.cpp:139-160 on the oracle
    fields, parities, inits, truth = window()
    y2, c2 = S3.decode_window_with_inits(orc32, fields, parities, inits, S3.WindowCfg(cg_iterations=40,
                                                                                       temporal_strength=-1.0))
    y3, c3 = S3.decode_window_with_inits(orc32, fields, parities, inits, S3.WindowCfg(cg_iterations=40,
                                                                                       temporal_strength=2.0))
    assert _psnr(c3, truth) > _psnr(c2, truth)
    for (s, car), y, c in zip(fields, y3, c3):
        assert np.abs(s - (y + (c * car).real)).max() < 1e-3
