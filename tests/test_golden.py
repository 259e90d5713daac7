"""The oracle against golden vectors written by hvd-core's own engine code
(tests/golden/make_golden.py, compiled from the reference checkout).  These
pin the oracle on machines without the reference."""
import pathlib

import numpy as np
import pytest

from conftest import TOL_IRE

import oracle as O
from hvd_amd import synth

GOLD = np.load(pathlib.Path(__file__).resolve().parent / "golden" / "hvdcore_vectors.npz")


def burst_field(g, burst_raw):
    f = np.zeros((g.field_height, g.field_width), np.float32)
    f[:, g.colour_burst_start:g.colour_burst_end] = synth.ire_from_raw(burst_raw)
    return f


def test_golden_lockin(orc32, orc64):
    f = burst_field(synth.NTSC, GOLD["ntsc_burst_raw"])
    assert np.array_equal(orc32.burst_lockin_phase(f, synth.NTSC), GOLD["ntsc_theta"])
    assert np.abs(orc64.burst_lockin_phase(f, synth.NTSC) - GOLD["ntsc_theta"]).max() < 2e-4
    assert orc32.burst_amplitude_ire(f, synth.NTSC) == GOLD["ntsc_burst_amp"]
    f = burst_field(synth.PAL, GOLD["pal_burst_raw"])
    th, vs = orc32.burst_lockin_phase_pal(f, synth.PAL)
    assert np.array_equal(th, GOLD["pal_theta"]) and np.array_equal(vs, GOLD["pal_vswitch"])
    th, vs = orc64.burst_lockin_phase_pal(f, synth.PAL)
    assert np.array_equal(vs, GOLD["pal_vswitch"]) and np.abs(th - GOLD["pal_theta"]).max() < 3e-4
    assert orc32.burst_amplitude_ire(f, synth.PAL) == GOLD["pal_burst_amp"]


def test_golden_refine(orc32, orc64):
    cfg = O.Cfg.make(cg_iterations=40, irls_outer=4, fast=0, cg_tol=1e-12)
    s, car, chi0 = GOLD["refine_s"], GOLD["refine_carrier"], GOLD["refine_chi0"]
    y32, c32, _, _ = orc32.variational_refine(s, car, chi0, cfg)
    assert np.array_equal(y32, GOLD["refine_y"])
    assert np.array_equal(c32.astype(np.complex64), GOLD["refine_chi"])
    y64, c64, _, _ = orc64.variational_refine(s, car, chi0, cfg)
    assert np.abs(y64 - GOLD["refine_y"]).max() <= TOL_IRE
    assert np.abs(c64 - GOLD["refine_chi"]).max() <= TOL_IRE


def test_golden_motion(orc32):
    got = orc32.estimate_motion(GOLD["motion_ref"], GOLD["motion_cur"], 32, 8)
    for a, k in zip(got, ("motion_dy", "motion_dx", "motion_conf")):
        assert np.array_equal(a, GOLD[k])
