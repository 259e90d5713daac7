"""The oracle's 3-D solve (VariationalRefine with temporal neighbours,
variational.cpp:198-556 with :255-281, :337-346, :385-397, :452-517),
pinned against hvd-core's own VariationalRefine built from its sources
(oracle/_ref) and against identities that reduce it to the 2-D solve."""
import numpy as np
import pytest

from conftest import TOL_IRE
from temporal_cases import temporal_scene
import oracle as O

CFG = dict(cg_iterations=40, irls_outer=4, fast=0, chroma_aniso=0.7)


@pytest.mark.parametrize("nu,eps_t", [(0.5, 2.0), (1.5, 0.7)])
def test_neighbours_match_hvd_core(orc64, orc32, ref, nu, eps_t):
    s, car, chi0, nbrs = temporal_scene(24, 40, 1)
    cfg = O.Cfg.make(**CFG)
    y_r, c_r = ref.variational_refine_nbr(s, car, chi0, cfg, nbrs, nu, eps_t)
    y3, c3, _, _ = orc32.variational_refine_nbr(s, car, chi0, cfg, nbrs, nu, eps_t)
    # the FP32 restatement performs hvd-core's float operations in its order
    assert np.array_equal(y3.astype(np.float32), y_r) and np.array_equal(c3.astype(np.complex64), c_r)
    y6, c6, _, _ = orc64.variational_refine_nbr(s, car, chi0, cfg, nbrs, nu, eps_t)
    assert np.abs(y6 - y_r).max() <= TOL_IRE and np.abs(c6 - c_r).max() <= TOL_IRE
    # and the neighbours do change the answer
    _, c2, _, _ = orc64.variational_refine(s, car, chi0, cfg)
    assert np.abs(c6 - c2).max() > 1.0


def test_neighbour_identities(orc64):
    # zero confidence or zero strength: exactly the 2-D solve
    s, car, chi0, nbrs = temporal_scene(20, 33, 2, n_nbr=3)
    cfg = O.Cfg.make(**CFG)
    y2, c2, _, _ = orc64.variational_refine(s, car, chi0, cfg)
    off = [(a, b, np.zeros_like(cf)) for a, b, cf in nbrs]
    y0, c0, _, _ = orc64.variational_refine_nbr(s, car, chi0, cfg, off, 1.0, 1.0)
    assert np.array_equal(y0, y2) and np.array_equal(c0, c2)
    y0, c0, _, _ = orc64.variational_refine_nbr(s, car, chi0, cfg, nbrs, 0.0, 1.0)
    assert np.array_equal(y0, y2) and np.array_equal(c0, c2)


def test_neighbour_that_agrees_is_invisible(orc64):
    # a neighbour that sees exactly the same composite through the same
    # carrier has rt = Re[chi * 0] = 0 and dc = 0: it adds nothing
    s, car, chi0, _ = temporal_scene(16, 25, 3)
    cfg = O.Cfg.make(**CFG)
    y2, c2, _, _ = orc64.variational_refine(s, car, chi0, cfg)
    same = [(s, car, np.ones_like(s))]
    y1, c1, _, _ = orc64.variational_refine_nbr(s, car, chi0, cfg, same, 2.0, 1.0)
    assert np.array_equal(c1, c2) and np.array_equal(y1, y2)
