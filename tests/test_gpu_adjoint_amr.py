"""Adjoint-weighted refinement on the GPU: dJ/du0 of the inflow pulse case
(charm_gpu_adjoint) -> charm_amr_adjoint_weights -> charm_adapt_weighted.

* unit weights reproduce charm_adapt leaf for leaf;
* the adjoint weights give the forest that the reference's criteria
  (charm_amr_flags) give on the weighted estimator, followed by the
  forest's own refine / coarsen / 2:1 balance, and it differs from the
  unweighted one;
* the projection onto the adapted mesh keeps the conserved totals."""
import ctypes as C

import numpy as np
import pytest

import charm_3d_amd as CA
import charm_3d_amd.workloads as W

pytestmark = pytest.mark.gpu

MAX_ERR, MIN_L, MAX_L = 0.01, 2, 4
LO, HI = (0.0, 0.0, 0.0), (0.45, 1.0, 1.0)


def solver():
    cfg, lv, u0, per = W.c5_pulse(3, CA.FP64)
    cfg.max_err, cfg.min_level, cfg.max_level = MAX_ERR, MIN_L, MAX_L
    f = CA.Forest.from_leaves(lv, per)
    s = CA.Solver(f, cfg)
    s.set_state(u0)
    print(f"[gpu] adjoint-weighted refinement n={f.n}", flush=True)
    return s, f, u0, per


def indicator(s, n):
    err2, rho_c = np.zeros(n), np.zeros(n)
    CA.check(CA.lib().charm_gpu_amr_indicator(s.gpu, err2.ctypes.data_as(C.POINTER(C.c_double)),
                                              rho_c.ctypes.data_as(C.POINTER(C.c_double)), n))
    return err2, rho_c


def test_unit_weights_equal_plain_adapt():
    s1, f1, _, _ = solver()
    n0 = f1.n
    s1.adapt()
    s2, f2, _, _ = solver()
    s2.adapt_weighted(np.ones(f2.n))
    lv = np.asarray(f1.leaves())[:, 3]
    print("levels after adapt:", np.bincount(lv), "from", n0, "leaves at level 3")
    assert (lv == MAX_L).any() and (lv == MIN_L).any()          # refined near the pulse, coarsened elsewhere
    assert np.array_equal(f1.leaves(), f2.leaves())
    np.testing.assert_array_equal(s1.get_state(), s2.get_state())
    s1.close()
    s2.close()


def test_adjoint_weights_drive_the_refinement():
    s, f, u0, per = solver()
    _, g, _ = s.adjoint(0.01, 16, LO, HI, ckpt=4)
    s.set_state(u0)
    leaves0 = np.asarray(f.leaves())
    w = CA.amr_adjoint_weights(leaves0, g)
    assert w.std() > 0.1
    err2, rho_c = indicator(s, f.n)
    flags = CA.amr_flags(leaves0, err2 * w, rho_c, MAX_ERR, MIN_L, MAX_L)
    ref = CA.Forest.from_leaves(leaves0, per)
    ref.adapt(flags, MAX_L, MIN_L)
    m0 = s.conservation_sums()
    s.adapt_weighted(w)
    assert np.array_equal(np.asarray(f.leaves()), np.asarray(ref.leaves()))
    m1 = s.conservation_sums()
    np.testing.assert_allclose(m1, m0, rtol=1e-12, atol=1e-14)
    plain, fp, _, _ = solver()
    plain.adapt()
    assert not np.array_equal(np.asarray(fp.leaves()), np.asarray(f.leaves()))
    print("leaves: initial %d, unweighted %d, adjoint-weighted %d" % (len(leaves0), fp.n, f.n))
    s.close()
    plain.close()
