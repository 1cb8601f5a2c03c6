"""Adjoint weights for goal-oriented refinement (charm_amr_adjoint_weights,
no counterpart in the reference): w_i = lambda_i / mean(lambda), lambda_i =
sum_v |dJ/dU_{i,v,0}| / V_i.  Properties on the CPU: mean 1, invariance to
the scale and sign of the gradient, blindness to the slope coefficients,
the 1 / V scaling between levels, and the error for a zero gradient."""
import numpy as np
import pytest

import charm_3d_amd as CA
import mesh_util as MU


def test_weight_properties():
    rng = np.random.default_rng(4)
    leaves = np.asarray(MU.random_adapted(1, 3, seed=6, frac=0.4))
    n = len(leaves)
    g = rng.normal(size=(n, 5, 4))
    w = CA.amr_adjoint_weights(leaves, g)
    assert w.shape == (n,) and (w >= 0).all() and abs(w.mean() - 1) < 1e-14
    np.testing.assert_allclose(CA.amr_adjoint_weights(leaves, -3.7 * g), w, rtol=1e-14)
    g2 = g.copy()
    g2[:, :, 1:] = rng.normal(size=(n, 5, 3))
    np.testing.assert_array_equal(CA.amr_adjoint_weights(leaves, g2), w)
    # the same mean-coefficient gradient on a leaf one level finer: 8 x the weight
    i, j = int(np.argmin(leaves[:, 3])), int(np.argmax(leaves[:, 3]))
    assert leaves[j, 3] == leaves[i, 3] + 1 or leaves[j, 3] > leaves[i, 3]
    g3 = np.zeros_like(g)
    g3[i, :, 0] = 1.0
    g3[j, :, 0] = 1.0
    w3 = CA.amr_adjoint_weights(leaves, g3)
    assert abs(w3[j] / w3[i] - 8.0 ** (leaves[j, 3] - leaves[i, 3])) < 1e-12
    with pytest.raises(CA.CharmError, match="zero"):
        CA.amr_adjoint_weights(leaves, np.zeros_like(g))
