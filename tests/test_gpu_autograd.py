"""The adjoint VJP kernels behind a torch.autograd.Function
(tests/torch_adjoint.py), on the inflow pulse case (FP64 device path):

* backward reproduces the adjoint entry point's gradients, scaled by the
  incoming gradient;
* torch.autograd.gradcheck (fast mode: random-direction products) passes
  for both inputs;
* 32 random directions of the initial state with h = 1e-6 agree with central
  differences of the forward objective to 1e-5 relative."""
import numpy as np
import pytest
import torch

import charm_3d_amd as CA
import charm_3d_amd.workloads as W
import mesh_util as MU
from torch_adjoint import PulseRollout, RolloutObjective

pytestmark = pytest.mark.gpu

THETA = (0.5, 1.0, 1.0)


def problem(adapted):
    lv = MU.random_adapted(2, 3, seed=8, frac=0.3) if adapted else CA.Forest.uniform(3).leaves()
    _, lv, u0, _ = W.c5_pulse(precision=CA.FP64, inlet=THETA, leaves=lv)
    print(f"[gpu] autograd pulse n={len(lv)} adapted={adapted}", flush=True)
    return PulseRollout(lv), u0


@pytest.mark.parametrize("adapted", [False, True])
def test_backward_returns_the_adjoint_gradients(adapted):
    prob, u0 = problem(adapted)
    u = torch.tensor(u0, dtype=torch.float64, requires_grad=True)
    th = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    j = RolloutObjective.apply(u, th, prob)
    (2.5 * j).backward()
    j_ref, g, gth = prob.gradient(u0, np.array(THETA))
    assert abs(float(j.detach()) - j_ref) <= 1e-14 * abs(j_ref)
    assert torch.equal(u.grad, 2.5 * torch.from_numpy(g))
    assert torch.equal(th.grad, 2.5 * torch.from_numpy(gth))


def test_gradcheck_fast_mode():
    prob, u0 = problem(True)
    u = torch.tensor(u0, dtype=torch.float64, requires_grad=True)
    th = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: RolloutObjective.apply(a, b, prob), (u, th),
                                    eps=1e-6, atol=1e-10, rtol=1e-5, fast_mode=True)


def test_32_random_directions_match_central_differences():
    prob, u0 = problem(False)
    u = torch.tensor(u0, dtype=torch.float64, requires_grad=True)
    th = torch.tensor(THETA, dtype=torch.float64)
    RolloutObjective.apply(u, th, prob).backward()
    g = u.grad.numpy()
    rng = np.random.default_rng(32)
    scale = np.abs(u0).max(axis=0, keepdims=True)
    worst = 0.0
    for _ in range(32):
        v = rng.normal(size=u0.shape) * scale
        h = 1e-6
        with torch.no_grad():
            jp = RolloutObjective.apply(torch.from_numpy(u0 + h * v), th, prob)
            jm = RolloutObjective.apply(torch.from_numpy(u0 - h * v), th, prob)
        fd = float(jp - jm) / (2 * h)
        ad = float((g * v).sum())
        worst = max(worst, abs(ad - fd) / abs(fd))
    print("32 directions: worst relative difference %.2e" % worst)
    assert worst <= 1e-5
