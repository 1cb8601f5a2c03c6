"""The torch.autograd.Function wrapper of the adjoint (tests/torch_adjoint.py)
on the CPU, with a stand-in problem whose objective and gradient are closed
forms (J = sum w u^2 / 2 + c . theta): backward scales both gradients by the
incoming gradient, returns tensors of the inputs' dtype, and passes
torch.autograd.gradcheck."""
import numpy as np
import torch

from torch_adjoint import RolloutObjective


class Quadratic:
    def __init__(self, shape, seed=0):
        rng = np.random.default_rng(seed)
        self.w = rng.uniform(0.5, 2.0, shape)
        self.c = rng.normal(size=3)
        self.calls = 0

    def objective(self, u0, theta):
        return float(0.5 * (self.w * u0 ** 2).sum() + self.c @ theta)

    def gradient(self, u0, theta):
        self.calls += 1
        return self.objective(u0, theta), self.w * u0, self.c.copy()


def test_backward_scales_the_closed_form_gradient():
    prob = Quadratic((6, 5, 4))
    u0 = np.random.default_rng(1).normal(size=(6, 5, 4))
    u = torch.tensor(u0, requires_grad=True)
    th = torch.tensor([0.5, 1.0, 1.0], dtype=torch.float64, requires_grad=True)
    j = RolloutObjective.apply(u, th, prob)
    assert abs(float(j) - prob.objective(u0, th.detach().numpy())) < 1e-12
    (-3.0 * j).backward()
    assert u.grad.dtype == torch.float64 and u.grad.shape == u.shape
    np.testing.assert_allclose(u.grad.numpy(), -3.0 * prob.w * u0, rtol=1e-15)
    np.testing.assert_allclose(th.grad.numpy(), -3.0 * prob.c, rtol=1e-15)
    assert prob.calls == 1


def test_gradcheck_passes_on_the_closed_form():
    prob = Quadratic((3, 5, 4), seed=2)
    u = torch.tensor(np.random.default_rng(3).normal(size=(3, 5, 4)), requires_grad=True)
    th = torch.tensor([0.5, 1.0, 1.0], dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: RolloutObjective.apply(a, b, prob), (u, th), eps=1e-6, atol=1e-8)
