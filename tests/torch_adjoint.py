"""A torch.autograd.Function around the GPU discrete adjoint (test side).

J(u0, theta) = sum_n dt * int_box p over an SSP-RK3 rollout of the inflow
pulse case (Euler P1, HLLC, inlet on the -x face), theta = the inlet's
(u, T, p).  forward runs the device objective (charm_gpu_objective);
backward runs charm_gpu_adjoint, which repeats the forward pass with
checkpoints, sweeps the hand-written VJP kernels back through the steps
and returns dJ/du0 and dJ/dtheta; both are scaled by the incoming
gradient.  The state crosses the C-ABI as host FP64 arrays, so tensors of
any device are accepted and the gradients come back on the input's
device."""
import numpy as np
import torch

import charm_3d_amd as CA
import charm_3d_amd.workloads as W


class PulseRollout:
    """The fixed part of the problem: leaves, arithmetic, dt, number of
    steps, the objective box and the checkpoint period."""

    def __init__(self, leaves, precision=CA.FP64, dt=0.01, steps=16,
                 lo=(0.0, 0.0, 0.0), hi=(0.45, 1.0, 1.0), ckpt=5):
        self.leaves = leaves
        self.precision = precision
        self.dt, self.steps, self.ckpt = dt, steps, ckpt
        self.lo, self.hi = lo, hi

    def device(self, u0, theta):
        cfg, lv, _, per = W.c5_pulse(precision=self.precision, inlet=tuple(float(t) for t in theta),
                                     leaves=self.leaves)
        ctx = CA.GpuContext(cfg)
        ce, cf = CA.recommended_chunks(cfg)
        ctx.upload_mesh(CA.Forest.from_leaves(lv, per).flatten(ce, cf))
        ctx.set_state(u0)
        return ctx

    def objective(self, u0, theta):
        ctx = self.device(u0, theta)
        try:
            return ctx.objective(self.dt, self.steps, self.lo, self.hi)
        finally:
            ctx.close()

    def gradient(self, u0, theta):
        """(J, dJ/du0, dJ/dtheta) from the adjoint entry point"""
        ctx = self.device(u0, theta)
        try:
            j, g, th, _, _ = ctx.adjoint(self.dt, self.steps, self.lo, self.hi, ckpt=self.ckpt, theta_face=0)
        finally:
            ctx.close()
        return j, g, th


def _host(t):
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


class RolloutObjective(torch.autograd.Function):
    """J = RolloutObjective.apply(u0, theta, problem); u0 is [n][V][4]
    (Taylor coefficients per leaf and variable), theta the inlet (u, T, p)."""

    @staticmethod
    def forward(ctx, u0, theta, problem):
        ctx.problem = problem
        ctx.save_for_backward(u0, theta)
        return u0.new_tensor(problem.objective(_host(u0), _host(theta)))

    @staticmethod
    def backward(ctx, grad_j):
        u0, theta = ctx.saved_tensors
        _, g, th = ctx.problem.gradient(_host(u0), _host(theta))
        gu = torch.from_numpy(g).to(u0) * grad_j
        gt = torch.from_numpy(th).to(theta) * grad_j
        return gu, gt, None
