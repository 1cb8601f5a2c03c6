"""The stage VJP on its own (charm_gpu_stage_vjp): for the forward-Euler stage
S(u) = u - dt A^-1 int(u) (charm_gpu_step with the Euler model, no limiter),
S_u^T x through the C-ABI, wrapped in a torch.autograd.Function whose
forward is the device stage itself.

* torch.autograd.gradcheck (fast mode) passes on a 2:1 mesh with inlet,
  outlet and wall faces, for HLLC and LF (FP64);
* <S_u^T x, v> equals the central difference of <x, S(u)> along v (random
  x, v), 1e-7 relative;
* the FP32 stage VJP (chunked face-centric kernel) is within 1e-5 of the
  FP64 one (element-centric kernel)."""
import numpy as np
import pytest
import torch

import charm_3d_amd as CA
import mesh_util as MU
from test_gpu_parity import device, random_state

pytestmark = pytest.mark.gpu

BND = ("INLET", "OUTLET", "WALL_SLIP", "OUTLET", "WALL_NO_SLIP", "OUTLET")
PARAMS = [(0.4, 0.1, -0.05, 1.0, 1.1), None, None, None, (1.05,), None]


def setup(flux, precision=CA.FP64):
    leaves = MU.random_adapted(1, 3, seed=31, frac=0.3)
    cfg = CA.make_config(precision=precision, model="NS", flux=flux, limiter=False, bnd=BND, bnd_param=PARAMS,
                         comp_m=(CA.GAS_R,), comp_cp=(3.5,))
    u0 = random_state(leaves, np.random.default_rng(7))
    ctx = device(cfg, leaves, (False,) * 3, u0)
    dt = 0.5 * ctx.get_dt(1)
    return ctx, u0, dt


class EulerStage(torch.autograd.Function):
    """S(u) on the device; backward = charm_gpu_stage_vjp"""

    @staticmethod
    def forward(fctx, u, dev, dt):
        fctx.dev, fctx.dt = dev, dt
        fctx.save_for_backward(u)
        dev.set_state(u.detach().numpy())
        dev.step("EULER", dt, 1)
        return torch.from_numpy(dev.get_state())

    @staticmethod
    def backward(fctx, g):
        (u,) = fctx.saved_tensors
        return torch.from_numpy(fctx.dev.stage_vjp(fctx.dt, u.detach().numpy(), g.detach().numpy())), None, None


@pytest.mark.parametrize("flux", ["HLLC", "LF"])
def test_gradcheck_of_the_stage(flux):
    dev, u0, dt = setup(flux)
    u = torch.tensor(u0, requires_grad=True)
    assert torch.autograd.gradcheck(lambda a: EulerStage.apply(a, dev, dt), (u,), eps=1e-7, atol=1e-9,
                                    rtol=1e-6, fast_mode=True)
    dev.close()


def test_dot_product_against_central_differences():
    dev, u0, dt = setup("HLLC")
    rng = np.random.default_rng(3)
    scale = np.abs(u0).max(axis=0, keepdims=True)
    for _ in range(4):
        x = rng.normal(size=u0.shape)
        v = rng.normal(size=u0.shape) * scale
        ad = float((dev.stage_vjp(dt, u0, x) * v).sum())
        h = 1e-6
        sp = EulerStage.apply(torch.from_numpy(u0 + h * v), dev, dt).numpy()
        sm = EulerStage.apply(torch.from_numpy(u0 - h * v), dev, dt).numpy()
        fd = float((x * (sp - sm)).sum()) / (2 * h)
        print("stage VJP: adjoint %.12e, central difference %.12e" % (ad, fd))
        assert abs(ad - fd) <= 1e-7 * abs(fd)
    dev.close()


def test_fp32_stage_vjp_close_to_fp64():
    out = {}
    x = np.random.default_rng(5).normal(size=(0,))
    for prec in (CA.FP64, CA.FP32):
        dev, u0, dt = setup("HLLC", prec)
        if x.size == 0:
            x = np.random.default_rng(5).normal(size=u0.shape)
        out[prec] = dev.stage_vjp(dt, u0, x)
        dev.close()
    d = np.linalg.norm(out[CA.FP32] - out[CA.FP64]) / np.linalg.norm(out[CA.FP64])
    print("FP32 vs FP64 stage VJP: relative L2 %.2e" % d)
    assert d < 1e-5
