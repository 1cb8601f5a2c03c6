"""The pointwise face-flux VJP through the C-ABI (charm_gpu_flux_vjp_batch)
behind a torch.autograd.Function whose forward is the device flux batch
(charm_gpu_flux_batch): torch.autograd.gradcheck over both sides' primitive
rows (r, u, v, w, p, c) at random unit normals, for LF, HLLC (subsonic
points of either star side and supersonic points of either direction) and
the central flux, one and two species (FP64)."""
import numpy as np
import pytest
import torch

import charm_3d_amd as CA
import mesh_util as MU
from test_gpu_parity import device, random_state

pytestmark = pytest.mark.gpu


class Flux(torch.autograd.Function):
    @staticmethod
    def forward(fctx, pl, pr, dev, kind, normals):
        fctx.dev, fctx.kind, fctx.normals = dev, kind, normals
        fctx.save_for_backward(pl, pr)
        return torch.from_numpy(dev.flux_batch(kind, pl.detach().numpy(), pr.detach().numpy(), normals))

    @staticmethod
    def backward(fctx, g):
        pl, pr = fctx.saved_tensors
        bl, br = fctx.dev.flux_vjp_batch(fctx.kind, pl.detach().numpy(), pr.detach().numpy(), fctx.normals,
                                         g.detach().numpy())
        return torch.from_numpy(bl), torch.from_numpy(br), None, None, None


def rows(rng, n, nc, vel_scale):
    r = 1.0 + 0.3 * rng.uniform(size=n)
    vel = vel_scale * rng.normal(size=(n, 3))
    p = 1.0 + 0.3 * rng.uniform(size=n)
    c = rng.dirichlet(np.ones(nc), size=n) if nc > 1 else np.ones((n, 1))
    return np.column_stack([r, vel, p, c])


@pytest.mark.parametrize("kind, nc", [("HLLC", 1), ("LF", 1), ("CD", 1), ("HLLC", 2)])
def test_flux_vjp_gradcheck(kind, nc):
    rng = np.random.default_rng(11 + nc)
    comp_m = (CA.GAS_R, 2.0 * CA.GAS_R)[:nc]
    comp_cp = (3.5, 2.5)[:nc]
    cfg = CA.make_config(precision=CA.FP64, flux=kind, eos="IDEAL" if nc == 1 else "MIX",
                         comp_m=comp_m, comp_cp=comp_cp)
    leaves = MU.uniform_leaves(1)
    dev = device(cfg, leaves, (False,) * 3, random_state(leaves, rng, nc=nc))
    n = 24
    # subsonic points and supersonic points in both directions along the normal
    pl = np.vstack([rows(rng, n, nc, 0.3), rows(rng, n // 2, nc, 0.1), rows(rng, n // 2, nc, 0.1)])
    pr = np.vstack([rows(rng, n, nc, 0.3), rows(rng, n // 2, nc, 0.1), rows(rng, n // 2, nc, 0.1)])
    nrm = rng.normal(size=(2 * n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    for blk, s in ((slice(n, n + n // 2), 3.0), (slice(n + n // 2, 2 * n), -3.0)):
        pl[blk, 1:4] += s * nrm[blk]
        pr[blk, 1:4] += s * nrm[blk]
    a = torch.tensor(pl, requires_grad=True)
    b = torch.tensor(pr, requires_grad=True)
    # the dual-number Jacobian is exact to rounding; the tolerance is the
    # central difference's: eps_mach |q| / eps with |q| up to ~20 at the
    # supersonic points (1e-8 here), plus its O(eps^2) truncation
    assert torch.autograd.gradcheck(lambda x, y: Flux.apply(x, y, dev, kind, nrm), (a, b), eps=1e-6,
                                    atol=1e-7, rtol=1e-6)
    dev.close()
