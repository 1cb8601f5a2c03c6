"""The configurations at their full sizes, through properties that do not
depend on an oracle run of that size (FP64 and FP32 device paths through
the C-ABI):

* C1 in full (level 4, 4,096 hexes, 100 steps, LF, periodic) against the
  oracle, which finishes it in seconds: relative 1e-11;
* Riemann box at level 7 (2,097,152 hexes):
  - free-stream preservation: a uniform state has a zero residual (to
    rounding of the face and volume terms, relative to F h^2);
  - the FP32 residual against the FP64 one (2e-5, as at small sizes);
  - three steps run twice give the same bits (atomic-free gather);
  - chunks of one leaf, the default chunks and chunks at the LDS limit give
    the same FP64 residual (1e-12);
* sine wave at level 7, periodic: the sum of the cell-mean residuals
  telescopes to zero (conservation) up to the rounding of each leaf's sum."""
import numpy as np
import pytest

import charm_3d_amd as CA
import charm_3d_amd.workloads as W
import oracle_lib as OL
from test_gpu_parity import oracle_cfg, relerr

pytestmark = pytest.mark.gpu


def context(cfg, level, periodic, u0, chunks=None):
    print(f"[gpu] full size level={level} prec={cfg.precision} chunks={chunks}", flush=True)
    ctx = CA.GpuContext(cfg)
    ce, cf = chunks if chunks else CA.recommended_chunks(cfg)
    ctx.upload_mesh(CA.Forest.uniform(level, periodic).flatten(ce, cf))
    ctx.set_state(u0)
    return ctx


def test_c1_in_full_matches_oracle():
    cfg, leaves, u0, periodic = W.c1(level=4, precision=CA.FP64)
    ctx = context(cfg, 4, periodic, u0)
    dt = ctx.get_dt(mode=1)
    ctx.run(dt, 100)
    sim = OL.OracleSim(oracle_cfg(cfg), leaves)
    sim.set_state(u0)
    sim.run(dt, 100)
    err = relerr(ctx.get_state(), sim.get_state())
    print("C1, 100 steps at level 4: device vs oracle %.2e" % err)
    assert err < 1e-11
    ctx.close()


@pytest.mark.parametrize("prec, tol", [(CA.FP64, 1e-13), (CA.FP32, 2e-6)])
def test_free_stream_preserved_at_level7(prec, tol):
    cfg, leaves, _, periodic = W.c2(level=7, precision=prec)
    n = len(leaves)
    u0 = np.zeros((n, 5, 4))
    vel = np.array([0.3, -0.2, 0.1])
    u0[:, :, 0] = W.prim_to_cons(np.ones(1), vel[None, :], np.ones(1))[0]
    ctx = context(cfg, 7, periodic, u0)
    r = ctx.residual()
    h = 1.0 / 128
    fscale = 1.0 + vel @ vel                           # |F| ~ p + rho |u|^2
    print("free stream, level 7: max |r| / (F h^2) = %.2e" % (np.abs(r).max() / (fscale * h * h)))
    assert np.abs(r).max() <= tol * fscale * h * h
    ctx.close()


def test_fp32_residual_close_to_fp64_at_level7():
    out = {}
    for prec in (CA.FP64, CA.FP32):
        cfg, leaves, u0, periodic = W.c2(level=7, precision=prec)
        ctx = context(cfg, 7, periodic, u0)
        out[prec] = ctx.residual()
        ctx.close()
    d = np.linalg.norm(out[CA.FP32] - out[CA.FP64]) / np.linalg.norm(out[CA.FP64])
    print("Riemann box level 7: FP32 vs FP64 residual, relative L2 %.2e" % d)
    assert d < 2e-5


def test_steps_are_deterministic_at_level7():
    cfg, leaves, u0, periodic = W.c2(level=7, precision=CA.FP32)
    states = []
    for _ in range(2):
        ctx = context(cfg, 7, periodic, u0)
        dt = ctx.get_dt(mode=1)
        ctx.run(dt, 3)
        states.append(ctx.get_state())
        ctx.close()
    assert np.array_equal(states[0], states[1])


def test_chunking_does_not_change_the_residual_at_level7():
    cfg, leaves, u0, periodic = W.c2(level=7, precision=CA.FP64)
    ce, cf = CA.recommended_chunks(cfg)
    ref = None
    for chunks in ((ce, cf), (1, 24), (256, cf)):
        ctx = context(cfg, 7, periodic, u0, chunks)
        r = ctx.residual()
        ctx.close()
        if ref is None:
            ref = r
        else:
            assert relerr(r, ref) < 1e-12, chunks


# Each face moment enters its two leaves with opposite signs, so the sum
# telescopes; what is left is the rounding of each leaf's sum of six
# moments, at most about 6 eps |M| per leaf, where |M| / |r| ~ 1 / (k h) ~ 20
# for this wave at level 7: 7e-15 (FP64) and 7e-6 (FP32) of sum |r|.
@pytest.mark.parametrize("prec, tol", [(CA.FP64, 1e-13), (CA.FP32, 1e-5)])
def test_periodic_mean_residuals_telescope_at_level7(prec, tol):
    cfg, leaves, u0, periodic = W.c1(level=7, precision=prec)
    ctx = context(cfg, 7, periodic, u0)
    r = ctx.residual()
    ctx.close()
    for v in range(5):
        s, a = r[:, v, 0].sum(), np.abs(r[:, v, 0]).sum()
        print("var %d: sum of mean residuals / sum of |mean residual| = %.2e" % (v, abs(s) / a))
        assert abs(s) <= tol * a
