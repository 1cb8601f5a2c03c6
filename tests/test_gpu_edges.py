"""Edge cases of the mesh and of the chunking, FP64 device against the
oracle (1e-12 relative) through the C-ABI:

* a single leaf (level 0) with six boundary faces, and a single leaf that is
  its own neighbour across every periodic axis (each face joins the leaf to
  itself: both sides are gathered by the same leaf);
* one periodic axis on a two-leaf-wide mesh;
* the smallest chunks the flattening allows (one leaf per workgroup) and
  chunks at the LDS limit give the default chunking's residual;
* ragged adapted meshes (levels 1-4, 2:1 balanced) with the smallest chunks;
* a uniform state is preserved on the single-leaf periodic mesh."""
import numpy as np
import pytest

import charm_3d_amd as CA
import mesh_util as MU
import oracle_lib as OL
from test_gpu_parity import REL64, oracle_cfg, random_state, relerr

pytestmark = pytest.mark.gpu


def residual(cfg, leaves, periodic, u0, chunks=None):
    print(f"[gpu] n={len(leaves)} periodic={periodic} chunks={chunks}", flush=True)
    ctx = CA.GpuContext(cfg)
    ce, cf = chunks if chunks else CA.recommended_chunks(cfg)
    ctx.upload_mesh(CA.Forest.from_leaves(leaves, periodic).flatten(ce, cf))
    ctx.set_state(u0)
    out = ctx.residual()
    ctx.close()
    return out


def oracle_residual(cfg, leaves, u0):
    sim = OL.OracleSim(oracle_cfg(cfg), leaves)
    sim.set_state(u0)
    return sim.residual()


CASES = {
    "single_leaf_boundary": (0, (False,) * 3, ("OUTLET", "WALL_SLIP", "INLET", "OUTLET", "WALL_NO_SLIP", "OUTLET")),
    "single_leaf_periodic": (0, (True,) * 3, ("PERIODIC",) * 6),
    "single_leaf_periodic_x": (0, (True, False, False), ("PERIODIC", "PERIODIC", "OUTLET", "WALL_SLIP", "OUTLET", "OUTLET")),
    "two_wide_periodic_y": (1, (False, True, False), ("OUTLET", "WALL_SLIP", "PERIODIC", "PERIODIC", "INLET", "OUTLET")),
}


def bnd_params(bnd):
    """inlet (u, v, w, T, p) and no-slip wall T for the sides that have them"""
    return [(0.1, -0.2, 0.05, 1.1, 1.2) if b == "INLET" else ((0.9,) if b == "WALL_NO_SLIP" else None)
            for b in bnd]


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("flux", ["LF", "HLLC"])
def test_small_meshes_match_oracle(name, flux):
    level, periodic, bnd = CASES[name]
    cfg = CA.make_config(precision=CA.FP64, flux=flux, bnd=bnd, bnd_param=bnd_params(bnd), grav=(0.1, -0.3, 0.2))
    leaves = MU.uniform_leaves(level)
    u0 = random_state(leaves, np.random.default_rng(17))
    got = residual(cfg, leaves, periodic, u0)
    assert relerr(got, oracle_residual(cfg, leaves, u0)) < REL64


def test_single_leaf_periodic_preserves_a_uniform_state():
    cfg = CA.make_config(precision=CA.FP64, flux="HLLC", bnd=("PERIODIC",) * 6)
    leaves = MU.uniform_leaves(0)
    rho, vel, p = 1.2, np.array([0.3, -0.1, 0.2]), 0.9
    u0 = np.zeros((1, 5, 4))
    u0[0, :, 0] = list(rho * vel) + [p / 0.4 + 0.5 * rho * vel @ vel, rho]      # gamma 1.4
    got = residual(cfg, leaves, (True,) * 3, u0)
    assert np.abs(got).max() < 1e-14


@pytest.mark.parametrize("mesh", ["uniform3", "adapt", "adapt_p"])
def test_chunk_sizes_do_not_change_the_residual(mesh):
    """one leaf per workgroup, chunks at the LDS limit and the default give
    the oracle's residual (the gather order per leaf is the slot order in
    every case, so FP64 results agree to rounding)"""
    gens = {"uniform3": (lambda: MU.uniform_leaves(3), (False,) * 3),
            "adapt": (lambda: MU.random_adapted(1, 4, seed=23, frac=0.3), (False,) * 3),
            "adapt_p": (lambda: MU.random_adapted(2, 4, seed=29, frac=0.2, periodic=(True,) * 3), (True,) * 3)}
    gen, periodic = gens[mesh]
    leaves = gen()
    bnd = ("PERIODIC",) * 6 if periodic[0] else ("OUTLET", "WALL_SLIP", "OUTLET", "INLET", "WALL_NO_SLIP", "OUTLET")
    cfg = CA.make_config(precision=CA.FP64, flux="HLLC", bnd=bnd, bnd_param=bnd_params(bnd))
    u0 = random_state(leaves, np.random.default_rng(19))
    ref = oracle_residual(cfg, leaves, u0)
    ce, cf = CA.recommended_chunks(cfg)
    for chunks in ((1, 24), (ce, cf), (256, cf)):
        got = residual(cfg, leaves, periodic, u0, chunks)
        assert relerr(got, ref) < REL64, chunks
