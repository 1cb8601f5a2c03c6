"""GPU parity of the multi-tree (general hexahedral) path against the oracle's
multi-tree restatement, FP64 to 1e-12 (residuals) and 1e-11 (steps), FP32
against FP64 with the bound written in each test:

* a distorted box of 2 x 2 x 2 trees, each relabelled by a random cube
  rotation, at level 1, with inlet, outlet, both walls and gravity, every
  flux; residual, forward-Euler steps with BJ and SSP-RK3 steps;
* two species and mass-flow boundaries on the same mesh;
* the reference's sample (samples/klin_amr: a Mach-2 wedge channel of 800
  trees, HLLC + BJ, Euler model; tests/golden/klin_mesh.npz) at its own
  settings for 5 steps, and at level 1;
* the octree-only entry points refuse a multi-tree context.
"""
import os

import numpy as np
import pytest

import charm_3d_amd as CA
import charm_3d_amd.workloads as W
import gmesh_util as GU
import oracle_lib as OL
from test_gpu_parity import oracle_cfg, relerr

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KLIN = os.path.join(HERE, "golden", "klin_mesh.npz")
BND6 = ("OUTLET", "INLET", "WALL_SLIP", "OUTLET", "WALL_NO_SLIP", "OUTLET")
PAR6 = [None, (-0.2, 0.1, 0.05, 1.1, 1.0), None, None, (0.9,), None]


def var_err(a, b):
    """max over variables of max|a - b| / max|b|"""
    return max(np.abs(a[:, v] - b[:, v]).max() / max(np.abs(b[:, v]).max(), 1e-300) for v in range(a.shape[1]))


def phys_err(a, b):
    """as var_err with one scale for the three momenta: in a flow along x the
    y and z momenta are rounding noise, and a per-variable scale would make
    1e-15 input changes read as O(1) (the oracle itself moves by 1e-15 of
    this scale under such a change in 5 steps of the wedge case)"""
    sc = np.abs(b).max(axis=(0, 2))
    sc[0:3] = sc[0:3].max()
    return max(np.abs(a[:, v] - b[:, v]).max() / max(sc[v], 1e-300) for v in range(a.shape[1]))


def gdevice(cfg, v, t2v, tb, level, u0):
    print(f"[gmesh] trees={len(t2v)} level={level} prec={cfg.precision} flux={cfg.flux} lim={cfg.limiter}",
          flush=True)
    gm = CA.GMesh(v, t2v, tb, level)
    ctx = CA.GpuContext(cfg)
    ctx.upload_gmesh(gm)
    ctx.set_state(u0)
    return ctx, gm


def sim_of(cfg, v, t2v, tb, level):
    return OL.OracleSim(oracle_cfg(cfg), None, trees=(GU.tree_vertices(v, t2v), tb, level))


def box_state(n, seed, nc=1, amp=0.1):
    rng = np.random.default_rng(seed)
    rho = 1.0 + amp * rng.uniform(-1, 1, n)
    vel = 0.3 * rng.normal(size=(n, 3))
    p = 1.0 + amp * rng.uniform(-1, 1, n)
    u = np.zeros((n, 4 + nc, 4))
    base = W.prim_to_cons(rho, vel, p)
    u[:, :4, 0] = base[:, :4]
    if nc == 1:
        u[:, 4, 0] = rho
    else:
        y = rng.uniform(0.2, 0.8, n)
        u[:, 4, 0] = rho * y
        u[:, 5, 0] = rho * (1 - y)
    u[:, :, 1:] = 0.05 * rng.normal(size=(n, 4 + nc, 3)) * np.abs(u[:, :, :1])
    return u


BOX = GU.box_trees(2, 2, 2, distort=0.15, seed=5)


@pytest.mark.parametrize("flux", ["LF", "HLLC", "CD", "GODUNOV"])
def test_gmesh_residual_matches_oracle(flux):
    v, t2v, tb = BOX
    cfg = CA.make_config(precision=CA.FP64, flux=flux, bnd=BND6, bnd_param=PAR6, grav=(0.1, -0.2, 0.3))
    u0 = box_state(64, 1)
    ctx, _ = gdevice(cfg, v, t2v, tb, 1, u0)
    got = ctx.residual()
    sim = sim_of(cfg, v, t2v, tb, 1)
    sim.set_state(u0)
    want = sim.residual()
    tol = 1e-9 if flux == "GODUNOV" else 1e-12
    assert var_err(got, want) <= tol, var_err(got, want)


@pytest.mark.parametrize("model,lim", [("EULER", True), ("NS", True), ("NS", False)])
def test_gmesh_steps_match_oracle(model, lim):
    v, t2v, tb = BOX
    cfg = CA.make_config(precision=CA.FP64, model=model, flux="HLLC", limiter=lim, bnd=BND6, bnd_param=PAR6,
                         grav=(0.1, -0.2, 0.3))
    u0 = box_state(64, 2)
    ctx, _ = gdevice(cfg, v, t2v, tb, 1, u0)
    ctx.run(2e-3, 5)
    sim = sim_of(cfg, v, t2v, tb, 1)
    sim.set_state(u0)
    sim.run(2e-3, 5)
    err = var_err(ctx.get_state(), sim.get_state())
    assert err <= 1e-11, err


@pytest.mark.parametrize("nc", [1, 2])
def test_gmesh_mass_flow_two_species(nc):
    v, t2v, tb = BOX
    mf = [0.35, 1.3, 1.6, -0.6, 0.8, 0.0, 0.95] + ([1.0] if nc == 1 else [0.75, 0.25])
    bnd = ("MASS_FLOW", "MASS_FLOW", "MASS_FLOW", "MASS_FLOW", "OUTLET", "WALL_SLIP")
    kw = {} if nc == 1 else dict(eos="MIX", comp_m=(CA.GAS_R, 0.5 * CA.GAS_R), comp_cp=(3.5, 5.0))
    u0 = box_state(64, 3, nc=nc)
    out = {}
    for prec in (CA.FP64, CA.FP32):
        cfg = CA.make_config(precision=prec, flux="HLLC", limiter=True, bnd=bnd, bnd_param=[mf] * 4 + [None, None],
                             **kw)
        ctx, _ = gdevice(cfg, v, t2v, tb, 1, u0)
        out[prec] = [ctx.residual()]
        ctx.run(2e-3, 3)
        out[prec].append(ctx.get_state())
        if prec == CA.FP64:
            sim = sim_of(cfg, v, t2v, tb, 1)
            sim.set_state(u0)
            assert var_err(out[prec][0], sim.residual()) <= 1e-12
            sim.run(2e-3, 3)
            assert var_err(out[prec][1], sim.get_state()) <= 1e-11
    assert relerr(out[CA.FP32][0], out[CA.FP64][0]) < 2e-5
    assert relerr(out[CA.FP32][1], out[CA.FP64][1]) < 2e-5


def klin_case(prec, level):
    d = np.load(KLIN)
    kinds = {v: k for k, v in CA.BND.items()}
    bnd = tuple(kinds[int(t)] for t in d["bnd_type"])
    cfg = CA.make_config(precision=prec, model={0: "EULER", 1: "NS"}[int(d["model"])],
                         flux={v: k for k, v in CA.FLUX.items()}[int(d["flux"])], limiter=bool(d["limiter"]),
                         comp_m=tuple(d["comp_m"]), comp_cp=(float(d["comp_cp"][0]),), bnd=bnd,
                         bnd_param=[list(r) for r in d["bnd_param"]])
    gm = CA.GMesh(d["vertices"], d["tree_to_vertex"], d["tree_bnd"], level)
    # the task's region: uniform V, T, P (the initial condition of src/charm_init.c:12-73)
    m, cp = float(d["comp_m"][0]), float(d["comp_cp"][0])
    cv = cp - CA.GAS_R / m
    t, p = float(d["region_t"]), float(d["region_p"])
    rho = p * m / (CA.GAS_R * t)
    vel = d["region_v"]
    u0 = np.zeros((gm.n, 5, 4))
    u0[:, 0:3, 0] = rho * vel
    u0[:, 3, 0] = rho * (cv * t + 0.5 * (vel ** 2).sum())
    u0[:, 4, 0] = rho
    c = np.sqrt(cp / cv * p / rho)
    dt = min(float(d["tau"]), float(d["cfl"]) * gm.geometry()["dh"].min() / (np.abs(vel).max() + c))
    return d, cfg, gm, u0, dt


@pytest.mark.parametrize("level", [0, 1])
def test_reference_sample_steps_match_oracle(level):
    """the reference's own sample at its settings: the wedge turns the Mach-2
    inflow, so five steps build the oblique shock at the wedge's corner"""
    d, cfg, gm, u0, dt = klin_case(CA.FP64, level)
    print(f"[klin] level {level} leaves {gm.n} dt {dt:.4g}", flush=True)
    ctx = CA.GpuContext(cfg)
    ctx.upload_gmesh(gm)
    ctx.set_state(u0)
    ctx.run(dt, 5)
    got = ctx.get_state()
    assert not np.allclose(got, u0, rtol=1e-6, atol=0)              # the flow has moved
    sim = OL.OracleSim(oracle_cfg(cfg), None,
                       trees=(GU.tree_vertices(d["vertices"], d["tree_to_vertex"]), d["tree_bnd"], level))
    sim.set_state(u0)
    sim.run(dt, 5)
    err = phys_err(got, sim.get_state())
    assert err <= 1e-11, err


def test_reference_sample_fp32_close_to_fp64():
    res = {}
    for prec in (CA.FP64, CA.FP32):
        d, cfg, gm, u0, dt = klin_case(prec, 0)
        ctx = CA.GpuContext(cfg)
        ctx.upload_gmesh(gm)
        ctx.set_state(u0)
        ctx.run(dt, 20)
        res[prec] = ctx.get_state()
    err = np.sqrt(((res[CA.FP32] - res[CA.FP64]) ** 2).sum() / (res[CA.FP64] ** 2).sum())
    assert err < 1e-5, err


def test_octree_entry_points_refuse_a_multi_tree_context():
    v, t2v, tb = BOX
    cfg = CA.make_config(precision=CA.FP64, flux="HLLC", bnd=BND6, bnd_param=PAR6)
    ctx, _ = gdevice(cfg, v, t2v, tb, 0, box_state(8, 4))
    with pytest.raises(CA.CharmError, match="multi-tree"):
        ctx.get_dt(1)
    with pytest.raises(CA.CharmError, match="multi-tree"):
        ctx.conservation_sums()
    # a viscous configuration is refused at upload
    cfg2 = CA.make_config(precision=CA.FP64, flux="HLLC", bnd=BND6, bnd_param=PAR6, use_visc=True,
                          transport=[dict(ml0=0.05, kp0=0.1)])
    ctx2 = CA.GpuContext(cfg2)
    with pytest.raises(CA.CharmError, match="multi-tree"):
        ctx2.upload_gmesh(CA.GMesh(v, t2v, tb, 0))
