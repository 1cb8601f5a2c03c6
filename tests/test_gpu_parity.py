"""GPU parity: the HIP path (through the C-ABI) against the FP64 oracle.

FP64 device results must match the oracle to 1e-12 relative (the bar the
FP64 path is held to); the FP32 bulk path is compared to the FP64 device
path with the tolerances written in each test.
"""
import os

import numpy as np
import pytest

import charm_3d_amd as CA
import charm_3d_amd.workloads as W
import mesh_util as MU
import oracle_lib as OL

pytestmark = pytest.mark.gpu

REL64 = 1e-12


def oracle_cfg(cfg):
    """The oracle's config for a device config (every field this path reads)."""
    kinds = {v: k for k, v in CA.BND.items()}
    bnd = tuple(kinds[cfg.bnd_type[i]] for i in range(6))
    params = [[cfg.bnd_param[i][j] for j in range(16)] for i in range(6)]
    comps = [cfg.comp[i] for i in range(cfg.n_comp)]
    tr = [dict(ml0=q.ml0, kp0=q.kp0, t0=q.t0, ts=q.ts, h0=q.h0, sig=q.sig, ek=q.ek,
               ml_type=q.ml_type, kp_type=q.kp_type, cp_type=q.cp_type) for q in comps]
    rx = [(tuple(cfg.react[j].left), tuple(cfg.react[j].right), cfg.react[j].a, cfg.react[j].e,
           cfg.react[j].n) for j in range(cfg.n_react)]
    return OL.make_config(n_comp=cfg.n_comp, comp_m=[q.m for q in comps],
                          comp_cp=[tuple(q.cp[k] for k in range(q.n_cp)) for q in comps],
                          eos="IDEAL" if cfg.eos == 0 else "MIX",
                          flux={v: k for k, v in CA.FLUX.items()}[cfg.flux],
                          limiter=bool(cfg.limiter), model="NS" if cfg.model == 1 else "EULER",
                          bnd=bnd, bnd_param=params, grav=[cfg.grav[i] for i in range(3)],
                          cfl=cfg.cfl, transport=tr, use_visc=bool(cfg.use_visc),
                          use_diff=bool(cfg.use_diff), t_ref=cfg.t_ref, reactions=rx,
                          chem_parity=bool(cfg.chem_parity),
                          chem_energy_molar=bool(cfg.chem_energy_molar),
                          visc_br2_eta=cfg.visc_br2_eta)


def device(cfg, leaves, periodic, u0):
    print(f"[gpu] n={len(leaves)} prec={cfg.precision} flux={cfg.flux} lim={cfg.limiter}", flush=True)
    ctx = CA.GpuContext(cfg)
    f = CA.Forest.from_leaves(leaves, periodic)
    ce, cf = CA.recommended_chunks(cfg)
    ctx.upload_mesh(f.flatten(ce, cf))
    ctx.set_state(u0)
    return ctx


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def random_state(leaves, rng, amp=0.1, nc=1):
    n = len(leaves)
    rho = 1.0 + amp * rng.uniform(-1, 1, n)
    vel = 0.3 * rng.normal(size=(n, 3))
    p = 1.0 + amp * rng.uniform(-1, 1, n)
    u = np.zeros((n, 4 + nc, 4))
    base = W.prim_to_cons(rho, vel, p)
    u[:, :4, 0] = base[:, :4]
    if nc == 1:
        u[:, 4, 0] = rho
    else:
        frac = rng.dirichlet(np.ones(nc), size=n)
        u[:, 4:, 0] = rho[:, None] * frac
    u[:, :, 1:] = 0.2 * amp * rng.normal(size=(n, 4 + nc, 3)) * np.abs(u[:, :, :1])
    return u


MESHES = {
    "uniform3": (lambda: MU.uniform_leaves(3), (False,) * 3),
    "uniform3p": (lambda: MU.uniform_leaves(3), (True,) * 3),
    "adapt": (lambda: MU.random_adapted(1, 3, seed=11), (False,) * 3),
    "adapt_p": (lambda: MU.random_adapted(2, 4, seed=13, frac=0.15, periodic=(True,) * 3), (True,) * 3),
}


@pytest.mark.parametrize("flux", ["LF", "HLLC", "CD", "GODUNOV"])
def test_pointwise_flux_fp64(flux):
    cfg = CA.make_config(precision=CA.FP64, flux=flux)
    ctx = CA.GpuContext(cfg)
    ocfg = oracle_cfg(cfg)
    rng = np.random.default_rng(7)
    n = 256
    pl = np.column_stack([rng.uniform(0.3, 2, n), rng.normal(size=(n, 3)), rng.uniform(0.3, 3, n), np.ones(n)])
    pr = np.column_stack([rng.uniform(0.3, 2, n), rng.normal(size=(n, 3)), rng.uniform(0.3, 3, n), np.ones(n)])
    nn = rng.normal(size=(n, 3))
    nn[:32] = np.eye(3)[rng.integers(0, 3, 32)] * rng.choice([-1, 1], (32, 1))
    nn /= np.linalg.norm(nn, axis=1)[:, None]
    got = ctx.flux_batch(flux, pl, pr, nn)
    for i in range(n):
        a = OL.prim_from(ocfg, *pl[i, :5])
        b = OL.prim_from(ocfg, *pr[i, :5])
        ref = OL.flux(ocfg, flux, a, b, nn[i])
        tol = 1e-9 if flux == "GODUNOV" else REL64
        assert np.allclose(got[i], ref, rtol=tol, atol=tol * np.abs(ref).max()), (i, got[i], ref)


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("flux", ["LF", "HLLC", "GODUNOV"])
def test_residual_fp64_matches_oracle(mesh, flux):
    gen, periodic = MESHES[mesh]
    leaves = gen()
    bnd = ("PERIODIC",) * 6 if periodic[0] else ("OUTLET", "WALL_SLIP", "OUTLET", "INLET", "WALL_NO_SLIP", "OUTLET")
    params = [None, None, None, (0.1, -0.2, 0.05, 1.1, 1.2), (0.9,), None]
    cfg = CA.make_config(precision=CA.FP64, flux=flux, bnd=bnd, bnd_param=params, grav=(0.1, -0.3, 0.2))
    u0 = random_state(leaves, np.random.default_rng(3))
    ctx = device(cfg, leaves, periodic, u0)
    got = ctx.residual()
    sim = OL.OracleSim(oracle_cfg(cfg), leaves)
    sim.set_state(u0)
    ref = sim.residual()
    tol = 1e-9 if flux == "GODUNOV" else REL64
    assert relerr(got, ref) < tol


@pytest.mark.parametrize("nc,flux", [(1, "HLLC"), (1, "LF"), (2, "HLLC")])
def test_mass_flow_boundary_matches_oracle(nc, flux):
    """BOUND_MASS_FLOW (src/charm_bnd_cond.c:101-146) on a 2:1 mesh: the
    direction cosines enter through +x and -y (inflow: the parameters'
    composition, total temperature and mass flux) and leave through -x and
    +y (outflow pressure); residual and three SSP-RK3 steps with the
    limiter against the oracle, FP64, and FP32 against FP64."""
    leaves = MU.random_adapted(1, 3, seed=11)
    # (flux, static p, total T, cosines, outflow p, composition)
    mf = [0.35, 1.3, 1.6, -0.6, 0.8, 0.0, 0.95] + ([1.0] if nc == 1 else [0.75, 0.25])
    bnd = ("MASS_FLOW", "MASS_FLOW", "MASS_FLOW", "MASS_FLOW", "OUTLET", "WALL_SLIP")
    params = [mf] * 4 + [None, None]
    if nc == 1:
        kw = {}
    else:
        kw = dict(eos="MIX", comp_m=(CA.GAS_R, 0.5 * CA.GAS_R), comp_cp=(3.5, 5.0))
    res = {}
    for prec in (CA.FP64, CA.FP32):
        cfg = CA.make_config(precision=prec, flux=flux, bnd=bnd, bnd_param=params, limiter=True, **kw)
        u0 = random_state(leaves, np.random.default_rng(21), nc=nc)
        ctx = device(cfg, leaves, (False,) * 3, u0)
        res[prec] = [ctx.residual()]
        ctx.run(2e-3, 3)
        res[prec].append(ctx.get_state())
        if prec == CA.FP64:
            sim = OL.OracleSim(oracle_cfg(cfg), leaves)
            sim.set_state(u0)
            assert relerr(res[prec][0], sim.residual()) < REL64
            sim.run(2e-3, 3)
            assert relerr(res[prec][1], sim.get_state()) < 1e-11
    assert relerr(res[CA.FP32][0], res[CA.FP64][0]) < 2e-5
    assert relerr(res[CA.FP32][1], res[CA.FP64][1]) < 2e-5


@pytest.mark.parametrize("mesh", ["uniform3", "adapt", "adapt_p"])
def test_limiter_fp64_matches_oracle(mesh):
    gen, periodic = MESHES[mesh]
    leaves = gen()
    bnd = ("PERIODIC",) * 6 if periodic[0] else ("OUTLET", "WALL_SLIP", "OUTLET", "INLET", "OUTLET", "OUTLET")
    params = [None, None, None, (0.1, -0.2, 0.05, 1.1, 1.2), None, None]
    cfg = CA.make_config(precision=CA.FP64, limiter=True, bnd=bnd, bnd_param=params)
    u0 = random_state(leaves, np.random.default_rng(5), amp=0.3)
    ctx = device(cfg, leaves, periodic, u0)
    ctx.limiter_bj()
    sim = OL.OracleSim(oracle_cfg(cfg), leaves)
    sim.set_state(u0)
    sim.limiter()
    ref = sim.get_state()
    got = ctx.get_state()
    assert relerr(got, ref) < REL64
    assert (np.abs(got[:, :, 1:]) < np.abs(u0[:, :, 1:]) - 1e-14).any()


@pytest.mark.parametrize("case", ["c1_l3", "c2_l4", "c2_adapt", "euler_lim"])
def test_steps_fp64_match_oracle(case):
    if case == "c1_l3":
        cfg, leaves, u0, periodic = W.c1(level=3, precision=CA.FP64)
        nsteps = 10
    elif case == "c2_l4":
        cfg, leaves, u0, periodic = W.c2(level=4, precision=CA.FP64)
        nsteps = 10
    elif case == "c2_adapt":
        leaves = MU.random_adapted(2, 4, seed=17, frac=0.1)
        cfg = CA.make_config(precision=CA.FP64, flux="HLLC", limiter=True)
        c, _ = W.centers(leaves)
        left = c[:, 0] < 0.5
        u0 = np.zeros((len(leaves), 5, 4))
        u0[:, :, 0] = W.prim_to_cons(np.where(left, 1.0, 0.125), np.zeros((len(c), 3)), np.where(left, 1.0, 0.1))
        periodic = (False,) * 3
        nsteps = 8
    else:
        cfg, leaves, u0, periodic = W.c2(level=3, precision=CA.FP64)
        cfg.model = CA.MODEL["EULER"]
        nsteps = 10
    ctx = device(cfg, leaves, periodic, u0)
    dt = ctx.get_dt(mode=1)
    sim = OL.OracleSim(oracle_cfg(cfg), leaves)
    sim.set_state(u0)
    assert abs(dt - sim.cfl_dt(1)) <= REL64 * dt
    ctx.run(dt, nsteps)
    sim.run(dt, nsteps)
    ref = sim.get_state()
    assert relerr(ctx.get_state(), ref) < 1e-11
    sums = ctx.conservation_sums()
    assert np.allclose(sums, sim.conservation_sums(), rtol=1e-12, atol=1e-13)


def test_fp32_bulk_vs_fp64_small_c2():
    """FP32 bulk against FP64 on the Riemann box (level 4, 20 steps):
    relative L2 over all coefficients below 1e-5 here; the full-size figure
    is reported by bench.py."""
    cfg64, leaves, u0, periodic = W.c2(level=4, precision=CA.FP64)
    cfg32, _, _, _ = W.c2(level=4, precision=CA.FP32)
    a = device(cfg64, leaves, periodic, u0)
    b = device(cfg32, leaves, periodic, u0)
    dt = a.get_dt(1)
    a.run(dt, 20)
    b.run(dt, 20)
    x, y = a.get_state(), b.get_state()
    l2 = np.linalg.norm(y - x) / np.linalg.norm(x)
    assert l2 < 1e-5, l2
    m0 = (u0[:, 4, 0]).sum()
    assert abs(b.conservation_sums()[4] - a.conservation_sums()[4]) < 1e-5 * abs(a.conservation_sums()[4])
    assert m0 > 0


def test_fp32_deviation_form_meets_the_gate():
    """The FP32 residual in deviation form (fluxes minus the chunk
    reference's flux): on the perturbed Riemann box the residual is within
    a few 1e-7 of the FP64 one, 20 steps stay below the metric's L2 1e-6
    gate, and a uniform state (deviations at the rounding of the reference's
    u_r = m / rho to FP32) has a residual at FP64 rounding of the fluxes."""
    cfg64, leaves, u0, periodic = W.c2(level=5, precision=CA.FP64)
    cfg32, _, _, _ = W.c2(level=5, precision=CA.FP32)
    a = device(cfg64, leaves, periodic, u0)
    b = device(cfg32, leaves, periodic, u0)
    r64, r32 = a.residual(), b.residual()
    rerr = relerr(r32, r64)
    dt = a.get_dt(1)
    a.run(dt, 20)
    b.run(dt, 20)
    x, y = a.get_state(), b.get_state()
    l2 = np.linalg.norm(y - x) / np.linalg.norm(x)
    print("level 5 Riemann box: FP32 residual error %.2e, 20 steps L2 %.2e" % (rerr, l2))
    assert rerr < 1e-6, rerr
    assert l2 < 1e-6, l2
    uni = np.zeros_like(u0)
    uni[:, :, 0] = u0[0, :, 0]
    b.set_state(uni)
    h2 = (1.0 / 32) ** 2                       # face area; the fluxes are O(1)
    assert np.abs(b.residual()).max() < 1e-13 * h2


def test_fp32_residual_close_to_fp64():
    leaves = MU.random_adapted(1, 3, seed=11)
    u0 = random_state(leaves, np.random.default_rng(9))
    r64 = device(CA.make_config(precision=CA.FP64), leaves, (False,) * 3, u0).residual()
    r32 = device(CA.make_config(precision=CA.FP32), leaves, (False,) * 3, u0).residual()
    assert relerr(r32, r64) < 2e-5


def test_nan_state_is_reported():
    leaves = MU.uniform_leaves(2)
    u0 = random_state(leaves, np.random.default_rng(1))
    u0[5, 3, 0] = np.nan
    ctx = device(CA.make_config(precision=CA.FP32), leaves, (False,) * 3, u0)
    with pytest.raises(CA.CharmError, match="SL <= S"):
        ctx.run(1e-3, 1)


def test_reference_entry_points_drive_the_device():
    """charm_model_ns_timestep_single / charm_timesteps through the forest's
    user pointer give the same state as charm_gpu_step."""
    import ctypes as C
    cfg, leaves, u0, periodic = W.c2(level=3, precision=CA.FP64)
    L = CA.lib()
    f = CA.Forest.from_leaves(leaves, periodic)
    dt = 1e-3
    cfg.dt, cfg.tmax = dt, 2.5 * dt
    h = C.c_void_p()
    CA.check(L.charm_ctx_create(f.h, C.byref(cfg), C.byref(h)))
    gpu = L.charm_ctx_gpu(h)
    u = np.ascontiguousarray(u0)
    CA.check(L.charm_gpu_set_state(gpu, u.ctypes.data_as(C.POINTER(C.c_double)), len(u)))
    L.charm_timesteps(f.h)                       # t = 0, dt, 2dt -> 3 steps
    got = np.zeros_like(u)
    CA.check(L.charm_gpu_get_state(gpu, got.ctypes.data_as(C.POINTER(C.c_double)), len(u)))
    L.charm_ctx_destroy(h)
    ref = device(cfg, leaves, periodic, u0)
    ref.run(dt, 3)
    assert np.array_equal(got, ref.get_state())


def test_mixed_precision_residual_is_the_fp32_residual():
    """FP64 state, FP32 residual arithmetic: the residual is the FP32 path's
    on the same (FP32-rounded) state, up to the FP32 rounding of the two
    compiled variants (the compiler may fuse multiply-adds differently):
    per variable, max |difference| <= 2e-6 of the largest |residual|."""
    rng = np.random.default_rng(9)
    leaves = MU.random_adapted(2, 4, seed=9, frac=0.3)
    u0 = random_state(leaves, rng)
    r32 = device(CA.make_config(precision=CA.FP32), leaves, (False,) * 3, u0).residual()
    rmx = device(CA.make_config(precision=CA.FP32_MIX), leaves, (False,) * 3, u0).residual()
    d = max(np.abs(rmx[:, v] - r32[:, v]).max() / np.abs(r32[:, v]).max() for v in range(r32.shape[1]))
    print("mixed vs FP32 residual: max relative difference %.2e" % d)
    assert d <= 2e-6


@pytest.mark.parametrize("case", ["riemann_box", "sine_wave"])
def test_reduced_precision_steps_close_to_fp64(case):
    """FP32 and mixed (FP64 state / update / limiter, FP32 residual) paths
    against the FP64 path: the Riemann box (level 4, 20 steps, HLLC + BJ)
    and the smooth advected sine wave (level 5, 50 steps, LF, no limiter),
    relative L2 over all coefficients."""
    mk = (lambda p: W.c2(level=4, precision=p)) if case == "riemann_box" else (lambda p: W.c1(level=5, precision=p))
    steps = 20 if case == "riemann_box" else 50
    cfg, leaves, u0, periodic = mk(CA.FP64)
    out = {}
    for prec in (CA.FP64, CA.FP32_MIX, CA.FP32):
        c, _, _, _ = mk(prec)
        ctx = device(c, leaves, periodic, u0)
        dt = ctx.get_dt(1)
        ctx.run(dt, steps)
        out[prec] = ctx.get_state()
    ref = out[CA.FP64]
    err = {p: np.linalg.norm(out[p] - ref) / np.linalg.norm(ref) for p in (CA.FP32_MIX, CA.FP32)}
    print("%s L2 vs FP64: mixed %.2e, fp32 %.2e" % (case, err[CA.FP32_MIX], err[CA.FP32]))
    if case == "riemann_box":
        # at the shock the limiter's switches amplify any rounding: both
        # reduced-precision paths land at a few 1e-6
        assert max(err.values()) < 1e-5
    else:
        # smooth flow: the FP64 state removes most of the FP32 error
        assert err[CA.FP32_MIX] < 1e-6 and err[CA.FP32_MIX] < err[CA.FP32] / 3


def test_write_solution_every_file_output_step(tmp_path):
    """charm_timesteps with FILE_OUTPUT_STEP = 2 writes <prefix>_<step>
    before steps 0, 2 and 4 (src/models/euler/charm_model.c:507-510); the
    last file holds the state after 4 steps, as charm_write_vtk of that
    state writes it."""
    import filecmp
    cfg, leaves, u0, periodic = W.c2(level=2, precision=CA.FP64)
    dt = 1e-3
    cfg.dt, cfg.tmax = dt, 4.5 * dt
    cfg.write_period = 2
    prefix = str(tmp_path / "run")
    cfg.out_prefix = prefix.encode()
    f = CA.Forest.from_leaves(leaves, periodic)
    s = CA.Solver(f, cfg)
    s.set_state(u0)
    s.timesteps()                                   # steps 0..4
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == ["run_%08d%s" % (k, sfx) for k in (0, 2, 4) for sfx in (".pvtu", "_0000.vtu")]
    ref = device(cfg, leaves, periodic, u0)
    ref.run(dt, 4)
    CA.write_vtk(str(tmp_path / "ref"), leaves, ref.get_state(), cfg)
    assert filecmp.cmp(str(tmp_path / "run_00000004_0000.vtu"), str(tmp_path / "ref_0000.vtu"), shallow=False)
    s.close()


def test_charm_3d_program_runs_a_task(tmp_path):
    """the charm_3d program on a unit-cube task.yaml (level 2, TAU 1e-3,
    TMAX 4.5e-3: five steps, files every 2): its last file equals the file
    charm_write_vtk writes from the same task run through the API"""
    import subprocess
    import test_task_yaml as TY
    text = (TY.EULER_TASK.replace("MIN_LEVEL: 3", "MIN_LEVEL: 2").replace("MAX_LEVEL: 5", "MAX_LEVEL: 2")
            .replace("REFINE_PERIOD: 10", "REFINE_PERIOD: 0").replace("TAU: 1.0e-2", "TAU: 1.0e-6")
            .replace("TMAX: 50 ", "TMAX: 4.5e-6 ").replace("FILE_OUTPUT_STEP: 100", "FILE_OUTPUT_STEP: 2")
            .replace("LOG_OUTPUT_STEP: 10", "LOG_OUTPUT_STEP: 1"))
    (tmp_path / "task.yaml").write_text(text)
    exe = os.path.join(os.path.dirname(CA.__file__), "charm_3d")
    r = subprocess.run([exe], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("STEP =") == 5
    names = sorted(p.name for p in tmp_path.iterdir() if p.name.startswith("charm_3d_"))
    assert names == ["charm_3d_%08d%s" % (k, s) for k in (0, 2, 4) for s in (".pvtu", "_0000.vtu")]
    t = CA.load_task(str(tmp_path / "task.yaml"))
    f = CA.task_forest(t)
    ctx = device(t.cfg, f.leaves(), f.periodic, CA.task_initial_state(t, f))
    ctx.run(t.cfg.dt, 4)
    CA.write_vtk(str(tmp_path / "api"), f.leaves(), ctx.get_state(), t.cfg)
    assert (tmp_path / "charm_3d_00000004_0000.vtu").read_bytes() == (tmp_path / "api_0000.vtu").read_bytes()
