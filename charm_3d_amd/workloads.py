"""Synthetic workloads of the benchmark configurations (initial states, BCs).

Initial conditions are FP64 L2 projections onto the P1 Taylor basis with the
8-point Gauss rule (the projection template of charm_replace_quads,
src/charm_amr.c:279-304); the reference only sets coefficient 0 from
piecewise-constant regions (src/charm_init.c:46-68).
"""
import numpy as np

from . import FP32, GAS_R, ROOT_LEN, Forest, GpuContext, make_config, recommended_chunks

GAMMA = 1.4
S3 = 1.0 / np.sqrt(3.0)
QGP = np.array([[sx, sy, sz] for sz in (-S3, S3) for sy in (-S3, S3) for sx in (-S3, S3)])


def centers(leaves):
    lv = np.asarray(leaves, dtype=np.float64)
    h = (ROOT_LEN >> lv[:, 3].astype(np.int64)) / ROOT_LEN
    return lv[:, :3] / ROOT_LEN + 0.5 * h[:, None], h


def project_p1(leaves, fn, nv):
    """u[e][v][k] = A^-1 sum_gp wJ phi_k f(x_gp) with A = V diag(1, 1/12, 1/12, 1/12)."""
    c, h = centers(leaves)
    n = len(c)
    acc = np.zeros((n, nv, 4))
    for g in QGP:
        x = c + 0.5 * h[:, None] * g[None, :]
        f = fn(x)
        acc[:, :, 0] += f
        acc[:, :, 1:] += f[:, :, None] * (0.5 * g)[None, None, :]
    u = np.empty_like(acc)
    u[:, :, 0] = acc[:, :, 0] / 8.0
    u[:, :, 1:] = acc[:, :, 1:] * (12.0 / 8.0)
    return u


def prim_to_cons(rho, vel, p, gamma=GAMMA):
    out = np.zeros(rho.shape + (5,))
    out[..., 0:3] = rho[..., None] * vel
    out[..., 3] = p / (gamma - 1) + 0.5 * rho * (vel ** 2).sum(-1)
    out[..., 4] = rho
    return out


def c1_exact(x, t=0.0):
    """rho = 1 + 0.2 sin(2 pi (x + y + z - 1.75 t)), u = (1, 0.5, 0.25), p = 1."""
    rho = 1 + 0.2 * np.sin(2 * np.pi * (x[:, 0] + x[:, 1] + x[:, 2] - 1.75 * t))
    vel = np.broadcast_to(np.array([1.0, 0.5, 0.25]), x.shape)
    return prim_to_cons(rho, vel, np.ones_like(rho))


def c1(level=4, precision=1):
    """Euler ideal gas, periodic unit cube, LF, no limiter, SSP-RK3."""
    leaves = Forest.uniform(level).leaves()
    cfg = make_config(precision=precision, model="NS", flux="LF", limiter=False,
                      bnd=("PERIODIC",) * 6, comp_m=(GAS_R,), comp_cp=(3.5,), cfl=0.3)
    return cfg, leaves, project_p1(leaves, c1_exact, 5), (True, True, True)


def c2(level=7, precision=1, seed=2):
    """Riemann box: (rho, p) = (1, 1) | (0.125, 0.1) at x = 0.5, u = 0, cell means
    of rho and p scaled by 1 + 0.01 U(-1, 1) (seeded), P1 slopes 0,
    transmissive boundaries, HLLC + Barth-Jespersen, SSP-RK3."""
    leaves = Forest.uniform(level).leaves()
    c, _ = centers(leaves)
    rng = np.random.default_rng(seed)
    left = c[:, 0] < 0.5
    rho = np.where(left, 1.0, 0.125) * (1 + 0.01 * rng.uniform(-1, 1, len(c)))
    p = np.where(left, 1.0, 0.1) * (1 + 0.01 * rng.uniform(-1, 1, len(c)))
    u = np.zeros((len(c), 5, 4))
    u[:, :, 0] = prim_to_cons(rho, np.zeros((len(c), 3)), p)
    cfg = make_config(precision=precision, model="NS", flux="HLLC", limiter=True,
                      bnd=("OUTLET",) * 6, comp_m=(GAS_R,), comp_cp=(3.5,), cfl=0.3)
    return cfg, leaves, u, (False, False, False)


def distorted_box_trees(n=4, distort=0.1, seed=2):
    """A unit box of n^3 hexahedral trees whose vertices are displaced by up
    to `distort` of the tree size (interior vertices in 3-D, vertices on a
    side inside that side's plane), so leaves are non-affine hexahedra with
    non-planar faces.  Returns vertices [nv][3], tree_to_vertex [nt][8]
    (z-order corners), tree_bnd [nt][6] (the box side's slot 0-5 on the
    box's sides, -1 between trees)."""
    rng = np.random.default_rng(seed)
    g = np.linspace(0.0, 1.0, n + 1)
    idx = np.stack(np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij"), -1)
    v = g[idx].reshape(-1, 3).copy()
    free = ((idx > 0) & (idx < n)).reshape(-1, 3)         # axes along which a vertex may move
    v += np.where(free, distort / n * rng.uniform(-1, 1, v.shape), 0.0)
    vid = np.arange((n + 1) ** 3).reshape(n + 1, n + 1, n + 1)
    t = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    t2v = np.stack([vid[t[:, 0] + (c & 1), t[:, 1] + ((c >> 1) & 1), t[:, 2] + (c >> 2)] for c in range(8)], 1)
    tb = np.full((len(t), 6), -1, np.int32)
    for f in range(6):
        a, hi = f // 2, f % 2
        tb[t[:, a] == (n - 1 if hi else 0), f] = f
    return v, t2v.astype(np.int32), tb


def c2_multi_tree(n=4, level=5, precision=1, seed=2, distort=0.1):
    """The Riemann box of c2 on a multi-tree mesh: n^3 distorted trees
    (distorted_box_trees) refined uniformly to `level` (n = 4, level 5:
    2,097,152 leaves, the size of the level-7 octree), cell means from the
    leaf centres, P1 slopes 0, HLLC + BJ, SSP-RK3, transmissive boundaries.
    dt = 0.3 min(dh) / max(|u| + c), dh the leaves' bounding-box extents.
    Returns (cfg, GMesh, u0, dt)."""
    from . import GMesh
    v, t2v, tb = distorted_box_trees(n, distort, seed)
    gm = GMesh(v, t2v, tb, level)
    c = gm.geometry()["c"]
    rng = np.random.default_rng(seed)
    left = c[:, 0] < 0.5
    rho = np.where(left, 1.0, 0.125) * (1 + 0.01 * rng.uniform(-1, 1, len(c)))
    p = np.where(left, 1.0, 0.1) * (1 + 0.01 * rng.uniform(-1, 1, len(c)))
    u = np.zeros((len(c), 5, 4))
    u[:, :, 0] = prim_to_cons(rho, np.zeros((len(c), 3)), p)
    cfg = make_config(precision=precision, model="NS", flux="HLLC", limiter=True,
                      bnd=("OUTLET",) * 6, comp_m=(GAS_R,), comp_cp=(3.5,), cfl=0.3)
    dt = 0.3 * gm.geometry()["dh"].min() / np.sqrt(GAMMA * p / rho).max()
    return cfg, gm, u, dt


C5_BOX = ((0.6, 0.6, 0.6), (0.8, 0.8, 0.8))


def c5_pulse(level=6, precision=1, inlet=(0.5, 1.0, 1.0), leaves=None):
    """Gaussian density pulse rho = 1 + 0.1 exp(-|x - x0|^2 / 0.1^2),
    x0 = (0.3, 0.3, 0.3), p = 1, at rest; Euler terms, HLLC, no limiter,
    SSP-RK3; inlet on the -x face with (u, T, p) = inlet (v = w = 0; the
    adjoint's parameters), transmissive elsewhere.  Nondimensional gas:
    R / M = 1, cp 3.5.  Without a limiter the P1 scheme needs CFL 0.1
    (dt = 0.1 h / max(|u| + c)); at 0.3 it diverges within 50 steps."""
    if leaves is None:
        leaves = Forest.uniform(level).leaves()

    def fn(x):
        r2 = ((x - 0.3) ** 2).sum(1)
        rho = 1 + 0.1 * np.exp(-r2 / 0.01)
        return prim_to_cons(rho, np.zeros((len(x), 3)), np.ones(len(x)))

    u, t, p = inlet
    cfg = make_config(precision=precision, model="NS", flux="HLLC", limiter=False,
                      bnd=("INLET",) + ("OUTLET",) * 5, bnd_param=[(u, 0.0, 0.0, t, p)] + [None] * 5,
                      comp_m=(GAS_R,), comp_cp=(3.5,), cfl=0.1)
    return cfg, leaves, project_p1(leaves, fn, 5), (False, False, False)


# two-species bubble case: species A (gamma 1.4, M 0.029) with 16 bubbles of
# species B (gamma 1.67, M 0.004); cp_i = gamma_i R / ((gamma_i - 1) M_i)
C3_GAMMA = (1.4, 1.67)
C3_M = (0.029, 0.004)
C3_CP = tuple(g * GAS_R / ((g - 1.0) * m) for g, m in zip(C3_GAMMA, C3_M))


def c3_bubbles(seed=3, n=16):
    """Bubble centres U(0, 1)^3 and radii U(0.05, 0.15) (seeded)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, (n, 3)), rng.uniform(0.05, 0.15, n)


def c3_state_fn(centres, radii, width, p=1.0e5, t=300.0, vel=(100.0, 50.0, 25.0)):
    """Conserved variables (ru, rv, rw, re, rc_A, rc_B) at points: species-B mass
    fraction 0.5 (1 + tanh((r_b - d_b) / width)), maximum over the bubbles,
    uniform p, T and velocity; mixture EOS (cp = sum c_i cp_i,
    M = 1 / sum c_i / M_i, cv = cp - R / M)."""
    vel = np.asarray(vel, dtype=np.float64)

    def fn(x):
        cb = np.zeros(len(x))
        for c, r in zip(centres, radii):
            d = np.sqrt(((x - c) ** 2).sum(1))
            cb = np.maximum(cb, 0.5 * (1.0 + np.tanh((r - d) / width)))
        ca = 1.0 - cb
        m = 1.0 / (ca / C3_M[0] + cb / C3_M[1])
        rho = p * m / (GAS_R * t)
        cv = ca * C3_CP[0] + cb * C3_CP[1] - GAS_R / m
        out = np.empty((len(x), 6))
        out[:, 0:3] = rho[:, None] * vel[None, :]
        out[:, 3] = rho * (cv * t + 0.5 * (vel ** 2).sum())
        out[:, 4] = rho * ca
        out[:, 5] = rho * cb
        return out
    return fn


def c3_config(precision=FP32, max_err=0.04, min_level=5, max_level=8, refine_period=20):
    """Exact Riemann solver, mixture EOS, Barth-Jespersen, SSP-RK3, transmissive
    boundaries, re-adaptation every `refine_period` steps."""
    return make_config(precision=precision, model="NS", flux="GODUNOV", limiter=True, eos="MIX",
                       comp_m=C3_M, comp_cp=C3_CP, bnd=("OUTLET",) * 6, cfl=0.3,
                       max_err=max_err, min_level=min_level, max_level=max_level,
                       refine_period=refine_period)


def c3(precision=FP32, seed=3, min_level=5, max_level=8, max_err=0.04, width=None):
    """Initial leaf set of the bubble case: from a uniform min_level forest,
    refine (then 2:1 balance) every leaf whose device estimator exceeds the
    adaptation threshold (err2 > MAX_ERROR^2 V, i.e. |grad rho| h >
    sqrt(12) MAX_ERROR) on the projected initial state, one level per pass,
    re-projecting the analytic state after each pass.  Needs a GPU.
    Returns (cfg, forest, u0)."""
    centres, radii = c3_bubbles(seed)
    if width is None:
        width = 3.0 / (1 << max_level)
    fn = c3_state_fn(centres, radii, width)
    cfg = c3_config(precision, max_err, min_level, max_level)
    forest = Forest.uniform(min_level)
    ce, cf = recommended_chunks(cfg)
    for _ in range(max_level - min_level):
        leaves = forest.leaves()
        ctx = GpuContext(cfg)
        ctx.upload_mesh(forest.flatten(ce, cf))
        ctx.set_state(project_p1(leaves, fn, 6))
        err2, _ = ctx.amr_indicator()
        ctx.close()
        vol = ((ROOT_LEN >> leaves[:, 3].astype(np.int64)) / ROOT_LEN) ** 3
        flags = ((err2 > max_err ** 2 * vol) & (leaves[:, 3] < max_level)).astype(np.int8)
        if not flags.any():
            break
        forest.adapt(flags, max_level, min_level)
    return cfg, forest, project_p1(forest.leaves(), fn, 6)


# reacting hot-spot case: H2, O2, H2O, N2.  cp_i(T) is a power series
# (R / M_i)(a1 + a2 T + a3 T^2 + a4 T^3 + a5 T^4): the least-squares quartic
# through the two-range NASA 7-coefficient polynomials (GRI-Mech 3.0 thermo
# data, 200-1000 K and 1000-3500 K) sampled over 250-1700 K.  A single range
# does not serve a 300-1500 K hot spot: the 200-1000 K fit of H2 extrapolates
# to cp / R = 1.18 at 1500 K (3.88 from the high range), which makes the
# mixture energy (cp(T) - R / M) T non-monotonic, and the reference's Newton
# temperature then lands on a spurious root.  Sutherland viscosity and
# conductivity with KP0 = ML0 cp(T0) / 0.72, Lennard-Jones diffusion
# parameters, and one reaction 2 H2 + O2 -> 2 H2O with k = A exp(-15000 / T),
# A = 1e9 (the reference's left / right species lists hold H2 and H2O twice;
# the rate is k C_H2^2 C_O2 and mass is conserved).
# The reference's energy source multiplies molar rates (mol / m^3 / s) by
# mass enthalpies (J / kg); at A = 1e9 that source releases about twenty
# times the local internal energy per step at the hot-spot peak and the
# state fails within two steps (on the oracle as on the device).  The case
# therefore runs with chem_energy_molar = 1: the rates times the molar
# enthalpies M_i h_i, whose release is bounded by the fuel present.
C4_SPECIES = ("H2", "O2", "H2O", "N2")
C4_M = (0.002016, 0.031999, 0.018015, 0.028014)
C4_NASA_LOW = ((2.34433112, 7.98052075e-03, -1.94781510e-05, 2.01572094e-08, -7.37611761e-12),
               (3.78245636, -2.99673416e-03, 9.84730201e-06, -9.68129509e-09, 3.24372837e-12),
               (4.19864056, -2.03643410e-03, 6.52040211e-06, -5.48797062e-09, 1.77197817e-12),
               (3.29867700, 1.40824040e-03, -3.96322200e-06, 5.64151500e-09, -2.44485400e-12))
C4_NASA_HIGH = ((3.33727920, -4.94024731e-05, 4.99456778e-07, -1.79566394e-10, 2.00255376e-14),
                (3.28253784, 1.48308754e-03, -7.57966669e-07, 2.09470555e-10, -2.16717794e-14),
                (3.03399249, 2.17691804e-03, -1.64072518e-07, -9.70419870e-11, 1.68200992e-14),
                (2.92664000, 1.48797680e-03, -5.68476000e-07, 1.00970380e-10, -6.75335100e-15))


def c4_cp_over_r():
    """quartic cp / R per species (T in K), fitted to both NASA ranges"""
    t = np.linspace(250.0, 1700.0, 400)
    x = t / 1000.0
    a = np.vstack([x ** k for k in range(5)]).T
    out = []
    for lo, hi in zip(C4_NASA_LOW, C4_NASA_HIGH):
        y = np.where(t <= 1000.0, sum(c * t ** k for k, c in enumerate(lo)),
                     sum(c * t ** k for k, c in enumerate(hi)))
        coef = np.linalg.lstsq(a, y, rcond=None)[0]
        out.append(tuple(float(coef[k]) / 1000.0 ** k for k in range(5)))
    return tuple(out)


C4_A = 1.0e9
C4_H0 = (0.0, 0.0, -241826.0 / 0.018015, 0.0)            # formation enthalpy at 298.15 K, J/kg
C4_ML = ((8.41e-6, 72.0), (1.919e-5, 127.0), (1.12e-5, 1064.0), (1.663e-5, 111.0))   # mu(273 K), S
C4_LJ = ((2.92, 38.0), (3.458, 107.4), (2.605, 572.4), (3.621, 97.53))               # sigma A, eps/k K
C4_Y = (0.02, 0.22, 0.0, 0.76)
# BR2 viscous flux (the case names a BR2-type flux; the reference's own is
# BR1): penalty factor = the number of faces of a hexahedron
C4_BR2_ETA = 6.0


def c4_config(precision=FP32, limiter=False, a=None, chem_energy_molar=True, visc_br2_eta=C4_BR2_ETA):
    cps = tuple(tuple(a * GAS_R / m for a in fit) for fit, m in zip(c4_cp_over_r(), C4_M))
    t0 = 273.0
    tr = []
    for i in range(4):
        cp0 = sum(c * t0 ** k for k, c in enumerate(cps[i]))
        ml0, ts = C4_ML[i]
        tr.append(dict(ml0=ml0, kp0=ml0 * cp0 / 0.72, t0=t0, ts=ts, ml_type=1, kp_type=1,
                       sig=C4_LJ[i][0], ek=C4_LJ[i][1], h0=C4_H0[i]))
    rx = [((0, 0, 1), (2, 2, -1), C4_A if a is None else a, 15000.0 * GAS_R, 0.0)]
    return make_config(precision=precision, model="NS", flux="HLLC", limiter=limiter, eos="MIX",
                       comp_m=C4_M, comp_cp=cps, bnd=("OUTLET",) * 6, cfl=0.3, transport=tr,
                       use_visc=True, use_diff=True, t_ref=298.15, reactions=rx,
                       chem_energy_molar=chem_energy_molar, visc_br2_eta=visc_br2_eta)


def c4_state_fn(cfg, r0=0.1, p=101325.0):
    """T = 300 + 1200 exp(-r^2 / r0^2) about the centre, p = 1 atm, at rest,
    premixed mass fractions; e = (cp(T) - R / M) T as the mixture EOS"""
    y = np.asarray(C4_Y)
    ms = np.asarray(C4_M)
    m = 1.0 / (y / ms).sum()
    cps = [[cfg.comp[i].cp[k] for k in range(cfg.comp[i].n_cp)] for i in range(4)]

    def fn(x):
        r2 = ((x - 0.5) ** 2).sum(1)
        t = 300.0 + 1200.0 * np.exp(-r2 / r0 ** 2)
        cp = sum(y[i] * sum(c * t ** k for k, c in enumerate(cps[i])) for i in range(4))
        e = (cp - GAS_R / m) * t
        rho = p * m / (GAS_R * t)
        out = np.zeros((len(x), 8))
        out[:, 3] = rho * e
        out[:, 4:] = rho[:, None] * y[None, :]
        return out
    return fn


def c4(level=6, precision=FP32, **kw):
    """Reacting hot spot: level-6 uniform cube (262,144 leaves), four
    species, HLLC + BR2 viscous terms + diffusion + chemistry, SSP-RK3.
    kw: c4_config's a (rate constant), chem_energy_molar and visc_br2_eta
    (0: the reference's BR1)."""
    leaves = Forest.uniform(level).leaves()
    cfg = c4_config(precision, **kw)
    return cfg, leaves, project_p1(leaves, c4_state_fn(cfg), 8), (False, False, False)


def dof_updates(n_elem, nv, stages, steps):
    """one DOF advanced through one RK stage: E * 4 * V * stages * steps"""
    return n_elem * 4 * nv * stages * steps
