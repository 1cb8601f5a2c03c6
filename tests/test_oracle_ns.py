"""Navier-Stokes pieces of the oracle (viscous BR1 terms, diffusion
coefficients, chemistry), pinned on the CPU by closed forms, published
values and invariants:

* auxiliary stress / heat flux: a linear velocity field gives the exact
  Newtonian stress mu (grad u + grad u^T) - 2/3 mu (div u) I, a linear
  temperature field the exact Fourier flux -k grad T, on every leaf of a
  hanging mesh (outlet boundaries copy the state, so the face averages are
  exact there too);
* viscous residual: a shear flow u = (g (y - 1/2), 0, 0) with constant
  stress has zero viscous momentum residual on interior leaves and the
  viscous-heating energy residual -V mu g^2;
* diffusion: the binary coefficient equals the Chapman-Enskog value for
  O2-N2 at 273.2 K and 1 atm (0.1754 cm^2/s with the tabulated collision
  integral; 0.181 cm^2/s measured), its temperature
  dependence follows the tabulated collision integral, and equal pair
  coefficients give d_i = D for any composition (Wilke's mixture rule);
* chemistry: the semi-implicit map decays a first-order reactant as
  exp(-k dt) to third order, conserves mass to the same order, and the
  energy source equals the reaction rate times the enthalpy change.
"""
import numpy as np
import pytest

import mesh_util as MU
import oracle_lib as OL

GAS_R = OL.GAS_R


def cfg_visc(ml0=0.1, kp0=0.2, nc=1, use_visc=True, **kw):
    tr = [dict(ml0=ml0, kp0=kp0, t0=273.0, ts=110.0, ml_type=0, kp_type=0)] * nc
    return OL.make_config(n_comp=nc, comp_m=(GAS_R,) * nc, comp_cp=((3.5,),) * nc,
                          eos="IDEAL" if nc == 1 else "MIX", transport=tr, use_visc=use_visc, **kw)


def state_from(leaves, fn, nv=5):
    return MU.project_p1(leaves, fn, nv)


def test_aux_stress_of_linear_velocity_is_exact():
    lv = MU.random_adapted(2, 3, seed=4, frac=0.3)
    G = np.array([[0.3, -0.2, 0.5], [0.1, 0.4, -0.3], [-0.25, 0.2, 0.15]])
    mu = 0.1

    def fn(x):
        u = (x - 0.5) @ G.T
        out = np.zeros((len(x), 5))
        out[:, 0:3] = u                       # rho = 1
        out[:, 3] = 2.5 + 0.5 * (u ** 2).sum(1)
        out[:, 4] = 1.0
        return out
    sim = OL.OracleSim(cfg_visc(ml0=mu), lv)
    sim.set_state(state_from(lv, fn))
    aux = sim.ns_aux()
    div = np.trace(G)
    tau = mu * (G + G.T) - 2.0 / 3.0 * mu * div * np.eye(3)
    want = np.array([tau[0, 0], tau[1, 1], tau[2, 2], tau[0, 1], tau[0, 2], tau[1, 2]])
    np.testing.assert_allclose(aux[:, :6, 0], np.broadcast_to(want, (len(lv), 6)), rtol=0, atol=1e-13)
    np.testing.assert_allclose(aux[:, :6, 1:], 0.0, atol=1e-13)


def test_aux_heat_flux_of_linear_temperature_is_exact():
    lv = MU.random_adapted(2, 3, seed=5, frac=0.3)
    b = np.array([40.0, -25.0, 10.0])
    kp = 0.2

    def fn(x):
        t = 300.0 + (x - 0.5) @ b                 # cv = 2.5, M = R: e = 2.5 T
        out = np.zeros((len(x), 5))
        out[:, 3] = 2.5 * t
        out[:, 4] = 1.0
        return out
    sim = OL.OracleSim(cfg_visc(kp0=kp), lv)
    sim.set_state(state_from(lv, fn))
    aux = sim.ns_aux()
    np.testing.assert_allclose(aux[:, 6:, 0], np.broadcast_to(-kp * b, (len(lv), 3)), rtol=1e-12)
    np.testing.assert_allclose(aux[:, 6:, 1:], 0.0, atol=1e-9)
    np.testing.assert_allclose(aux[:, :6], 0.0, atol=1e-12)


def test_viscous_residual_shear_flow():
    lv = MU.uniform_leaves(3)
    g, mu = 0.8, 0.05

    def fn(x):
        out = np.zeros((len(x), 5))
        out[:, 0] = g * (x[:, 1] - 0.5)
        out[:, 3] = 2.5 + 0.5 * out[:, 0] ** 2
        out[:, 4] = 1.0
        return out
    u = state_from(lv, fn)
    r = {}
    for visc in (True, False):
        sim = OL.OracleSim(cfg_visc(ml0=mu, kp0=0.0, use_visc=visc), lv)
        sim.set_state(u)
        r[visc] = sim.ns_residual()
    rv = r[True] - r[False]
    c, h = MU.centers(lv)
    inner = np.all((c > h[:, None]) & (c < 1 - h[:, None]), axis=1)
    assert inner.sum() == 6 ** 3
    np.testing.assert_allclose(rv[inner, :3, :], 0.0, atol=1e-14)
    np.testing.assert_allclose(rv[inner, 3, 0], -h[inner] ** 3 * mu * g * g, rtol=1e-11)


def test_viscous_terms_vanish_for_uniform_flow():
    lv = MU.random_adapted(2, 3, seed=6, frac=0.3)

    def fn(x):
        out = np.zeros((len(x), 5))
        out[:, 0:3] = [0.3, -0.1, 0.2]
        out[:, 3] = 2.5 + 0.5 * 0.14
        out[:, 4] = 1.0
        return out
    u = state_from(lv, fn)
    r = {}
    for visc in (True, False):
        sim = OL.OracleSim(cfg_visc(use_visc=visc), lv)
        sim.set_state(u)
        r[visc] = sim.ns_residual()
    np.testing.assert_allclose(r[True], r[False], rtol=0, atol=1e-14)


def diff_cfg(ms, sig, ek, t_press=None):
    nc = len(ms)
    tr = [dict(ml0=1e-5, kp0=0.02, sig=sig[i], ek=ek[i]) for i in range(nc)]
    return OL.make_config(n_comp=nc, comp_m=ms, comp_cp=((1000.0,),) * nc, eos="MIX",
                          transport=tr, use_visc=True, use_diff=True)


def mix_state(leaves, cfg, y, t, p):
    """uniform mixture at rest: mass fractions y, temperature t, pressure p"""
    y = np.asarray(y, dtype=np.float64)
    ms = np.array([cfg.comp[i].m for i in range(cfg.n_comp)])
    m = 1.0 / (y / ms).sum()
    cp = 1000.0
    cv = cp - GAS_R / m
    rho = p * m / (GAS_R * t)
    u = np.zeros((len(leaves), 4 + len(y), 4))
    u[:, 3, 0] = rho * cv * t
    u[:, 4:, 0] = rho * y
    return u


def test_binary_diffusion_o2_n2():
    # O2 (M 0.032, sigma 3.467 A, eps/k 106.7 K) and N2 (0.028, 3.798 A, 71.4 K)
    cfg = diff_cfg((0.032, 0.028), (3.467, 3.798), (106.7, 71.4))
    lv = MU.uniform_leaves(1)
    sim = OL.OracleSim(cfg, lv)
    sim.set_state(mix_state(lv, cfg, (0.3, 0.7), 273.2, 101325.0))
    d = sim.ns_diff_coeffs()
    assert np.allclose(d[:, 0], d[:, 1], rtol=1e-14)          # binary: d_0 = d_1 = D_01
    # Chapman-Enskog, D = 0.0018583 sqrt(T^3 (1/M_A + 1/M_B)) / (p sigma_AB^2 Omega_D) cm^2/s
    # with sigma_AB = 3.6325 A, T* = 273.2 / 87.28 = 3.13, Omega_D(3.13) = 0.9383
    # (tabulated): 0.1754 cm^2/s; measured 0.181 cm^2/s
    assert abs(d[0, 0] / 1.754e-5 - 1.0) < 0.01
    assert abs(d[0, 0] / 1.81e-5 - 1.0) < 0.05
    # temperature dependence: D ~ T^1.5 / Omega_D(T*), Omega_D tabulated at
    # T* = 1, 2, 5 (1.440, 1.075, 0.8422)
    eps = np.sqrt(106.7 * 71.4)
    vals = {}
    for ts, om in ((1.0, 1.440), (2.0, 1.075), (5.0, 0.8422)):
        sim.set_state(mix_state(lv, cfg, (0.3, 0.7), ts * eps, 101325.0))
        vals[ts] = (sim.ns_diff_coeffs()[0, 0] / (ts * eps) ** 1.5, om)
    for ts in (2.0, 5.0):
        ratio = vals[ts][0] / vals[1.0][0]
        assert abs(ratio / (vals[1.0][1] / vals[ts][1]) - 1.0) < 3e-3, ts


def test_mixture_rule_equal_pairs():
    cfg = diff_cfg((0.03, 0.03, 0.03), (3.5, 3.5, 3.5), (90.0, 90.0, 90.0))
    lv = MU.uniform_leaves(1)
    sim = OL.OracleSim(cfg, lv)
    for y in ((0.2, 0.3, 0.5), (0.7, 0.1, 0.2)):
        sim.set_state(mix_state(lv, cfg, y, 400.0, 2e5))
        d = sim.ns_diff_coeffs()
        np.testing.assert_allclose(d, d[0, 0], rtol=1e-13)
    sim.set_state(mix_state(lv, cfg, (1.0, 0.0, 0.0), 400.0, 2e5))
    assert np.all(sim.ns_diff_coeffs()[:, 0] == 0.0)            # pure species: no diffusion


def chem_cfg(reactions, h0=(0.0, 0.0, 0.0), parity=False, ms=(0.03, 0.03, 0.03)):
    nc = len(ms)
    tr = [dict(h0=h0[i]) for i in range(nc)]
    return OL.make_config(n_comp=nc, comp_m=ms, comp_cp=((1000.0,),) * nc, eos="MIX",
                          transport=tr, reactions=reactions, t_ref=298.15, chem_parity=parity)


def test_chemistry_first_order_decay_and_mass():
    A, E = 2.0e3, 0.0
    cfg = chem_cfg([((0, -1, -1), (1, -1, -1), A, E, 0.0)])
    lv = MU.uniform_leaves(1)
    errs, dm = [], []
    for dt in (5e-5, 2.5e-5, 1.25e-5):              # k dt = 0.1, 0.05, 0.025
        u = mix_state(lv, cfg, (0.8, 0.2, 0.0), 500.0, 1e5)
        u[:, 4:, 1:] = 0.01                                    # slopes are zeroed by the step
        sim = OL.OracleSim(cfg, lv)
        sim.set_state(u)
        sim.ns_chem_apply(dt)
        v = sim.get_state()
        ca0 = u[0, 4, 0] / 0.03
        errs.append(abs(v[0, 4, 0] / 0.03 - ca0 * np.exp(-A * dt)) / ca0)
        dm.append(abs(v[0, 4:, 0].sum() - u[0, 4:, 0].sum()) / u[0, 4:, 0].sum())
        assert np.all(v[:, 4:, 1:] == 0.0)
        assert np.array_equal(v[:, :4], u[:, :4])
    # third-order local error: halving dt divides it by ~8
    assert 6.5 < errs[0] / errs[1] < 9.5 and 6.5 < errs[1] / errs[2] < 9.5
    assert 6.5 < dm[0] / dm[1] < 9.5


def test_chemistry_energy_source():
    A = 5.0e2
    h0 = (0.0, -2.0e6, 0.0)
    lv = MU.uniform_leaves(1)
    for parity in (False, True):
        cfg = chem_cfg([((0, -1, -1), (1, -1, -1), A, 0.0, 0.0)], h0=h0, parity=parity)
        u = mix_state(lv, cfg, (0.6, 0.4, 0.0), 600.0, 1e5)
        sim = OL.OracleSim(cfg, lv)
        sim.set_state(u)
        rhs = sim.ns_chem_apply(1e-4)
        ca = sim.get_state()[0, 4, 0] / 0.03
        # w = k C_A, h_B - h_A = h0_B - h0_A at equal constant cp
        np.testing.assert_allclose(rhs, A * ca * (h0[1] - h0[0]), rtol=1e-12)
    # A + B -> C: the reference subtracts chem_w[species] (index bug); for
    # species 1 >= reactions the parity mode reads 0 there
    rx = [((0, 1, -1), (2, -1, -1), A, 0.0, 0.0)]
    out = {}
    for parity in (False, True):
        cfg = chem_cfg(rx, h0=(1e5, 3e5, -2e6), parity=parity)
        u = mix_state(lv, cfg, (0.3, 0.3, 0.4), 600.0, 1e5)
        sim = OL.OracleSim(cfg, lv)
        sim.set_state(u)
        out[parity] = (sim.ns_chem_apply(1e-4), sim.get_state())
    st = out[False][1]
    w = A * (st[0, 4, 0] / 0.03) * (st[0, 5, 0] / 0.03)
    dh = 1000.0 * (600.0 - 298.15)                        # h_i = h0_i + cp (T - T_ref)
    hA, hB, hC = 1e5 + dh, 3e5 + dh, -2e6 + dh
    np.testing.assert_allclose(out[False][0], w * (hC - hA - hB), rtol=1e-9)
    np.testing.assert_allclose(out[True][0], w * (hC - hA), rtol=1e-9)


def test_chemistry_molar_energy_source_repeated_species():
    """2 A + B -> 2 C written as left (A, A, B), right (C, C): rate w = k C_A^2
    C_B, net molar rates -2w, -w, +2w; with chem_energy_molar the source is
    sum_i w_i M_i h_i (J/m^3/s), the reference's mode sum_i w_i h_i.  Molar
    masses with 2 M_A + M_B = 2 M_C conserve mass to the map's third order."""
    A = 40.0
    ms = (0.002, 0.032, 0.018)
    h0 = (0.0, 0.0, -1.3e7)
    lv = MU.uniform_leaves(1)
    rx = [((0, 0, 1), (2, 2, -1), A, 0.0, 0.0)]
    out = {}
    for molar in (False, True):
        tr = [dict(h0=h0[i]) for i in range(3)]
        cfg = OL.make_config(n_comp=3, comp_m=ms, comp_cp=((1000.0,),) * 3, eos="MIX", transport=tr,
                             reactions=rx, t_ref=298.15, chem_energy_molar=molar)
        u = mix_state(lv, cfg, (0.05, 0.55, 0.4), 600.0, 1e5)
        sim = OL.OracleSim(cfg, lv)
        sim.set_state(u)
        dt = 1e-5
        out[molar] = (sim.ns_chem_apply(dt), sim.get_state(), u)
    rhs_ref, st, u = out[False]
    ca, cb = st[0, 4, 0] / ms[0], st[0, 5, 0] / ms[1]
    w = A * ca * ca * cb
    dh = 1000.0 * (600.0 - 298.15)
    h = np.array(h0) + dh
    wi = np.array([-2 * w, -w, 2 * w])
    np.testing.assert_allclose(rhs_ref, (wi * h).sum(), rtol=1e-9)
    np.testing.assert_allclose(out[True][0], (wi * np.array(ms) * h).sum(), rtol=1e-9)
    # the species update is the same in both modes
    assert np.array_equal(out[True][1], st)
    # consumption: d C_A = 2 d C_B at first order, mass conserved
    dca = (st[0, 4, 0] - u[0, 4, 0]) / ms[0]
    dcb = (st[0, 5, 0] - u[0, 5, 0]) / ms[1]
    assert abs(dca / dcb - 2.0) < 0.05
    assert abs(st[0, 4:, 0].sum() - u[0, 4:, 0].sum()) < 1e-4 * u[0, 4:, 0].sum()


def test_sutherland_viscosity_of_air_in_stress():
    """Sutherland's law for air (mu0 1.716e-5 Pa s at 273.15 K, S = 110.4 K)
    gives 1.846e-5 Pa s at 300 K (tabulated air viscosity)."""
    lv = MU.uniform_leaves(2)
    G = np.zeros((3, 3))
    G[0, 1] = 1e-3
    tr = [dict(ml0=1.716e-5, kp0=0.0, t0=273.15, ts=110.4, ml_type=1, kp_type=0)]
    cfg = OL.make_config(comp_m=(0.02897,), comp_cp=((1005.0,),), transport=tr, use_visc=True)
    cv = 1005.0 - GAS_R / 0.02897

    def fn(x):
        u = (x - 0.5) @ G.T
        out = np.zeros((len(x), 5))
        out[:, 0:3] = u
        out[:, 3] = cv * 300.0 + 0.5 * (u ** 2).sum(1)
        out[:, 4] = 1.0
        return out
    sim = OL.OracleSim(cfg, lv)
    sim.set_state(state_from(lv, fn))
    tau_xy = sim.ns_aux()[:, 3, 0]
    np.testing.assert_allclose(tau_xy / 1e-3, 1.846e-5, rtol=1e-3)
