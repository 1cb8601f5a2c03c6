"""CPU checks that pin the parts of the oracle that consistency, conservation
and symmetry tests leave open: the HLLC star region and the Barth-Jespersen
vertex rule.  Each is written from a source independent of the reference's
code and is chosen so that a plausible restatement mistake fails it
(tests/test_oracle_mutants.py builds such mistakes into the oracle and runs
this file against them).

HLLC: Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics (3rd
ed.), section 10.4: the star states U*_K (eq. 10.39) and F*_K = F_K +
S_K (U*_K - U_K) (eq. 10.38), the contact speed S* (eq. 10.37) and the
pressure-based wave speeds (eqs. 10.59-10.61).  The reference writes the
same flux in the F*_K = [S*(S_K U_K - F_K) + S_K (p_K + rho_K (S_K - u_K)
(S* - u_K)) D*] / (S_K - S*) form (src/charm_fluxes.c:319-437); the two are
equal algebraically, not in their arithmetic.

BJ: the limiter bounds the reconstruction at the eight vertices of the cell
(src/limiters/charm_limiter_bj.c:279-299).  For a P1 Taylor field u0 +
s_x phi_x + s_y phi_y with phi_d in [-1/2, 1/2] on the cell, the vertex
extremes are u0 +- (|s_x| + |s_y|) / 2, which fixes psi by hand.
"""
import numpy as np
import pytest

import mesh_util as MU
import oracle_lib as OL

R = OL.GAS_R


# ------------------------------------------------------------ HLLC
def toro_hllc(PL, PR, n, nc):
    """3-D HLLC flux from Toro's star states, frame-free: the star velocity
    is S* n plus side K's tangential velocity."""
    n = np.asarray(n, dtype=float)
    sides = []
    for P in (PL, PR):
        v = np.array([P.u, P.v, P.w])
        c = np.array([P.c[i] for i in range(nc)])
        E = P.r * P.e_tot
        sides.append(dict(r=P.r, v=v, un=v @ n, p=P.p, E=E, a=P.cz, g=P.gam, c=c))
    L, Rr = sides
    # PVRS pressure estimate (eq. 10.61) and wave speeds (eqs. 10.59-10.60)
    rbar, abar = 0.5 * (L["r"] + Rr["r"]), 0.5 * (L["a"] + Rr["a"])
    pstar = max(0.0, 0.5 * (L["p"] + Rr["p"]) - 0.5 * (Rr["un"] - L["un"]) * rbar * abar)

    def qk(s):
        if pstar <= s["p"]:
            return 1.0
        return np.sqrt(1.0 + (s["g"] + 1.0) / (2.0 * s["g"]) * (pstar / s["p"] - 1.0))

    SL = L["un"] - L["a"] * qk(L)
    SR = Rr["un"] + Rr["a"] * qk(Rr)
    # contact speed (eq. 10.37)
    SS = (Rr["p"] - L["p"] + L["r"] * L["un"] * (SL - L["un"]) - Rr["r"] * Rr["un"] * (SR - Rr["un"])) / \
         (L["r"] * (SL - L["un"]) - Rr["r"] * (SR - Rr["un"]))
    K, SK = (L, SL) if SS >= 0 else (Rr, SR)

    def U(s):
        return np.concatenate([s["r"] * s["v"], [s["E"]], s["r"] * s["c"]])

    def F(s):
        return np.concatenate([s["r"] * s["un"] * s["v"] + s["p"] * n, [(s["E"] + s["p"]) * s["un"]],
                               s["r"] * s["c"] * s["un"]])

    # star state (eq. 10.39)
    fac = K["r"] * (SK - K["un"]) / (SK - SS)
    vstar = K["v"] + (SS - K["un"]) * n
    Estar = K["E"] / K["r"] + (SS - K["un"]) * (SS + K["p"] / (K["r"] * (SK - K["un"])))
    Ustar = fac * np.concatenate([vstar, [Estar], K["c"]])
    return F(K) + SK * (Ustar - U(K)), dict(SL=SL, SR=SR, SS=SS)


def _gas(nc):
    if nc == 1:
        return OL.make_config(flux="HLLC")
    gams, ms = (1.4, 1.67), (0.029, 0.004)
    cps = tuple((g * R / ((g - 1) * m),) for g, m in zip(gams, ms))
    return OL.make_config(n_comp=2, comp_m=ms, comp_cp=cps, eos="MIX", flux="HLLC")


@pytest.mark.parametrize("nc", [1, 2])
def test_hllc_star_region_equals_toro_star_states(nc):
    """Random subsonic pairs on random normals, star region of either side,
    S* strictly inside (SL, SR) so the reference's clamp is inactive."""
    cfg = _gas(nc)
    rng = np.random.default_rng(41 + nc)
    seen = {0: 0, 1: 0}
    worst = 0.0
    while min(seen.values()) < 40:
        pscale = 1.0 if nc == 1 else 1e5
        rl, rr = rng.uniform(0.3, 2.0, 2)
        pl, pr = pscale * rng.uniform(0.3, 2.0, 2)
        cl = rng.dirichlet((1, 1)) if nc == 2 else None
        cr = rng.dirichlet((1, 1)) if nc == 2 else None
        PL0 = OL.prim_from(cfg, rl, 0, 0, 0, pl, c=cl)
        vs = 0.4 * PL0.cz
        PL = OL.prim_from(cfg, rl, *(vs * rng.normal(size=3)), pl, c=cl)
        PR = OL.prim_from(cfg, rr, *(vs * rng.normal(size=3)), pr, c=cr)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        ref, w = toro_hllc(PL, PR, n, nc)
        if not (w["SL"] < 0 < w["SR"] and w["SL"] < w["SS"] < w["SR"]):
            continue
        seen[0 if w["SS"] >= 0 else 1] += 1
        q = OL.flux(cfg, "HLLC", PL, PR, n)
        scale = np.abs(ref).max()
        err = np.abs(q - ref).max() / scale
        worst = max(worst, err)
        assert err < 1e-12, (err, w)


@pytest.mark.parametrize("un", [0.0, 0.35, -0.35])
def test_hllc_isolated_contact_is_exact(un):
    """An isolated contact (equal normal velocity and pressure, jumps in
    density, tangential velocity and composition) is resolved exactly: the
    flux is the physical flux of the side the contact leaves behind; at rest
    it is the pressure alone."""
    cfg = _gas(2)
    n = np.array([0.36, 0.48, 0.8])
    t = np.array([0.8, -0.6, 0.0])
    p = 1.3e5
    vl = un * 300.0 * n + 40.0 * t
    vr = un * 300.0 * n - 25.0 * t + 10.0 * np.cross(n, t)
    PL = OL.prim_from(cfg, 1.2, *vl, p, c=(0.9, 0.1))
    PR = OL.prim_from(cfg, 0.3, *vr, p, c=(0.2, 0.8))
    q = OL.flux(cfg, "HLLC", PL, PR, n)
    K = PL if un >= 0 else PR
    v = np.array([K.u, K.v, K.w])
    vn = v @ n
    exact = np.concatenate([K.r * vn * v + p * n, [(K.r * K.e_tot + p) * vn],
                            [K.r * K.c[i] * vn for i in range(2)]])
    # rounding scale per component: the terms S_K U_K and F_K the flux is made of
    UK = np.concatenate([K.r * v, [K.r * K.e_tot], [K.r * K.c[i] for i in range(2)]])
    smax = abs(vn) + max(PL.cz, PR.cz) * 2
    tol = 1e-13 * (smax * np.abs(UK) + np.abs(exact) + p)
    assert (np.abs(q - exact) <= tol).all(), (q, exact, tol)
    if un == 0.0:
        assert (np.abs(q - np.concatenate([p * n, [0.0, 0.0, 0.0]])) <= tol).all()


# ------------------------------------------------------------ BJ
def test_bj_vertex_psi_hand_value():
    """One interior leaf with slopes on two axes, neighbour means above and
    below it; psi_v = min(1, a_v / d_v, b_v / d_v) with d_v = (|s_x| +
    |s_y|) / 2 the vertex deviation.  Evaluating at face centres would give
    d_v = max(|s_x|, |s_y|) / 2 and a psi larger by 0.2 or more here."""
    level = 2
    leaves = MU.uniform_leaves(level)
    h = MU.ROOT >> level
    key = {tuple(q[:3]): i for i, q in enumerate(leaves.tolist())}
    e = key[(h, h, h)]
    ex_p, ex_m = key[(2 * h, h, h)], key[(0, h, h)]
    cfg = OL.make_config(limiter=True)
    base = np.array([0.1, 0.0, 0.0, 2.5, 1.0])      # ru rv rw re rho
    scale = np.maximum(np.abs(base), 0.1)
    u = np.zeros((len(leaves), 5, 4))
    u[:, :, 0] = base
    a = (0.030 + 0.004 * np.arange(5)) * scale       # +x neighbour above
    b = (0.045 + 0.003 * np.arange(5)) * scale       # -x neighbour below
    sx = 0.08 * scale * np.array([1, -1, 1, 1, -1])
    sy = 0.06 * scale * np.array([-1, 1, 1, -1, 1])
    u[ex_p, :, 0] += a
    u[ex_m, :, 0] -= b
    u[e, :, 1] = sx
    u[e, :, 2] = sy
    s = OL.OracleSim(cfg, leaves)
    s.set_state(u)
    s.limiter()
    out = s.get_state()
    d = 0.5 * (np.abs(sx) + np.abs(sy))
    psi = np.minimum(1.0, np.minimum(a, b) / d)
    psi_face = np.minimum(1.0, np.minimum(a, b) / (0.5 * np.maximum(np.abs(sx), np.abs(sy))))
    assert (psi < 1).all() and (psi_face - psi > 0.2).all()
    assert np.allclose(out[e, :, 0], u[e, :, 0], rtol=0, atol=0)
    assert np.allclose(out[e, :, 1], psi * sx, rtol=1e-13, atol=1e-16)
    assert np.allclose(out[e, :, 2], psi * sy, rtol=1e-13, atol=1e-16)
    assert (out[e, :, 3] == 0).all()
