"""BR2 option of the oracle's viscous flux (an extension: the reference's
viscous flux is BR1, src/models/ns/charm_model_ns_diff_integrals.c:202-382),
pinned on the CPU by closed forms derived from the lifting's definition,
l_f = A^-1 oint_f [Q(avg) - Q(own)] phi, rather than from the code:

* a shear layer v = v0 | v1 across x = 1/2 (periodic box, constant state per
  leaf, constant mu): the lifting of the jump face has the trace
  2 mu dv / h on both sides (mean mu dv / (2h) plus the x slope 3 mu dv / h at
  phi_x = 1/2), the own-trace stress is zero, so BR2's rho v mean residual is
  -/+ 2 eta mu dv h on the two leaves at the jump and zero elsewhere; BR1's is
  -/+ 5/2 mu dv h there and +/- 1/2 mu dv h one leaf further out (its
  auxiliary stress spreads to the neighbouring faces);
* the same layer with the jump face hanging (big leaves on the left): per
  sub-face the big side's lifting has mean mu dv / (8h), x slope
  3 mu dv / (4h) and y, z slopes +-3 mu dv / (8h), the small side's the
  conforming 2 mu dv / h_small trace; the big leaf's residual is
  -(75/32) eta mu dv h and each small one's +(75/128) eta mu dv h;
* without jumps (a linear field on a hanging mesh) every lifting vanishes
  and BR2 equals BR1;
* the surface flux is single-valued: the viscous mean residuals of a random
  state telescope to zero over a periodic hanging mesh.
"""
import numpy as np
import pytest

import mesh_util as MU
import oracle_lib as OL

GAS_R = OL.GAS_R
MU0 = 0.05


def cfg_visc(eta, bnd=("PERIODIC",) * 6, **kw):
    tr = [dict(ml0=MU0, kp0=0.0, t0=273.0, ts=110.0, ml_type=0, kp_type=0)]
    return OL.make_config(comp_m=(GAS_R,), comp_cp=((3.5,),), transport=tr, use_visc=True,
                          bnd=bnd, visc_br2_eta=eta, **kw)


def shear_state(leaves, v0, v1):
    c, _ = MU.centers(leaves)
    n = len(leaves)
    u = np.zeros((n, 5, 4))
    v = np.where(c[:, 0] < 0.5, v0, v1)
    u[:, 1, 0] = v                                   # rho = 1, p = 1
    u[:, 3, 0] = 2.5 + 0.5 * v * v
    u[:, 4, 0] = 1.0
    return u


def visc(leaves, u, eta, **kw):
    s = OL.OracleSim(cfg_visc(eta, **kw), leaves)
    s.set_state(u)
    return s.ns_visc()


@pytest.mark.parametrize("eta", [1.0, 3.0, 6.0])
def test_br2_shear_layer_closed_form(eta):
    lv = MU.uniform_leaves(3)
    h = 1.0 / 8
    v0, v1 = 0.2, -0.5
    dv = v1 - v0
    u = shear_state(lv, v0, v1)
    col = np.floor(MU.centers(lv)[0][:, 0] / h).astype(int)
    r2 = visc(lv, u, eta)[:, 1, 0]
    r1 = visc(lv, u, 0.0)[:, 1, 0]
    scale = MU0 * abs(dv) * h
    # BR2: only the leaves at the two jumps (x = 1/2 and the periodic x = 0)
    want2 = np.zeros(8)
    want2[[3, 0]] = -2 * eta * MU0 * dv * h             # left of the jump v0 -> v1 / right of v1 -> v0
    want2[[4, 7]] = 2 * eta * MU0 * dv * h
    np.testing.assert_allclose(r2, want2[col], rtol=0, atol=1e-13 * scale)
    # BR1 for contrast: the auxiliary stress reaches the next faces (columns
    # 2-5 around the jump at x = 1/2; 6, 7, 0, 1 around the periodic one,
    # which runs v1 -> v0)
    want1 = np.array([-2.5, 0.5, 0.5, -2.5, 2.5, -0.5, -0.5, 2.5]) * MU0 * dv * h
    np.testing.assert_allclose(r1, want1[col], rtol=0, atol=1e-13 * scale)


@pytest.mark.parametrize("eta", [1.0, 6.0])
def test_br2_hanging_shear_layer_closed_form(eta):
    l3 = MU.uniform_leaves(3)
    l4 = MU.uniform_leaves(4)
    c3 = MU.centers(l3)[0][:, 0]
    c4 = MU.centers(l4)[0][:, 0]
    lv = MU.sort_morton(np.concatenate([l3[c3 < 0.5], l4[c4 > 0.5]]))
    h = 1.0 / 8
    v0, v1 = 0.3, -0.1
    dv = v1 - v0
    u = shear_state(lv, v0, v1)
    c, hh = MU.centers(lv)
    r = visc(lv, u, eta)[:, 1, 0]
    big = np.isclose(c[:, 0], 0.5 - h / 2)
    small = np.isclose(c[:, 0], 0.5 + h / 4)
    assert big.sum() == 64 and small.sum() == 256
    scale = MU0 * abs(dv) * h
    np.testing.assert_allclose(r[big], -75.0 / 32.0 * eta * MU0 * dv * h, rtol=0, atol=1e-13 * scale)
    np.testing.assert_allclose(r[small], 75.0 / 128.0 * eta * MU0 * dv * h, rtol=0, atol=1e-13 * scale)
    # no jump at their faces, zero own-trace stress: nothing
    quiet = (np.isclose(c[:, 0], 0.5 - 3 * h / 2) | np.isclose(c[:, 0], 0.5 + 3 * h / 4))
    np.testing.assert_allclose(r[quiet], 0.0, atol=1e-13 * scale)


def test_br2_equals_br1_without_jumps():
    lv = MU.random_adapted(2, 3, seed=4, frac=0.3)
    G = np.array([[0.3, -0.2, 0.5], [0.1, 0.4, -0.3], [-0.25, 0.2, 0.15]])
    b = np.array([0.4, -0.25, 0.1])

    def fn(x):
        # every conserved field linear, so P1 holds it exactly and the two
        # traces at a face agree: no jumps (T itself is not linear)
        out = np.zeros((len(x), 5))
        out[:, 0:3] = (x - 0.5) @ G.T                  # rho = 1
        out[:, 3] = 3.0 + (x - 0.5) @ b
        out[:, 4] = 1.0
        return out
    u = MU.project_p1(lv, fn, 5)
    tr = [dict(ml0=MU0, kp0=0.2, t0=273.0, ts=110.0, ml_type=0, kp_type=0)]

    def run(eta):
        cfg = OL.make_config(comp_m=(GAS_R,), comp_cp=((3.5,),), transport=tr, use_visc=True,
                             bnd=("OUTLET",) * 6, visc_br2_eta=eta)
        s = OL.OracleSim(cfg, lv)
        s.set_state(u)
        return s.ns_visc(), s.ns_aux(), (s.ns_aux_own() if eta > 0 else None)
    r1, a1, _ = run(0.0)
    r2, a2, own = run(6.0)
    np.testing.assert_array_equal(a2, a1)             # tau_h is BR1's auxiliary field
    np.testing.assert_allclose(own, a1, rtol=0, atol=1e-12 * np.abs(a1).max())
    np.testing.assert_allclose(r2, r1, rtol=0, atol=1e-12 * np.abs(r1).max())


def test_br2_viscous_means_telescope_on_periodic_hanging_mesh():
    lv = MU.random_adapted(2, 3, seed=9, frac=0.35, periodic=(True, True, True))
    rng = np.random.default_rng(5)
    n = len(lv)
    u = np.zeros((n, 5, 4))
    r = 1.0 + 0.2 * rng.uniform(-1, 1, n)
    vel = 0.3 * rng.normal(size=(n, 3))
    p = 1.0 + 0.2 * rng.uniform(-1, 1, n)
    u[:, 0:3, 0] = r[:, None] * vel
    u[:, 3, 0] = p / 0.4 + 0.5 * r * (vel ** 2).sum(1)
    u[:, 4, 0] = r
    u[:, :, 1:] = 0.05 * rng.normal(size=(n, 5, 3)) * np.abs(u[:, :, :1])
    tr = [dict(ml0=MU0, kp0=0.2, t0=273.0, ts=110.0, ml_type=0, kp_type=0)]
    cfg = OL.make_config(comp_m=(GAS_R,), comp_cp=((3.5,),), transport=tr, use_visc=True,
                         bnd=("PERIODIC",) * 6, visc_br2_eta=6.0)
    s = OL.OracleSim(cfg, lv)
    s.set_state(u)
    rv = s.ns_visc()
    tot = rv[:, :, 0].sum(0)
    ab = np.abs(rv[:, :, 0]).sum(0)
    np.testing.assert_allclose(tot, 0.0, atol=1e-13 * ab.max())
    # and BR2 differs from BR1 on such a state
    s1 = OL.OracleSim(OL.make_config(comp_m=(GAS_R,), comp_cp=((3.5,),), transport=tr, use_visc=True,
                                     bnd=("PERIODIC",) * 6), lv)
    s1.set_state(u)
    assert np.abs(s1.ns_visc() - rv).max() > 1e-3 * np.abs(rv).max()
