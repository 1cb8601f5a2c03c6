"""Adaptation on the CPU: the oracle's estimator, criteria and projection
pinned by closed forms and invariants, and the product's host-side
refine / coarsen decisions (charm_amr_flags) against the oracle's.

Pinning (none of these restates the formula under test):
* estimator: for a globally linear density the Green-Gauss gradient equals
  the exact gradient on every leaf (divergence theorem, exact traces), with
  hanging faces present; a constant state against an inlet face has the
  one-sided closed form (rho - rho_b) / (2h);
* projection: refinement reproduces a linear field exactly, coarsening a
  refined field is the identity, coarsening conserves mass, and a parent
  refined again by the balance equals the closed-form restriction of the
  coarsened parent;
* criteria: hand-built families and thresholds.
"""
import numpy as np
import pytest

import charm_3d_amd as CA
import mesh_util as MU
import oracle_lib as OL

ROOT = MU.ROOT
GAM = 1.4


def linear_state(leaves, b, p=1.0, nc=1):
    """rho = 1 + b . (x - 1/2), u = 0, p constant: exact in the P1 basis."""
    def fn(x):
        rho = 1.0 + (x - 0.5) @ np.asarray(b)
        out = np.zeros((len(x), 4 + nc))
        out[:, 3] = p / (GAM - 1.0)
        out[:, 4] = rho
        return out
    return MU.project_p1(leaves, fn, 4 + nc)


def test_indicator_linear_density_is_exact_on_adapted_mesh():
    lv = MU.random_adapted(2, 4, seed=11, frac=0.3)
    assert len(set(lv[:, 3])) > 1                       # hanging faces present
    b = np.array([0.3, -0.2, 0.45])
    sim = OL.OracleSim(OL.make_config(), lv)
    sim.set_state(linear_state(lv, b))
    err2, rho_c = sim.amr_indicator()
    c, h = MU.centers(lv)
    vol = h ** 3
    want = (b @ b) / 12.0 * vol ** (5.0 / 3.0)
    np.testing.assert_allclose(err2, want, rtol=1e-11, atol=0)
    np.testing.assert_allclose(rho_c, 1.0 + (c - 0.5) @ b, rtol=1e-14)


def test_indicator_inlet_face_one_sided():
    lv = MU.uniform_leaves(2)
    rho0, rho_b = 1.3, 0.7
    # inlet on -x with (u, v, w, T, p): rho_b = p M / (R T) = p / T for M = R
    bnd = ("INLET",) + ("OUTLET",) * 5
    params = [[0.0, 0.0, 0.0, 1.0 / rho_b, 1.0]] + [None] * 5
    sim = OL.OracleSim(OL.make_config(bnd=bnd, bnd_param=params), lv)
    u = np.zeros((len(lv), 5, 4))
    u[:, 3, 0] = 1.0 / (GAM - 1.0)
    u[:, 4, 0] = rho0
    sim.set_state(u)
    err2, _ = sim.amr_indicator()
    h = 0.25
    at_inlet = lv[:, 0] == 0
    g = (rho0 - rho_b) / (2.0 * h)
    np.testing.assert_allclose(err2[at_inlet], g * g / 12.0 * h ** 5, rtol=1e-12)
    # elsewhere only the rounding of the outlet's EOS round trip remains
    assert np.all(err2[~at_inlet] < 1e-20 * g * g * h ** 5)


def _refine_all(leaves):
    out, mp = [], []
    for i, q in enumerate(leaves):
        out.extend(MU._children(tuple(int(v) for v in q)))
        mp.extend([i] * 8)
    out = np.array(out, dtype=np.int64)
    order = np.argsort(MU.morton_key(out[:, 0], out[:, 1], out[:, 2]), kind="stable")
    return out[order], np.array(mp, dtype=np.int64)[order]


def test_projection_refine_exact_and_coarsen_inverse():
    lv = MU.random_adapted(1, 3, seed=5, frac=0.4)
    rng = np.random.default_rng(1)
    b = rng.normal(size=(5, 3))
    a = rng.normal(size=5)

    def fn(x):
        return a[None, :] + (x - 0.5) @ b.T
    u = MU.project_p1(lv, fn, 5)
    fine, mp = _refine_all(lv)
    uf = OL.amr_project(5, lv, u, fine, mp)
    np.testing.assert_allclose(uf, MU.project_p1(fine, fn, 5), rtol=0, atol=1e-14)
    # coarsen every family back: first child index of each parent
    back = np.array([-1 - 8 * i for i in range(len(lv))], dtype=np.int64)
    # the fine leaves of parent i are 8i..8i+7 in Morton order
    assert np.array_equal(np.repeat(np.arange(len(lv)), 8), np.sort(mp))
    ur = np.random.default_rng(2).normal(size=uf.shape)
    uc = OL.amr_project(5, fine, ur, lv, back)
    uff = OL.amr_project(5, lv, uc, fine, mp)
    ucc = OL.amr_project(5, fine, uff, lv, back)
    np.testing.assert_allclose(ucc, uc, rtol=0, atol=1e-13)       # coarsen o refine = id
    vol_f, vol_c = MU.centers(fine)[1] ** 3, MU.centers(lv)[1] ** 3
    np.testing.assert_allclose((vol_c[:, None] * uc[:, :, 0]).sum(0),
                               (vol_f[:, None] * ur[:, :, 0]).sum(0), rtol=1e-13)


def test_projection_coarsened_parent_refined_again():
    """Coarsen a family, then project into one octant of the parent: the
    closed-form parent (means and first moments of the children) restricted
    to the octant."""
    fam = np.array(MU._children((0, 0, 0, 0)), dtype=np.int64)          # level 1
    fam = fam[np.argsort(MU.morton_key(fam[:, 0], fam[:, 1], fam[:, 2]))]
    rng = np.random.default_rng(3)
    u = rng.normal(size=(8, 5, 4))
    target = np.array([[ROOT // 4, 0, ROOT // 4, 2]], dtype=np.int64)   # inside child 4
    got = OL.amr_project(5, fam, u, target, np.array([-1], dtype=np.int64))
    c, h = MU.centers(fam)
    sig = np.sign(c - 0.5)
    par = np.zeros((5, 4))
    par[:, 0] = u[:, :, 0].mean(0)
    for d in range(3):
        par[:, 1 + d] = (0.375 * sig[:, d, None] * u[:, :, 0] + u[:, :, 1 + d] / 16.0).sum(0)
    tc, th = MU.centers(target)
    want = np.zeros((5, 4))
    want[:, 0] = par[:, 0] + par[:, 1:] @ (tc[0] - 0.5)
    want[:, 1:] = par[:, 1:] * th[0]
    np.testing.assert_allclose(got[0], want, rtol=0, atol=1e-14)


def _flags_case():
    """Two level-2 families inside the level-1 octant 0, the rest level 1."""
    lv = []
    for q in MU._children((0, 0, 0, 0)):
        if q == (0, 0, 0, 1):
            lv.extend(MU._children(q))
        else:
            lv.append(q)
    lv = np.array(lv, dtype=np.int64)
    return lv[np.argsort(MU.morton_key(lv[:, 0], lv[:, 1], lv[:, 2]))]


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_flags_hand_cases(impl):
    lv = _flags_case()
    n = len(lv)
    vol = (MU.centers(lv)[1]) ** 3
    max_err = 0.1

    def flags(err2, rho, min_level=0, max_level=4):
        if impl == "product":
            return CA.amr_flags(lv, err2, rho, max_err, min_level, max_level)
        sim = OL.OracleSim(OL.make_config(), lv)
        return sim.amr_flags(err2, rho, max_err, min_level, max_level)

    fam = lv[:, 3] == 2
    err2 = np.zeros(n)
    rho = np.ones(n)
    f = flags(err2, rho)
    assert np.all(f[fam] == -1) and np.all(f[~fam] == 0)          # smooth family coarsens
    assert np.all(flags(err2, rho, min_level=2)[fam] == 0)        # at MIN_LEVEL: kept
    e2 = err2.copy()
    e2[3] = 1.0001 * max_err ** 2 * vol[3]                         # one member refines
    f = flags(e2, rho)
    assert f[3] == 1 and np.all(f[fam & (np.arange(n) != 3)] == 0)
    assert np.all(flags(e2, rho, max_level=2)[fam] == 0)          # capped: kept, not coarsened
    # every member at 0.9 of its bound (sum 7.2 e^2 V): the centre-density
    # spread decides against 8 e^2 V
    e2 = np.where(fam, 0.9 * max_err ** 2 * vol, 0.0)
    alt = np.where(np.arange(8) % 2 == 0, 1, -1)
    r2 = rho.copy()
    r2[fam] += alt * 0.35 * max_err                                # 7.2 + 0.98 > 8: kept
    assert np.all(flags(e2, r2)[fam] == 0)
    r2[fam] = 1.0 + alt * 0.3 * max_err                            # 7.2 + 0.72 < 8: coarsened
    assert np.all(flags(e2, r2)[fam] == -1)
    e2[fam.nonzero()[0][2]] = 1.0001 * max_err ** 2 * vol[fam][0]
    f = flags(e2, r2)                                              # a member above its bound
    assert np.all(f[fam] == np.where(np.arange(8) == 2, 1, 0))
    r2 = rho.copy()
    r2[fam] += np.where(np.arange(8) % 2 == 0, 1, -1) * 0.11 * max_err   # jump: 8 d^2 V < 8 e^2 V
    assert np.all(flags(err2, r2)[fam] == -1)
    r2[fam] = 1.0 + np.where(np.arange(8) % 2 == 0, 1, -1) * 1.01 * max_err
    assert np.all(flags(err2, r2)[fam] == 0)


def test_flags_product_matches_oracle_random():
    rng = np.random.default_rng(7)
    seen_ref = seen_coa = 0
    for seed in range(4):
        lv = MU.random_adapted(1, 4, seed=seed, frac=0.35)
        vol = MU.centers(lv)[1] ** 3
        err2 = vol * 0.01 * rng.uniform(0.0, 0.3, len(lv)) * np.where(rng.random(len(lv)) < 0.1, 6.0, 1.0)
        rho = 1.0 + 0.01 * rng.normal(size=len(lv))
        sim = OL.OracleSim(OL.make_config(), lv)
        for me, lo, hi in ((0.1, 0, 5), (0.08, 2, 4), (0.2, 1, 3)):
            a = CA.amr_flags(lv, err2, rho, me, lo, hi)
            b = sim.amr_flags(err2, rho, me, lo, hi)
            assert np.array_equal(a, b)
            seen_ref += int((a == 1).sum())
            seen_coa += int((a == -1).sum())
    assert seen_ref > 50 and seen_coa > 500
