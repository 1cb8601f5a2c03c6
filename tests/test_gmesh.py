"""Multi-tree meshes on the CPU: the product's host builder (charm_gmesh_build:
exact face matching by canonical tree-face lattice keys, FP64 geometry) against
the oracle's multi-tree restatement (faces matched by physical face centres),
and the oracle's multi-tree traversal pinned against its single-tree one:

* distorted boxes of trees whose corners are relabelled by random cube
  rotations (every face orientation), levels 0-2, and the reference's sample
  mesh (samples/klin_amr, 800 trees; tests/golden/klin_mesh.npz): identical
  face sets (leaf, local face) and boundary slots, geometry to 1e-15;
* one unit-cube tree through the multi-tree constructor equals the octree
  oracle bit for bit;
* a distorted tree at level 1 equals its eight children as trees at level 0
  (corners = the parent's trilinear map; each child relabelled by a random
  rotation): residuals and three steps agree to rounding, so faces matched
  across trees behave as faces inside a tree.
"""
import os

import numpy as np
import pytest

import charm_3d_amd as CA
import gmesh_util as GU
import mesh_util as MU
import oracle_lib as OL

HERE = os.path.dirname(os.path.abspath(__file__))
KLIN = os.path.join(HERE, "golden", "klin_mesh.npz")
BND6 = ("OUTLET", "INLET", "WALL_SLIP", "OUTLET", "WALL_NO_SLIP", "OUTLET")
PAR6 = [None, (0.2, 0.1, 0.0, 1.1, 1.0), None, None, (0.9,), None]


def face_set_product(gm):
    out = set()
    for r in gm.faces():
        if r[1] >= 0:
            out.add(tuple(sorted([(int(r[0]), int(r[2])), (int(r[1]), int(r[3]))])))
        else:
            out.add(((int(r[0]), int(r[2])), ("b", int(-1 - r[1]))))
    return out


def face_set_oracle(sim):
    out = set()
    for r in sim.faces():
        if r[0] == 2:
            out.add(tuple(sorted([(int(r[5]), int(r[1])), (int(r[9]), int(r[2]))])))
        else:
            out.add(((int(r[5]), int(r[1])), ("b", int(r[13]))))
    return out


def oracle_sim(v, t2v, tb, level, cfg=None):
    cfg = cfg or OL.make_config(flux="HLLC", limiter=True, model="EULER", bnd=("OUTLET",) * 6)
    return OL.OracleSim(cfg, None, trees=(GU.tree_vertices(v, t2v), tb, level))


def random_state(n, seed, nv=5):
    rng = np.random.default_rng(seed)
    u = np.zeros((n, nv, 4))
    r = 1.0 + 0.2 * rng.uniform(-1, 1, n)
    vel = 0.3 * rng.normal(size=(n, 3))
    p = 1.0 + 0.2 * rng.uniform(-1, 1, n)
    u[:, 0:3, 0] = r[:, None] * vel
    u[:, 3, 0] = p / 0.4 + 0.5 * r * (vel ** 2).sum(1)
    u[:, 4, 0] = r
    u[:, :, 1:] = 0.05 * rng.normal(size=(n, nv, 3)) * np.abs(u[:, :, :1])
    return u


@pytest.mark.parametrize("shape,distort,level", [((1, 1, 1), 0.0, 1), ((2, 2, 2), 0.15, 0),
                                                 ((2, 2, 1), 0.15, 1), ((3, 2, 2), 0.2, 2)])
def test_gmesh_matches_oracle_multi_tree(shape, distort, level):
    v, t2v, tb = GU.box_trees(*shape, distort=distort, seed=3)
    gm = CA.GMesh(v, t2v, tb, level)
    sim = oracle_sim(v, t2v, tb, level)
    assert gm.n == sim.n == len(t2v) * 8 ** level
    assert face_set_product(gm) == face_set_oracle(sim)
    g = gm.geometry()
    for i in range(gm.n):
        o = sim.geom(i)
        np.testing.assert_allclose(g["c"][i], o["c"], rtol=0, atol=1e-15)
        np.testing.assert_allclose(g["dh"][i], o["dh"], rtol=1e-14)
        assert abs(g["volume"][i] - o["volume"]) <= 1e-14 * o["volume"]
        np.testing.assert_allclose(g["a_inv"][i], o["a_inv"], rtol=0, atol=1e-13 * np.abs(o["a_inv"]).max())
    # leaves: tree, Morton order inside it
    lv = gm.leaves()
    assert (np.diff(lv[:, 0]) >= 0).all() and (lv[:, 4] == level).all()


@pytest.mark.skipif(not os.path.exists(KLIN), reason="fixture absent")
@pytest.mark.parametrize("level", [0, 1])
def test_reference_sample_mesh_matches_oracle(level):
    d = np.load(KLIN)
    gm = CA.GMesh(d["vertices"], d["tree_to_vertex"], d["tree_bnd"], level)
    sim = oracle_sim(d["vertices"], d["tree_to_vertex"], d["tree_bnd"], level)
    assert gm.n == 800 * 8 ** level
    fs = face_set_product(gm)
    assert fs == face_set_oracle(sim)
    # 1,510 shared tree faces and 1,780 boundary quadrangles at level 0
    nb = sum(1 for f in fs if f[1][0] == "b")
    assert (len(fs) - nb, nb) == ((1510, 1780) if level == 0 else (1510 * 4 + 800 * 12, 1780 * 4))


def test_one_tree_equals_the_octree_oracle():
    unit = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]], float)
    cfg = OL.make_config(flux="HLLC", limiter=True, model="NS", bnd=BND6, bnd_param=PAR6)
    a = OL.OracleSim(cfg, MU.uniform_leaves(2))
    b = OL.OracleSim(cfg, None, trees=(unit[None], np.arange(6, dtype=np.int32)[None], 2))
    u = random_state(a.n, 0)
    for s in (a, b):
        s.set_state(u)
    np.testing.assert_array_equal(a.residual(), b.residual())
    for s in (a, b):
        s.run(1e-3, 3)
    np.testing.assert_array_equal(a.get_state(), b.get_state())


def children_as_trees(tv, tb, seed, rotate):
    """the 8 children of a tree as trees of their own (corners = the
    parent's trilinear map at the child corners), optionally each relabelled
    by a random rotation; tree_bnd follows the parent's on outer faces"""
    rng = np.random.default_rng(seed)

    def tmap(xi):
        w = [[1 - x, x] for x in xi]
        return sum(w[0][k & 1] * w[1][(k >> 1) & 1] * w[2][k >> 2] * tv[k] for k in range(8))
    verts, t2v, tbnd, key = [], [], [], {}

    def vid(p):
        k = tuple(p)
        if k not in key:
            key[k] = len(verts)
            verts.append(tmap([x / 2.0 for x in p]))
        return key[k]
    for c in range(8):
        o = (c & 1, (c >> 1) & 1, c >> 2)
        corner = {(a, b, d): vid((o[0] + a, o[1] + b, o[2] + d)) for a in (0, 1) for b in (0, 1) for d in (0, 1)}
        bnd = {f: (int(tb[f]) if o[f // 2] == f % 2 else -1) for f in range(6)}
        m = GU.ROT[rng.integers(len(GU.ROT))] if rotate else np.eye(3, dtype=int)
        nv_, nb_ = [], []
        for k in range(8):
            p = np.array([k & 1, (k >> 1) & 1, k >> 2]) - 0.5
            q = m @ p + 0.5
            nv_.append(corner[tuple(int(round(x)) for x in q)])
        for f in range(6):
            p = np.zeros(3)
            p[f // 2] = 0.5 if f % 2 else -0.5
            q = m @ p
            ax = int(np.argmax(np.abs(q)))
            nb_.append(bnd[2 * ax + (1 if q[ax] > 0 else 0)])
        t2v.append(nv_)
        tbnd.append(nb_)
    return np.array(verts), np.array(t2v, np.int32), np.array(tbnd, np.int32)


@pytest.mark.parametrize("affine", [False, True])
def test_children_trees_equal_parent_refined(affine):
    """The children carry the parent's corner labelling: the reference's face
    rule (two triangles split along the 1-2 diagonal, the normal from corners
    0, 1, 2) depends on the labelling, so a relabelled child would be a
    different discretisation of the same leaf.  A trilinear and an affine
    parent."""
    rng = np.random.default_rng(7)
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]], float)
    if affine:
        m = np.eye(3) + 0.2 * rng.uniform(-1, 1, (3, 3))
        tv = base @ m.T
    else:
        tv = base + 0.12 * rng.uniform(-1, 1, (8, 3))
    tb = np.arange(6, dtype=np.int32)
    cfg = OL.make_config(flux="HLLC", limiter=True, model="NS", bnd=BND6, bnd_param=PAR6)
    parent = OL.OracleSim(cfg, None, trees=(tv[None], tb[None], 1))
    v, t2v, tbnd = children_as_trees(tv, tb, 11, rotate=False)
    kids = OL.OracleSim(cfg, None, trees=(GU.tree_vertices(v, t2v), tbnd, 0))
    assert parent.n == kids.n == 8
    # parent leaf m (Morton) is child tree m
    cp = np.array([parent.geom(i)["c"] for i in range(8)])
    ck = np.array([kids.geom(i)["c"] for i in range(8)])
    np.testing.assert_allclose(ck, cp, rtol=0, atol=1e-15)
    u = random_state(8, 2)
    # the kids' leaves carry their own (rotated) frames only through geometry:
    # the Taylor basis lives in physical coordinates, so the same state applies
    for s in (parent, kids):
        s.set_state(u)
    ra, rb = parent.residual(), kids.residual()
    assert np.abs(ra - rb).max() <= 1e-13 * np.abs(ra).max()
    for s in (parent, kids):
        s.run(2e-3, 3)
    assert np.abs(parent.get_state() - kids.get_state()).max() <= 1e-12 * np.abs(parent.get_state()).max()


def test_gmesh_errors():
    v, t2v, tb = GU.box_trees(2, 1, 1, seed=1)
    bad = tb.copy()
    bad[bad == 1] = -1                   # a box side with no boundary and no partner
    with pytest.raises(CA.CharmError, match="no other tree"):
        CA.GMesh(v, t2v, bad, 0)
    t2 = t2v.copy()
    t2[0, 0] = len(v) + 5
    with pytest.raises(CA.CharmError, match="names vertex"):
        CA.GMesh(v, t2, tb, 0)
    flat = v.copy()
    flat[:, 2] = 0.0                     # zero-volume trees
    with pytest.raises(CA.CharmError, match="degenerate"):
        CA.GMesh(flat, t2v, tb, 0)
