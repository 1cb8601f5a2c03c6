"""CPU tests of libcharm3d.so: every symbol declared in include/charm_gpu.h is
exported, and the host forest/flattening agrees with the oracle's own face
traversal (content of the face list: sides, hanging sub-faces, boundaries)."""
import os
import re
import subprocess

import numpy as np
import pytest

import charm_3d_amd as CA
import mesh_util as MU
import oracle_lib as OL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    hdr = open(os.path.join(ROOT, "include", "charm_gpu.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set()
    for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(charm_\w+)\s*\(", hdr, flags=re.M):
        names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    CA.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", CA.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    decl = declared_functions()
    assert len(decl) > 40
    missing = [d for d in decl if d not in exported]
    assert not missing, missing


def oracle_records(faces):
    recs = set()
    for r in faces:
        ns, f0, f1, h0, h1 = (int(x) for x in r[:5])
        q0, q1 = r[5:9], r[9:13]
        bnd = int(r[13])
        if ns == 1:
            recs.add((int(q0[0]), -1 - bnd, f0 | (3 << 3)))
        elif h0:
            for j in range(4):
                recs.add((int(q0[j]), int(q1[0]), f0 | (1 << 3) | (j << 5)))
        elif h1:
            for j in range(4):
                recs.add((int(q0[0]), int(q1[j]), f0 | (2 << 3) | (j << 5)))
        else:
            recs.add((int(q0[0]), int(q1[0]), f0))
    return recs


def product_records(mesh):
    return set((int(a), int(b), int(c)) for a, b, c, _ in mesh.faces())


MESHES = [
    ("uniform3", lambda: MU.uniform_leaves(3), (False,) * 3),
    ("uniform3p", lambda: MU.uniform_leaves(3), (True,) * 3),
    ("adapt_a", lambda: MU.random_adapted(1, 3, seed=11), (False,) * 3),
    ("adapt_b", lambda: MU.random_adapted(1, 4, seed=12, frac=0.2, periodic=(True, False, True)),
     (True, False, True)),
    ("adapt_c", lambda: MU.random_adapted(2, 4, seed=13, frac=0.15, periodic=(True,) * 3), (True,) * 3),
]


@pytest.mark.parametrize("name,gen,periodic", MESHES, ids=[m[0] for m in MESHES])
def test_flatten_matches_oracle_traversal(name, gen, periodic):
    leaves = gen()
    f = CA.Forest.from_leaves(leaves, periodic)
    assert np.array_equal(f.leaves(), leaves)          # same Morton order
    mesh = f.flatten(128, 512)
    bnd = tuple("PERIODIC" if periodic[a // 2] else "OUTLET" for a in range(6))
    osim = OL.OracleSim(OL.make_config(bnd=bnd), leaves)
    assert product_records(mesh) == oracle_records(osim.faces())
    assert mesh.num_faces == len(product_records(mesh))


@pytest.mark.parametrize("level", [2, 3, 4])
def test_uniform_forest_and_counts(level):
    f = CA.Forest.uniform(level)
    assert np.array_equal(f.leaves(), MU.uniform_leaves(level))
    N = 1 << level
    faces = f.flatten().faces()
    kind = (faces[:, 2] >> 3) & 3
    assert (kind == 0).sum() == 3 * N * N * (N - 1)
    assert (kind == 3).sum() == 6 * N * N


def test_balance_matches_bruteforce():
    rng = np.random.default_rng(21)
    leaves = [tuple(int(v) for v in q) for q in MU.uniform_leaves(1)]
    for _ in range(3):                         # unbalanced random refinement
        nxt = []
        for q in leaves:
            if rng.random() < 0.35 and q[3] < 5:
                nxt.extend(MU._children(q))
            else:
                nxt.append(q)
        leaves = nxt
    raw = MU.sort_morton(np.array(leaves))
    f = CA.Forest.from_leaves(raw)
    f.balance()
    assert np.array_equal(f.leaves(), MU.balance(raw))


def flags_applied(leaves, flags, max_level, min_level):
    """refine (flag > 0 below max_level), then coarsen every complete family
    of eight leaves flagged < 0 above min_level (charm_forest_adapt's rules,
    src/charm_amr.c:601-606), by sets"""
    fl = {tuple(int(v) for v in q): int(g) for q, g in zip(leaves, flags)}

    def parent(q):
        h = MU.ROOT >> (q[3] - 1)
        return (q[0] // h * h, q[1] // h * h, q[2] // h * h, q[3] - 1)

    fams = {}
    for q in fl:
        if q[3] > 0:
            fams.setdefault(parent(q), []).append(q)
    gone = {p for p, kids in fams.items()
            if len(kids) == 8 and kids[0][3] > min_level and all(fl[k] < 0 for k in kids)}
    out = list(gone)
    for q, g in fl.items():
        if g > 0 and q[3] < max_level:
            out.extend(MU._children(q))
        elif not (q[3] > 0 and parent(q) in gone):
            out.append(q)
    return MU.sort_morton(np.array(out))


@pytest.mark.parametrize("periodic", [False, True])
def test_adapt_rounds_match_bruteforce_balance(periodic):
    """Rounds of mixed random refine and coarsen flags on a forest that the
    adaptation keeps balanced (its balance passes visit only the leaves each
    rebuild made) give the leaves of flags -> brute-force balance, and the
    same leaves and old -> new map as the same flags on a forest rebuilt from
    the same leaves (whose first balance pass visits every leaf)"""
    rng = np.random.default_rng(7 + periodic)
    per = (periodic,) * 3
    f = CA.Forest.uniform(2, per)
    coarsened = refined = 0
    for _ in range(6):
        lv = f.leaves()
        flags = rng.choice(np.array([-1, 0, 1], dtype=np.int8), size=len(lv), p=[0.55, 0.3, 0.15])
        slow = CA.Forest.from_leaves(lv, per)
        m_fast = f.adapt(flags, 5, 1)
        m_slow = slow.adapt(flags, 5, 1)
        got = f.leaves()
        np.testing.assert_array_equal(got, slow.leaves())
        np.testing.assert_array_equal(m_fast, m_slow)
        np.testing.assert_array_equal(got, MU.balance(flags_applied(lv, flags, 5, 1), per))
        coarsened += int((m_fast < 0).sum())
        refined += int(len(got) > len(lv))
    assert coarsened > 0 and refined > 0


def test_neighbor_table_covers_every_face():
    leaves = MU.random_adapted(1, 3, seed=31)
    mesh = CA.Forest.from_leaves(leaves).flatten()
    nbr = mesh.neighbors()
    assert (nbr != np.iinfo(np.int32).min).all()
    c, h = MU.centers(leaves)
    for e in range(len(leaves)):
        for d in range(6):
            v = nbr[e, d]
            a, s = d // 2, (1 if d % 2 else -1)
            if -8 < v < 0:
                assert v == -1 - d
                edge = c[e, a] + s * 0.5 * h[e]
                assert abs(edge - (1.0 if s > 0 else 0.0)) < 1e-12
            elif v >= 0:
                # the neighbour touches the face plane
                assert abs((c[v, a] - s * 0.5 * h[v]) - (c[e, a] + s * 0.5 * h[e])) < 1e-12


def test_adapt_refine_coarsen_and_map():
    f = CA.Forest.uniform(2)
    flags = np.zeros(f.n, dtype=np.int8)
    flags[0] = 1
    flags[8:16] = -1                     # one whole family
    mp = f.adapt(flags, max_level=4, min_level=0)
    lv = f.leaves()
    assert len(mp) == len(lv)
    assert (lv[:, 3] == 3).sum() == 8 and (lv[:, 3] == 1).sum() == 1
    assert (mp[:8] == 0).all() and (mp[8:15] == np.arange(1, 8)).all() and mp[15] == -1 - 8
    # the result is 2:1 balanced and tiles the cube
    assert np.array_equal(lv, MU.balance(lv))


def test_errors_are_reported_not_fatal():
    with pytest.raises(CA.CharmError, match="tile|cover|aligned"):
        CA.Forest.from_leaves(np.array([[0, 0, 0, 1]], dtype=np.int32))
    unbalanced = MU.sort_morton(np.array(
        [q for q in MU.uniform_leaves(1).tolist() if q[:3] != [0, 0, 0]] +
        [q for q in MU.uniform_leaves(3).tolist() if max(q[:3]) < (1 << 18)]))
    f = CA.Forest.from_leaves(unbalanced)
    with pytest.raises(CA.CharmError, match="balanced"):
        f.flatten()


@pytest.mark.parametrize("level", [0, 3, 7, 12])
def test_level_mass_inverse_matches_oracle(level):
    a, ai = CA.level_mass_inverse(level)
    h = 2.0 ** -level
    assert np.allclose(a @ ai, np.eye(4), atol=1e-14)
    assert np.allclose(ai, np.diag([1, 12, 12, 12]) / h ** 3, rtol=1e-14, atol=1e-14 / h ** 3)
    # power-of-two scaling of the level-0 inverse is exact
    assert np.array_equal(ai, CA.level_mass_inverse(0)[1] * 8.0 ** level)
    if level <= 3:
        leaves = MU.uniform_leaves(level)
        g = OL.OracleSim(OL.make_config(), leaves).geom(len(leaves) // 2)
        assert np.allclose(g["a_inv"], ai, rtol=1e-13, atol=1e-13 * ai.max())


@pytest.mark.parametrize("eta", [-1.0, float("nan"), float("inf")])
def test_bad_br2_penalty_factor_is_rejected(eta):
    """the context is refused before any device call (CPU)"""
    cfg = CA.make_config(model="NS", use_visc=True, visc_br2_eta=eta)
    with pytest.raises(CA.CharmError, match="BR2 penalty factor"):
        CA.GpuContext(cfg)


def test_chunk_faces_below_one_leaf_is_an_error():
    """a chunk must hold every face of a leaf (six on a uniform mesh)"""
    with pytest.raises(CA.CharmError, match="chunk_faces"):
        CA.Forest.uniform(2).flatten(1, 5)


def check_chunks(f, ce, cf):
    """every invariant the stage kernel relies on: chunks partition the
    leaves in order with at most ce leaves and cf distinct faces, a chunk
    holds exactly the faces incident to its leaves, sorted by (axis, kind),
    and every leaf's slots decode to its own incidences (face, side)"""
    m = f.flatten(ce, cf)
    faces = m.faces()
    eb, fb, cf4, sp, sl = m.chunks()
    n = f.n
    assert eb[0] == 0 and eb[-1] == n and (np.diff(eb) > 0).all() and (np.diff(eb) <= ce).all()
    assert fb[0] == 0 and fb[-1] == len(cf4) and (np.diff(fb) <= cf).all()
    rows = {tuple(r): i for i, r in enumerate(faces.tolist())}
    assert len(rows) == len(faces)
    inc = [[] for _ in range(n)]
    for i, (e0, e1, _, _) in enumerate(faces.tolist()):
        inc[e0].append((i, 0))
        if e1 >= 0:
            inc[e1].append((i, 1))
    assert sp[-1] == sum(len(x) for x in inc)
    for c in range(len(eb) - 1):
        recs = cf4[fb[c]:fb[c + 1]]
        ids = [rows[tuple(r)] for r in recs.tolist()]
        assert len(set(ids)) == len(ids)
        key = [((r[2] & 7) >> 1) * 4 + ((r[2] >> 3) & 3) for r in recs.tolist()]
        assert key == sorted(key)
        want = set()
        for e in range(eb[c], eb[c + 1]):
            want |= {i for i, _ in inc[e]}
            got = [(ids[s >> 1], s & 1) for s in sl[sp[e]:sp[e + 1]].tolist()]
            assert sorted(got) == sorted(inc[e]), (c, e)
        assert set(ids) == want
    return m


@pytest.mark.parametrize("case", ["uniform4", "adapted", "adapted_periodic", "single_periodic"])
def test_chunks_hold_their_leaves_faces(case):
    if case == "uniform4":
        f = CA.Forest.uniform(4)
    elif case == "adapted":
        f = CA.Forest.from_leaves(MU.random_adapted(2, 5, seed=31, frac=0.3))
    elif case == "adapted_periodic":
        per = (True, False, True)
        f = CA.Forest.from_leaves(MU.random_adapted(2, 4, seed=37, frac=0.25, periodic=per), per)
    else:
        f = CA.Forest.uniform(0, (True, True, True))
    for ce, cf in ((1, 24), (7, 40), (128, 512), (256, 1024)):
        check_chunks(f, ce, cf)


def test_chunk_layout_spans_blocks():
    """a level-6 forest (262,144 leaves: eight independently chunked blocks
    of 32,768) gives the same face-copy count whatever the thread count, and
    no chunk crosses a block boundary"""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import charm_3d_amd as CA; "
            "m = CA.Forest.uniform(6).flatten(213, 640); eb = m.chunks()[0]; "
            "print(m.num_chunks, m.num_chunk_faces, int(((eb[:-1] // 32768) != ((eb[1:] - 1) // 32768)).sum()))"
            % os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    outs = set()
    for th in ("1", "3", "8"):
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                           env=dict(os.environ, CHARM_HOST_THREADS=th))
        assert r.returncode == 0, r.stderr
        outs.add(r.stdout.strip())
    assert len(outs) == 1
    nch, ncf, crossing = (int(x) for x in outs.pop().split())
    assert crossing == 0 and ncf > CA.Forest.uniform(6).flatten(213, 640).num_faces
