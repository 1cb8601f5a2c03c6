"""Gmsh connectivity reader (charm_conn_read_msh, the reference's
charm_conn_reader_msh, src/charm_connectivity.c:315-519) on the CPU.

Meshes are written here: structured boxes of hexahedra with the six sides
tagged LEFT RIGHT FRONT BACK BOTTOM TOP, shuffled node ids, and every
hexahedron's node order turned by a random proper rotation of the cube.
Pinned by geometry, not by the reader's own tables:

* counts: trees, vertices, boundary faces per side, interior links
  (closed forms of the box);
* every link joins faces with the same centroid, the neighbour's centre on
  the outer side, and the link is symmetric;
* the orientation r: corner 0 of the lower face number meets corner r of the
  higher one (face corners taken from the coordinates bits here);
* trees stay right-handed after the Gmsh -> z-order corner mapping;
* boundary patches sit on their geometric side and bind to the task's
  boundary slots by name;
* error cases: missing file, binary format, a face of three hexahedra, a
  quadrangle on no face, a patch name the task lacks.
The reference's sample mesh (samples/klin_amr/klin.msh: 800 hexahedra of a
wedge channel) is read when the reference tree is present."""
import itertools
import os

import numpy as np
import pytest

import charm_3d_amd as CA

SIDES = ("LEFT", "RIGHT", "FRONT", "BACK", "BOTTOM", "TOP")
Z_TO_GMSH = (0, 1, 3, 2, 4, 5, 7, 6)          # z-order corner j is Gmsh node Z_TO_GMSH[j] (an involution)


def rotations():
    """the 24 proper rotations of the cube as signed permutation matrices"""
    out = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3), dtype=int)
            for i in range(3):
                m[i, perm[i]] = signs[i]
            if round(np.linalg.det(m)) == 1:
                out.append(m)
    return out


def zcorner(bits):
    return bits[0] + 2 * bits[1] + 4 * bits[2]


def write_box(path, n, rng, rotate=True, extra_lines=(), binary=False, dup_hex=False, stray_quad=False,
              split_x=None):
    """n = (nx, ny, nz) unit hexahedra; returns nothing, writes a Gmsh 2.2 file.
    Hexahedra are in the physical volume "BOX", or with split_x those at
    i >= split_x in a second one, "BOX2"."""
    nx, ny, nz = n
    pts = [(i, j, k) for k in range(nz + 1) for j in range(ny + 1) for i in range(nx + 1)]
    ids = rng.permutation(len(pts)) * 3 + 7                      # sparse, shuffled node ids
    nid = {p: int(ids[q]) for q, p in enumerate(pts)}
    rots = rotations()
    elems = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                # z-order corners of the cell, relabelled by a rotation about its centre
                corner = [None] * 8
                r = rots[rng.integers(len(rots))] if rotate else np.eye(3, dtype=int)
                for c in range(8):
                    b = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1])
                    rb = (r @ (2 * b - 1) + 1) // 2                 # rotated corner bits
                    corner[c] = nid[(i + rb[0], j + rb[1], k + rb[2])]
                gm = [corner[Z_TO_GMSH[g]] for g in range(8)]
                elems.append((5, 9 if split_x is not None and i >= split_x else 1, gm))
    if dup_hex:
        elems.append(elems[0])
    for side in range(6):
        a, s = side // 2, side % 2
        lim = (nx, ny, nz)
        rng_t = [range(lim[t]) for t in range(3) if t != a]
        for u in rng_t[0]:
            for v in rng_t[1]:
                q = []
                for du, dv in ((0, 0), (1, 0), (1, 1), (0, 1)):
                    p = [0, 0, 0]
                    p[a] = s * lim[a]
                    o = [t for t in range(3) if t != a]
                    p[o[0]], p[o[1]] = u + du, v + dv
                    q.append(nid[tuple(p)])
                elems.append((3, 2 + side, q))
    if stray_quad:
        elems.append((3, 2, [nid[(0, 0, 0)], nid[(1, 0, 0)], nid[(1, 1, 1)], nid[(0, 1, 1)]]))
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 %d 8\n$EndMeshFormat\n" % (1 if binary else 0))
        f.write("$PhysicalNames\n%d\n" % (7 if split_x is None else 8))
        for s, name in enumerate(SIDES):
            f.write('2 %d "%s"\n' % (2 + s, name))
        f.write('3 1 "BOX"\n')
        if split_x is not None:
            f.write('3 9 "BOX2"\n')
        f.write('$EndPhysicalNames\n')
        f.write("$Nodes\n%d\n" % len(pts))
        for p in pts:
            f.write("%d %g %g %g\n" % (nid[p], *p))
        f.write("$EndNodes\n$Elements\n%d\n" % len(elems))
        for e, (t, tag, nodes) in enumerate(elems):
            f.write("%d %d 2 %d %d %s\n" % (e + 1, t, tag, tag, " ".join(map(str, nodes))))
        f.write("$EndElements\n")
        for line in extra_lines:
            f.write(line + "\n")


def face_corners_from_bits(f):
    """corners of face f (axis f // 2, side f % 2) in increasing z-order"""
    a, s = f // 2, f % 2
    return [c for c in range(8) if (c >> a) & 1 == s]


def cube_task():
    t = CA.Task()
    for s, name in enumerate(SIDES):
        t.bnd_name[s].value = name.encode()
    t.n_regions = 1
    t.region[0].name = b"BOX"
    return t


@pytest.mark.parametrize("n, rotate", [((3, 2, 4), False), ((3, 3, 3), True), ((5, 1, 2), True)])
def test_box_counts_links_and_orientation(tmp_path, n, rotate):
    rng = np.random.default_rng(sum(n) + rotate)
    path = tmp_path / "box.msh"
    write_box(path, n, rng, rotate=rotate)
    c = CA.Connectivity(path)
    nx, ny, nz = n
    nt = nx * ny * nz
    assert c.num_trees == nt and c.num_vertices == (nx + 1) * (ny + 1) * (nz + 1)
    v = c.vertices[c.tree_to_vertex]                               # [nt][8][3]
    centre = v.mean(1)
    # right-handed z-order corners: det(c1 - c0, c2 - c0, c4 - c0) > 0
    jac = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 4] - v[:, 0]], axis=2)
    assert (np.linalg.det(jac) > 0.5).all()
    self_link = (c.tree_to_tree == np.arange(nt)[:, None]) & (c.tree_to_face == np.arange(6)[None, :])
    n_bnd = 2 * (nx * ny + ny * nz + nx * nz)
    assert self_link.sum() == n_bnd and (~self_link).sum() == 6 * nt - n_bnd
    fcb = [face_corners_from_bits(f) for f in range(6)]
    for t in range(nt):
        for f in range(6):
            fcent = v[t, fcb[f]].mean(0)
            outward = fcent - centre[t]
            if self_link[t, f]:
                # on the box surface, tagged with the side it lies on
                side = int(np.argmax(np.abs(outward)))
                side = 2 * side + (1 if outward[side] > 0 else 0)
                assert c.patches[[p[1] for p in c.patches].index(c.tree_to_patch[t, f])][2] == SIDES[side]
                continue
            nt_, code = c.tree_to_tree[t, f], int(c.tree_to_face[t, f])
            nf, r = code % 6, code // 6
            assert np.allclose(v[nt_, fcb[nf]].mean(0), fcent)           # same face
            assert (centre[nt_] - centre[t]) @ outward > 0                 # on the outer side
            assert c.tree_to_tree[nt_, nf] == t and int(c.tree_to_face[nt_, nf]) == f + 6 * r
            lo_t, lo_f, hi_t, hi_f = (t, f, nt_, nf) if f <= nf else (nt_, nf, t, f)
            assert np.allclose(v[lo_t, fcb[lo_f][0]], v[hi_t, fcb[hi_f][r]])
            assert c.tree_to_patch[t, f] == -1
    if rotate:
        assert len({int(x) // 6 for x in c.tree_to_face[~self_link]}) > 1    # orientations other than 0 occur
    tb, tr = c.bind_task(cube_task())
    for t in range(nt):
        for f in range(6):
            pid = c.tree_to_patch[t, f]
            want = -1 if pid < 0 else SIDES.index([p[2] for p in c.patches if p[1] == pid][0])
            assert tb[t, f] == want
    assert (tr == 0).all()


def test_error_cases(tmp_path):
    rng = np.random.default_rng(3)
    with pytest.raises(CA.CharmError, match="Failed to open"):
        CA.Connectivity(tmp_path / "absent.msh")
    write_box(tmp_path / "bin.msh", (1, 1, 1), rng, binary=True)
    with pytest.raises(CA.CharmError, match="binary"):
        CA.Connectivity(tmp_path / "bin.msh")
    write_box(tmp_path / "dup.msh", (2, 1, 1), rng, dup_hex=True)
    with pytest.raises(CA.CharmError, match="more than two"):
        CA.Connectivity(tmp_path / "dup.msh")
    write_box(tmp_path / "stray.msh", (2, 2, 2), rng, stray_quad=True)
    with pytest.raises(CA.CharmError, match="matches no hexahedron face"):
        CA.Connectivity(tmp_path / "stray.msh")
    write_box(tmp_path / "ok.msh", (1, 1, 2), rng)
    c = CA.Connectivity(tmp_path / "ok.msh")
    t = cube_task()
    t.bnd_name[5].value = b"LID"
    with pytest.raises(CA.CharmError, match=r"Patch \(id=7, name='TOP'\) is nod defined in task.yaml"):
        c.bind_task(t)


KLIN = "/root/reference/samples/klin_amr"


@pytest.mark.skipif(not os.path.exists(KLIN + "/klin.msh"), reason="reference sample absent")
def test_reference_sample_mesh():
    """samples/klin_amr/klin.geo: three transfinite blocks of 10x10, 10x10 and
    60x10 quadrangles extruded by one layer: 800 hexahedra; patches FRONT /
    BACK the two 800-quadrangle planes, TOP / BOTTOM 10 + 10 + 60 edges, LEFT
    / RIGHT 10 each; 1,510 shared faces."""
    c = CA.Connectivity(KLIN + "/klin.msh")
    assert c.num_trees == 800 and c.num_vertices == 1782
    self_link = (c.tree_to_tree == np.arange(800)[:, None]) & (c.tree_to_face == np.arange(6)[None, :])
    assert self_link.sum() == 1780 and (~self_link).sum() == 2 * 1510
    names = {p[1]: p[2] for p in c.patches}
    got = {}
    for pid in c.tree_to_patch[self_link]:
        got[names[int(pid)]] = got.get(names[int(pid)], 0) + 1
    assert got == {"FRONT": 800, "BACK": 800, "TOP": 80, "BOTTOM": 80, "LEFT": 10, "RIGHT": 10}
    assert set(c.tree_region.tolist()) == {1} and names[1] == "KLIN"
    v = c.vertices[c.tree_to_vertex]
    jac = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 4] - v[:, 0]], axis=2)
    assert (np.linalg.det(jac) > 0).all()
    task = CA.load_task(KLIN + "/task.yaml")
    tb, tr = c.bind_task(task)
    assert (tb[~self_link] == -1).all() and (tb[self_link] >= 0).all() and (tr == 0).all()
