"""Multi-tree test meshes (independent of the product's C code): a box of
nx x ny x nz hexahedral trees with its interior vertices displaced at random
(non-affine, non-planar faces), each tree's corners relabelled by a random
proper rotation of the reference cube (so neighbouring trees meet in every
orientation), and the box's six sides as boundary slots 0-5.

Arrays are those of a p4est-style connectivity: vertices [nv][3],
tree_to_vertex [nt][8] (z-order corners), tree_bnd [nt][6] (the boundary slot
of a tree face, -1 where it meets another tree)."""
import itertools

import numpy as np


def _rotations():
    """the 24 proper rotations of the cube as signed axis permutations
    (new axis i = sign[i] * old axis perm[i])"""
    out = []
    for perm in itertools.permutations(range(3)):
        for sign in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3), int)
            for i in range(3):
                m[i, perm[i]] = sign[i]
            if round(np.linalg.det(m)) == 1:
                out.append(m)
    return out


ROT = _rotations()


def box_trees(nx, ny, nz, distort=0.0, seed=0, rotate=True, size=(1.0, 1.0, 1.0)):
    rng = np.random.default_rng(seed)
    gx, gy, gz = (np.linspace(0.0, s, k + 1) for s, k in zip(size, (nx, ny, nz)))
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    vid = np.arange(X.size).reshape(X.shape)
    v = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    h = min(s / k for s, k in zip(size, (nx, ny, nz)))
    inner = np.zeros(X.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    v[inner.ravel()] += distort * h * rng.uniform(-1, 1, (inner.sum(), 3))
    # boundary vertices move inside their side's plane only (the sides stay planar)
    for ax, n in enumerate((nx, ny, nz)):
        for end in (0, n):
            sl = [slice(1, -1)] * 3
            sl[ax] = end
            ids = vid[tuple(sl)].ravel()
            for t in range(3):
                if t != ax:
                    v[ids, t] += distort * h * rng.uniform(-1, 1, len(ids))
    t2v, tbnd = [], []
    for i, j, k in itertools.product(range(nx), range(ny), range(nz)):
        corner = {}
        for c in range(8):
            a, b, d = c & 1, (c >> 1) & 1, c >> 2
            corner[(a, b, d)] = vid[i + a, j + b, k + d]
        bnd = {}
        for f in range(6):
            ax, side = f // 2, f % 2
            idx = (i, j, k)[ax] + side
            n = (nx, ny, nz)[ax]
            bnd[f] = f if idx in (0, n) else -1
        m = ROT[rng.integers(len(ROT))] if rotate else np.eye(3, dtype=int)
        # new local corner p (in {0,1}^3) is the old local corner q = R (p - c) + c, c = 1/2
        newv, newb = [], []
        for c in range(8):
            p = np.array([c & 1, (c >> 1) & 1, c >> 2]) - 0.5
            q = m @ p + 0.5
            newv.append(corner[tuple(int(round(x)) for x in q)])
        for f in range(6):
            p = np.zeros(3)
            p[f // 2] = 0.5 if f % 2 else -0.5
            q = m @ p
            ax = int(np.argmax(np.abs(q)))
            newb.append(bnd[2 * ax + (1 if q[ax] > 0 else 0)])
        t2v.append(newv)
        tbnd.append(newb)
    return v, np.array(t2v, np.int32), np.array(tbnd, np.int32)


def tree_vertices(v, t2v):
    """[nt][8][3] corner coordinates"""
    return np.ascontiguousarray(v[t2v], dtype=np.float64)


def leaf_centres(v, t2v, level):
    """physical centres of the leaves of the uniform multi-tree forest, in the
    order of the oracle and the product (trees, then Morton inside a tree)"""
    tv = tree_vertices(v, t2v)
    n = 1 << level
    out = []
    for t in range(len(t2v)):
        for m in range(n ** 3):
            c = [0, 0, 0]
            for b in range(level):
                for a in range(3):
                    c[a] |= ((m >> (3 * b + a)) & 1) << b
            xi = [(c[a] + 0.5) / n for a in range(3)]
            w = [[1 - x, x] for x in xi]
            p = np.zeros(3)
            for k in range(8):
                p += w[0][k & 1] * w[1][(k >> 1) & 1] * w[2][k >> 2] * tv[t, k]
            out.append(p)
    return np.array(out)
