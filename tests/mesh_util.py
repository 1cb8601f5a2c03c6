"""Small octree helpers for tests (independent of the product's C forest).

Leaves are int32 rows (x, y, z, level) in root units (root length 2^19),
sorted in Morton order of their anchors.
"""
import numpy as np

MAXLEVEL = 19
ROOT = 1 << MAXLEVEL


def morton_key(x, y, z):
    x = np.asarray(x, dtype=np.uint64)
    y = np.asarray(y, dtype=np.uint64)
    z = np.asarray(z, dtype=np.uint64)
    key = np.zeros_like(x)
    for b in range(MAXLEVEL):
        key |= ((x >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
        key |= ((y >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + 1)
        key |= ((z >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + 2)
    return key


def sort_morton(leaves):
    leaves = np.asarray(leaves, dtype=np.int64)
    order = np.argsort(morton_key(leaves[:, 0], leaves[:, 1], leaves[:, 2]), kind="stable")
    return np.ascontiguousarray(leaves[order], dtype=np.int32)


def uniform_leaves(level):
    n = 1 << level
    h = ROOT >> level
    i = np.arange(n)
    x, y, z = np.meshgrid(i, i, i, indexing="ij")
    lv = np.stack([x.ravel() * h, y.ravel() * h, z.ravel() * h,
                   np.full(n ** 3, level)], axis=1)
    return sort_morton(lv)


def _children(q):
    x, y, z, l = q
    h = ROOT >> (l + 1)
    return [(x + (c & 1) * h, y + ((c >> 1) & 1) * h, z + ((c >> 2) & 1) * h, l + 1)
            for c in range(8)]


def _leaf_at(leafset, x, y, z, maxl):
    """Leaf containing point (x, y, z) (integer root coords), or None."""
    for l in range(maxl, -1, -1):
        h = ROOT >> l
        key = (x // h * h, y // h * h, z // h * h, l)
        if key in leafset:
            return key
    return None


def balance(leaves, periodic=(False, False, False)):
    """Brute-force 2:1 face balance: refine any leaf with a face neighbour
    more than one level finer, until nothing changes."""
    leafset = set(tuple(int(v) for v in q) for q in leaves)
    maxl = max(q[3] for q in leafset)
    changed = True
    while changed:
        changed = False
        for q in list(leafset):
            if q not in leafset:
                continue
            x, y, z, l = q
            h = ROOT >> l
            need = False
            for a in range(3):
                for s in (-1, 1):
                    # sample points just outside the face at the finest level
                    fine = ROOT >> maxl
                    for t1 in range(0, h, fine):
                        for t2 in range(0, h, fine):
                            p = [x, y, z]
                            p[a] = p[a] + h if s > 0 else p[a] - fine
                            ta, tb = [b for b in range(3) if b != a]
                            p[ta] += t1
                            p[tb] += t2
                            if p[a] < 0 or p[a] >= ROOT:
                                if not periodic[a]:
                                    continue
                                p[a] %= ROOT
                            nb = _leaf_at(leafset, p[0], p[1], p[2], maxl)
                            if nb is not None and nb[3] > l + 1:
                                need = True
                                break
                        if need:
                            break
                    if need:
                        break
                if need:
                    break
            if need:
                leafset.remove(q)
                for c in _children(q):
                    leafset.add(c)
                changed = True
    return sort_morton(np.array(sorted(leafset), dtype=np.int64))


def random_adapted(base_level, max_level, seed, frac=0.3, periodic=(False, False, False)):
    """Uniform base mesh, random leaves refined up to max_level, then
    2:1 balanced."""
    rng = np.random.default_rng(seed)
    leaves = [tuple(int(v) for v in q) for q in uniform_leaves(base_level)]
    for _ in range(max_level - base_level):
        nxt = []
        for q in leaves:
            if q[3] < max_level and rng.random() < frac:
                nxt.extend(_children(q))
            else:
                nxt.append(q)
        leaves = nxt
    return balance(np.array(leaves, dtype=np.int64), periodic)


def centers(leaves):
    leaves = np.asarray(leaves, dtype=np.float64)
    h = (ROOT >> leaves[:, 3].astype(np.int64)) / ROOT
    c = leaves[:, :3] / ROOT + 0.5 * h[:, None]
    return c, h


S3 = 1.0 / np.sqrt(3.0)
QGP_REF = np.array([[-S3, -S3, -S3], [S3, -S3, -S3], [-S3, S3, -S3], [S3, S3, -S3],
                    [-S3, -S3, S3], [S3, -S3, S3], [-S3, S3, S3], [S3, S3, S3]])


def project_p1(leaves, fn, nv):
    """L2 projection onto the P1 Taylor basis with the 8-point Gauss rule
    (mass matrix diag(V, V/12, V/12, V/12)).  fn(x[npts,3]) -> [npts, nv]."""
    c, h = centers(leaves)
    n = len(leaves)
    u = np.zeros((n, nv, 4))
    acc = np.zeros((n, nv, 4))
    for g in range(8):
        x = c + 0.5 * h[:, None] * QGP_REF[g][None, :]
        f = fn(x)
        phi = np.concatenate([np.ones((n, 1)), 0.5 * np.broadcast_to(QGP_REF[g], (n, 3))], axis=1)
        acc += f[:, :, None] * phi[:, None, :]
    # sum_g f phi_k wJ / (A_kk) with wJ = V/8, A = V diag(1, 1/12, 1/12, 1/12)
    u[:, :, 0] = acc[:, :, 0] / 8.0
    u[:, :, 1:] = acc[:, :, 1:] * 12.0 / 8.0
    return u
