/* charm_gmesh.c -- multi-tree forests refined uniformly, with the per-leaf
 * trilinear geometry of the reference, FP64 on the host.
 *
 * The reference builds a p4est forest over a Gmsh connectivity at MIN_LEVEL
 * (src/charm_3d.c:36-37) and computes every leaf's geometry from the
 * trilinear map of its tree's 8 corners (src/charm_geom.c:12-172, 496-530):
 * vertices and centre through p4est_qcoord_to_vertex, 8 Gauss points at
 * +-1/sqrt(3) of the reference cube with the Jacobian determinant, one normal
 * per face from the face's first three corners, 6 face points on the two
 * triangles (v0 v1 v2), (v1 v2 v3) with weight 1/6 and twice the triangle's
 * area, dh = the bounding box, the Taylor-basis mass matrix by the volume
 * rule and its inverse.  Leaves are numbered tree by tree, Morton order
 * inside a tree (p4est's global order).
 *
 * Faces: inside a tree by index arithmetic; a tree face with a boundary slot
 * is a boundary face; every other tree face meets another tree.  Two leaf
 * faces on a shared tree face are matched exactly, without coordinates: the
 * tree face is named by its corners' global vertex ids in a canonical order
 * (the smallest id, then its two face neighbours by id), and a leaf face by
 * the integer lattice position of its centre in that canonical frame.  Side 0
 * of a face is the leaf with the smaller index.
 *
 * For the device the geometry is reduced to what the kernels read: at every
 * volume point, face point and vertex the values of the basis (x - c) / dh
 * (computed in FP64 relative to the leaf's own centre, so FP32 keeps them to
 * its precision), the quadrature weights times Jacobians, the inverse mass
 * matrix, 1 / dh, the side-0 normal of every face, and the mean of each basis
 * function over the leaf.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "charm_gpu.h"
#include "charm_internal.h"

#define GM_MAXLEVEL 19
#define GM_ROOT ((int64_t) 1 << GM_MAXLEVEL)
#define GM_FACE_BLOCK 8192   /* side-0 leaves per face group block */

/* corners of each face in z-order (src/charm_geom.c:182-188) */
static const int gm_ftv[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};

struct charm_gmesh {
    int64_t n, nf, ntree;
    int level;
    int32_t *leaf;      /* [n][5] tree, x, y, z, level (root units 2^19) */
    double *geo;        /* [n][23] c, dh, volume, a_inv */
    double *lg;         /* [CHARM_GM_LEAFQ][n] */
    double *fg;         /* [CHARM_GM_FACEQ][nf] */
    double *fn;         /* [3][nf] side-0 outward normals */
    int32_t *fe;        /* [2][nf] side-0 leaf, side-1 leaf or -1 - boundary slot */
    int32_t *ff;        /* [2][nf] local faces (side 1: -1 on a boundary) */
    int32_t *lnb;       /* [6][n] neighbour leaf, or -1 - face index of a boundary face */
    double (*tv)[8][3]; /* [ntree] tree corners (z-order) */
    int32_t *treg;      /* [ntree] region index of each tree (a task's), or NULL */
};

typedef struct {
    double v[8][3], c[3], dh[3], vol;
    double n[6][3], fc[6][3], fgp[6][6][3], fwj[6][6];
    double qgp[8][3], qwj[8];
    double a[4][4], ainv[4][4];
} gm_geom_t;

/* p4est_qcoord_to_vertex over a tree's corners: trilinear weights
 * w1 = q / root, w0 = 1 - w1, summed z, y, x outer to inner */
static void gm_tree_point(const double tv[8][3], int64_t qx, int64_t qy, int64_t qz, double out[3])
{
    double wx[2], wy[2], wz[2];
    int xi, yi, zi, k = 0;
    wx[1] = (double) qx / (double) GM_ROOT; wx[0] = 1.0 - wx[1];
    wy[1] = (double) qy / (double) GM_ROOT; wy[0] = 1.0 - wy[1];
    wz[1] = (double) qz / (double) GM_ROOT; wz[0] = 1.0 - wz[1];
    out[0] = out[1] = out[2] = 0.0;
    for (zi = 0; zi < 2; zi++)
        for (yi = 0; yi < 2; yi++) {
            const double yf = wz[zi] * wy[yi];
            for (xi = 0; xi < 2; xi++, k++) {
                const double xf = yf * wx[xi];
                out[0] += xf * tv[k][0];
                out[1] += xf * tv[k][1];
                out[2] += xf * tv[k][2];
            }
        }
}

static double gm_basis(const gm_geom_t *g, const double x[3], int k)
{
    return k == 0 ? 1.0 : (x[k - 1] - g->c[k - 1]) / g->dh[k - 1];
}

/* src/charm_geom.c:40-128: the point of reference coordinates r in [-1,1]^3
 * and the Jacobian determinant there */
static void gm_volume_point(const double v[8][3], const double r[3], double gp[3], double *gj)
{
    double jac[3][3] = {{0}}, w[3][2], d[2] = {-1.0, 1.0};
    int a, k;
    for (a = 0; a < 3; a++) { w[a][0] = 1.0 - r[a]; w[a][1] = 1.0 + r[a]; }
    gp[0] = gp[1] = gp[2] = 0.0;
    for (k = 0; k < 8; k++) {
        const int ix = k & 1, iy = (k >> 1) & 1, iz = k >> 2;
        const double f = w[2][iz] * w[1][iy] * w[0][ix] / 8.0;
        const double fx = w[2][iz] * w[1][iy] * d[ix] / 8.0;
        const double fy = w[2][iz] * d[iy] * w[0][ix] / 8.0;
        const double fz = d[iz] * w[1][iy] * w[0][ix] / 8.0;
        for (a = 0; a < 3; a++) {
            gp[a] += f * v[k][a];
            jac[a][0] += fx * v[k][a];
            jac[a][1] += fy * v[k][a];
            jac[a][2] += fz * v[k][a];
        }
    }
    *gj = jac[0][0] * (jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1])
        - jac[0][1] * (jac[1][0] * jac[2][2] - jac[1][2] * jac[2][0])
        + jac[0][2] * (jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0]);
}

static void gm_geom(const double tv[8][3], const int32_t q[4], gm_geom_t *g)
{
    const int64_t l = GM_ROOT >> q[3], l2 = l / 2;
    const double s3 = 1.0 / sqrt(3.0);
    int f, i, j, k, a;
    for (k = 0; k < 8; k++)
        gm_tree_point(tv, q[0] + (k & 1) * l, q[1] + ((k >> 1) & 1) * l, q[2] + (k >> 2) * l, g->v[k]);
    gm_tree_point(tv, q[0] + l2, q[1] + l2, q[2] + l2, g->c);
    for (a = 0; a < 3; a++) {
        double lo = g->v[0][a], hi = g->v[0][a];
        for (k = 1; k < 8; k++) { lo = g->v[k][a] < lo ? g->v[k][a] : lo; hi = g->v[k][a] > hi ? g->v[k][a] : hi; }
        g->dh[a] = hi - lo;
    }
    for (f = 0; f < 6; f++) {
        const int ax = f / 2;
        int64_t o[3] = {l2, l2, l2};
        double x[4][3], e0[3], e1[3], nl;
        o[ax] = (f & 1) ? l : 0;
        gm_tree_point(tv, q[0] + o[0], q[1] + o[1], q[2] + o[2], g->fc[f]);
        for (i = 0; i < 4; i++) memcpy(x[i], g->v[gm_ftv[f][i]], sizeof(x[i]));
        for (a = 0; a < 3; a++) { e0[a] = x[1][a] - x[0][a]; e1[a] = x[2][a] - x[0][a]; }
        g->n[f][0] = e0[1] * e1[2] - e0[2] * e1[1];
        g->n[f][1] = e0[2] * e1[0] - e0[0] * e1[2];
        g->n[f][2] = e0[0] * e1[1] - e0[1] * e1[0];
        nl = sqrt(g->n[f][0] * g->n[f][0] + g->n[f][1] * g->n[f][1] + g->n[f][2] * g->n[f][2]);
        for (a = 0; a < 3; a++) g->n[f][a] /= nl;
        /* two triangles, three points each at (2/3, 1/6, 1/6) of (p0 - p2, p1 - p2)
         * (src/charm_geom.c:443-494) */
        for (i = 0; i < 2; i++) {
            const double *p0 = x[i], *p1 = x[i + 1], *p2 = x[i + 2];
            double d1[3], d2[3], E = 0.0, F = 0.0, G = 0.0, gj;
            const double wa[3][2] = {{2.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}, {1.0 / 6.0, 1.0 / 6.0}};
            for (a = 0; a < 3; a++) {
                d1[a] = p0[a] - p2[a]; d2[a] = p1[a] - p2[a];
                E += d1[a] * d1[a]; F += d2[a] * d2[a]; G += d1[a] * d2[a];
            }
            gj = sqrt(E * F - G * G);
            for (j = 0; j < 3; j++) {
                for (a = 0; a < 3; a++) g->fgp[f][3 * i + j][a] = wa[j][0] * d1[a] + wa[j][1] * d2[a] + p2[a];
                g->fwj[f][3 * i + j] = gj / 6.0;
            }
        }
    }
    g->vol = 0.0;
    for (k = 0; k < 8; k++) {
        const double r[3] = {(k & 1) ? s3 : -s3, (k & 2) ? s3 : -s3, (k & 4) ? s3 : -s3};
        gm_volume_point((const double (*)[3]) g->v, r, g->qgp[k], &g->qwj[k]);
        g->vol += g->qwj[k];
    }
    for (i = 0; i < 4; i++)
        for (j = 0; j < 4; j++) {
            g->a[i][j] = 0.0;
            for (k = 0; k < 8; k++) g->a[i][j] += g->qwj[k] * gm_basis(g, g->qgp[k], i) * gm_basis(g, g->qgp[k], j);
        }
}

/* side-0 outward normal of face f: the face normal flipped to point away
 * from the leaf's centre (src/models/euler/charm_model.c:332-344) */
static void gm_outward(const gm_geom_t *g, int f, double n[3])
{
    double s = 0.0;
    int a;
    for (a = 0; a < 3; a++) { n[a] = g->n[f][a]; s += n[a] * (g->fc[f][a] - g->c[a]); }
    if (s < 0.0) for (a = 0; a < 3; a++) n[a] = -n[a];
}

typedef struct { int64_t g0, g1, g2; int32_t s, t; int64_t leaf; int face; } gm_key_t;

static int gm_key_cmp(const void *pa, const void *pb)
{
    const gm_key_t *a = (const gm_key_t *) pa, *b = (const gm_key_t *) pb;
    if (a->g0 != b->g0) return a->g0 < b->g0 ? -1 : 1;
    if (a->g1 != b->g1) return a->g1 < b->g1 ? -1 : 1;
    if (a->g2 != b->g2) return a->g2 < b->g2 ? -1 : 1;
    if (a->s != b->s) return a->s < b->s ? -1 : 1;
    if (a->t != b->t) return a->t < b->t ? -1 : 1;
    if (a->leaf != b->leaf) return a->leaf < b->leaf ? -1 : 1;
    return a->face - b->face;
}

static int64_t gm_morton(int32_t x, int32_t y, int32_t z, int level)
{
    int64_t m = 0;
    int b;
    for (b = 0; b < level; b++)
        m |= ((int64_t) ((x >> b) & 1) << (3 * b)) | ((int64_t) ((y >> b) & 1) << (3 * b + 1)) |
             ((int64_t) ((z >> b) & 1) << (3 * b + 2));
    return m;
}

void charm_gmesh_destroy(charm_gmesh_t *m)
{
    if (!m) return;
    free(m->leaf); free(m->geo); free(m->lg); free(m->fg); free(m->fn); free(m->fe); free(m->ff);
    free(m->lnb); free(m->tv); free(m->treg);
    free(m);
}

int charm_gmesh_build(int64_t n_vertices, const double *vertices, int64_t n_trees, const int32_t *tree_to_vertex,
                      const int32_t *tree_bnd, int level, charm_gmesh_t **out)
{
    charm_gmesh_t *m;
    int64_t e, t, nper, side_n, nkey = 0, i, nf = 0;
    int64_t *pe = NULL;            /* [n][6]: partner leaf (>= 0), or -1 - slot */
    int64_t *next = NULL;          /* face group offsets */
    int8_t *pf = NULL;             /* [n][6]: partner face */
    gm_key_t *keys = NULL;
    double (*tvs)[8][3] = NULL;
    if (!out) { charm_set_error("gmesh_build: null output"); return CHARM_ERR_ARG; }
    *out = NULL;
    if (!vertices || !tree_to_vertex || !tree_bnd || n_vertices < 1 || n_trees < 1 || level < 0 || level > 6) {
        charm_set_error("gmesh_build: bad arguments (level 0..6, at least one tree)");
        return CHARM_ERR_ARG;
    }
    nper = (int64_t) 1 << (3 * level);
    side_n = (int64_t) 1 << level;
    if (n_trees * nper > ((int64_t) 1 << 30)) { charm_set_error("gmesh_build: too many leaves"); return CHARM_ERR_ARG; }
    for (i = 0; i < 8 * n_trees; i++)
        if (tree_to_vertex[i] < 0 || tree_to_vertex[i] >= n_vertices) {
            charm_set_error("gmesh_build: tree %lld names vertex %d of %lld", (long long) (i / 8),
                            tree_to_vertex[i], (long long) n_vertices);
            return CHARM_ERR_ARG;
        }
    for (i = 0; i < 6 * n_trees; i++)
        if (tree_bnd[i] < -1 || tree_bnd[i] > 5) {
            charm_set_error("gmesh_build: boundary slot %d (0..5, or -1 for a face between trees)", tree_bnd[i]);
            return CHARM_ERR_ARG;
        }
    m = (charm_gmesh_t *) calloc(1, sizeof(*m));
    if (!m) { charm_set_error("gmesh_build: out of memory"); return CHARM_ERR_MEM; }
    m->ntree = n_trees; m->level = level; m->n = n_trees * nper;
    m->leaf = (int32_t *) malloc((size_t) m->n * 5 * sizeof(int32_t));
    m->geo = (double *) malloc((size_t) m->n * 23 * sizeof(double));
    m->lg = (double *) malloc((size_t) m->n * CHARM_GM_LEAFQ * sizeof(double));
    m->lnb = (int32_t *) malloc((size_t) m->n * 6 * sizeof(int32_t));
    pe = (int64_t *) malloc((size_t) m->n * 6 * sizeof(int64_t));
    pf = (int8_t *) malloc((size_t) m->n * 6);
    keys = (gm_key_t *) malloc((size_t) (m->n * 6 + 1) * sizeof(gm_key_t));
    tvs = (double (*)[8][3]) malloc((size_t) n_trees * sizeof(*tvs));
    if (!m->leaf || !m->geo || !m->lg || !m->lnb || !pe || !pf || !keys || !tvs) goto oom;
    for (t = 0; t < n_trees; t++)
        for (i = 0; i < 8; i++) memcpy(tvs[t][i], vertices + 3 * (size_t) tree_to_vertex[8 * t + i], 3 * sizeof(double));

    /* leaves, compact geometry and the device leaf block */
    for (e = 0; e < m->n; e++) {
        const int64_t mm = e % nper, len = GM_ROOT >> level;
        int32_t c[3] = {0, 0, 0}, q[4];
        int b, a, k, j;
        gm_geom_t g;
        double *geo = m->geo + 23 * (size_t) e;
        t = e / nper;
        for (b = 0; b < level; b++)
            for (a = 0; a < 3; a++) c[a] |= (int32_t) (((mm >> (3 * b + a)) & 1) << b);
        q[0] = (int32_t) (c[0] * len); q[1] = (int32_t) (c[1] * len); q[2] = (int32_t) (c[2] * len); q[3] = level;
        m->leaf[5 * e] = (int32_t) t;
        memcpy(m->leaf + 5 * e + 1, q, sizeof(q));
        gm_geom((const double (*)[3]) tvs[t], q, &g);
        if (!(g.vol > 0.0) || !charm_gj_inverse4((const double (*)[4]) g.a, g.ainv)) {
            charm_set_error("gmesh_build: tree %lld has a degenerate or inverted leaf (volume %g)", (long long) t, g.vol);
            goto fail;
        }
        memcpy(geo, g.c, 3 * sizeof(double));
        memcpy(geo + 3, g.dh, 3 * sizeof(double));
        geo[6] = g.vol;
        memcpy(geo + 7, g.ainv, 16 * sizeof(double));
        /* device block: 1/dh, volume-point basis values, weights, A^-1,
         * vertex basis values, mean basis values */
        {
            double *L = m->lg;
            const int64_t n = m->n;
            double mb[3] = {0.0, 0.0, 0.0};
            for (a = 0; a < 3; a++) L[(size_t) a * n + e] = 1.0 / g.dh[a];
            for (k = 0; k < 8; k++) {
                for (a = 0; a < 3; a++) {
                    const double b3 = (g.qgp[k][a] - g.c[a]) / g.dh[a];
                    L[(size_t) (3 + 3 * k + a) * n + e] = b3;
                    mb[a] += b3 * g.qwj[k];
                }
                L[(size_t) (27 + k) * n + e] = g.qwj[k];
            }
            for (k = 0; k < 4; k++)
                for (j = 0; j < 4; j++) L[(size_t) (35 + 4 * k + j) * n + e] = g.ainv[k][j];
            for (k = 0; k < 8; k++)
                for (a = 0; a < 3; a++) L[(size_t) (51 + 3 * k + a) * n + e] = (g.v[k][a] - g.c[a]) / g.dh[a];
            for (a = 0; a < 3; a++) L[(size_t) (75 + a) * n + e] = mb[a] / g.vol;
        }
        /* partners */
        for (int f = 0; f < 6; f++) {
            const int ax = f / 2, hi = f & 1;
            const int64_t slot = 6 * e + f;
            if (hi ? c[ax] + 1 < side_n : c[ax] > 0) {
                int32_t p[3] = {c[0], c[1], c[2]};
                p[ax] += hi ? 1 : -1;
                pe[slot] = t * nper + gm_morton(p[0], p[1], p[2], level);
                pf[slot] = (int8_t) (f ^ 1);
            } else if (tree_bnd[6 * t + f] >= 0) {
                pe[slot] = -1 - tree_bnd[6 * t + f];
                pf[slot] = -1;
            } else {
                /* canonical frame of the tree face: corner i sits at lattice
                 * ((i & 1) 2N, (i >> 1) 2N); the leaf face's centre at (u, v) */
                const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
                const int64_t twoN = 2 * side_n, u = 2 * c[t1] + 1, v = 2 * c[t2] + 1;
                int64_t gid[4];
                int o = 0, A, B;
                for (i = 0; i < 4; i++) {
                    gid[i] = tree_to_vertex[8 * t + gm_ftv[f][i]];
                    if (gid[i] < gid[o]) o = (int) i;
                }
                A = o ^ 1; B = o ^ 2;
                {
                    const int64_t du = (o & 1) ? twoN - u : u, dv = (o & 2) ? twoN - v : v;
                    gm_key_t *kk = &keys[nkey++];
                    if (gid[B] < gid[A]) { int s = A; A = B; B = s; }
                    kk->g0 = gid[o]; kk->g1 = gid[A]; kk->g2 = gid[B];
                    /* A along the tree face's first axis (o ^ 1) or its second */
                    kk->s = (int32_t) (A == (o ^ 1) ? du : dv);
                    kk->t = (int32_t) (A == (o ^ 1) ? dv : du);
                    kk->leaf = e; kk->face = f;
                }
                pe[slot] = -100;      /* filled by the matching below */
                pf[slot] = -1;
            }
        }
    }
    qsort(keys, (size_t) nkey, sizeof(gm_key_t), gm_key_cmp);
    for (i = 0; i < nkey; i++) {
        const gm_key_t *a = &keys[i], *b = i + 1 < nkey ? &keys[i + 1] : NULL;
        const int same = b && a->g0 == b->g0 && a->g1 == b->g1 && a->g2 == b->g2 && a->s == b->s && a->t == b->t;
        const int third = same && i + 2 < nkey && keys[i + 2].g0 == a->g0 && keys[i + 2].g1 == a->g1 &&
                          keys[i + 2].g2 == a->g2 && keys[i + 2].s == a->s && keys[i + 2].t == a->t;
        if (!same || third) {
            charm_set_error("gmesh_build: a face of tree %lld (face %d) meets %s", (long long) (a->leaf / nper),
                            a->face, same ? "more than one other tree" : "no other tree and has no boundary");
            goto fail;
        }
        pe[6 * a->leaf + a->face] = b->leaf; pf[6 * a->leaf + a->face] = (int8_t) b->face;
        pe[6 * b->leaf + b->face] = a->leaf; pf[6 * b->leaf + b->face] = (int8_t) a->face;
        i++;
    }
    /* faces grouped by (block of GM_FACE_BLOCK side-0 leaves, side-0 local
     * face, boundary or not), leaf order inside a group: the face pass's
     * threads of a wave then read and write the slots of consecutive leaves
     * on one local face and take one branch, and the six sweeps of a block
     * find its leaves' states in L2 (one XCD walks a contiguous face range) */
    const int64_t nblk = (m->n + GM_FACE_BLOCK - 1) / GM_FACE_BLOCK;
    next = (int64_t *) calloc((size_t) nblk * 12, sizeof(int64_t));
    if (!next) goto oom;
    for (e = 0; e < m->n; e++)
        for (int f = 0; f < 6; f++) {
            const int64_t p = pe[6 * e + f];
            if (p < 0 || p > e || (p == e && pf[6 * e + f] > f))
                next[12 * (e / GM_FACE_BLOCK) + 2 * f + (p < 0)]++;
        }
    for (int64_t b = 0; b < nblk * 12; b++) { const int64_t c = next[b]; next[b] = nf; nf += c; }
    m->nf = nf;
    m->fg = (double *) malloc((size_t) nf * CHARM_GM_FACEQ * sizeof(double));
    m->fn = (double *) malloc((size_t) nf * 3 * sizeof(double));
    m->fe = (int32_t *) malloc((size_t) nf * 2 * sizeof(int32_t));
    m->ff = (int32_t *) malloc((size_t) nf * 2 * sizeof(int32_t));
    if (!m->fg || !m->fn || !m->fe || !m->ff) goto oom;
    {
        int64_t fi;
        for (e = 0; e < m->n; e++) {
            gm_geom_t g;
            int have = 0;
            for (int f = 0; f < 6; f++) {
                const int64_t p = pe[6 * e + f];
                int pt, a;
                double nn[3];
                if (!(p < 0 || p > e || (p == e && pf[6 * e + f] > f))) continue;
                fi = next[12 * (e / GM_FACE_BLOCK) + 2 * f + (p < 0)]++;
                if (!have) {
                    gm_geom((const double (*)[3]) tvs[m->leaf[5 * e]], m->leaf + 5 * e + 1, &g);
                    have = 1;
                }
                m->fe[fi] = (int32_t) e;
                m->fe[nf + fi] = (int32_t) p;
                m->ff[fi] = f;
                m->ff[nf + fi] = p >= 0 ? pf[6 * e + f] : -1;
                gm_outward(&g, f, nn);
                for (a = 0; a < 3; a++) m->fn[(size_t) a * nf + fi] = nn[a];
                for (pt = 0; pt < 6; pt++) {
                    const double *x = g.fgp[f][pt];
                    for (a = 0; a < 3; a++) m->fg[(size_t) (3 * pt + a) * nf + fi] = (x[a] - g.c[a]) / g.dh[a];
                    m->fg[(size_t) (24 + pt) * nf + fi] = g.fwj[f][pt];
                }
                /* side 1's basis (x - c1) / dh1 = b0 dh0 / dh1 + (c0 - c1) / dh1 */
                for (a = 0; a < 3; a++) {
                    const double *g1 = p >= 0 ? m->geo + 23 * (size_t) p : NULL;
                    m->fg[(size_t) (18 + a) * nf + fi] = g1 ? g.dh[a] / g1[3 + a] : 0.0;
                    m->fg[(size_t) (21 + a) * nf + fi] = g1 ? (g.c[a] - g1[a]) / g1[3 + a] : 0.0;
                }
            }
        }
        /* limiter neighbours: the leaf across each face, or -1 - the boundary face */
        for (fi = 0; fi < nf; fi++) {
            const int32_t e0 = m->fe[fi], e1 = m->fe[nf + fi];
            m->lnb[(size_t) m->ff[fi] * m->n + e0] = e1 >= 0 ? e1 : (int32_t) (-1 - fi);
            if (e1 >= 0) m->lnb[(size_t) m->ff[nf + fi] * m->n + e1] = e0;
        }
    }
    free(pe); free(pf); free(keys); free(next);
    m->tv = tvs;
    *out = m;
    return CHARM_OK;
oom:
    charm_set_error("gmesh_build: out of memory");
fail:
    free(pe); free(pf); free(keys); free(tvs); free(next);
    charm_gmesh_destroy(m);
    return CHARM_ERR_MEM;
}

int charm_gmesh_info(const charm_gmesh_t *m, int64_t *n_leaves, int64_t *n_faces)
{
    if (!m) { charm_set_error("gmesh_info: null mesh"); return CHARM_ERR_ARG; }
    if (n_leaves) *n_leaves = m->n;
    if (n_faces) *n_faces = m->nf;
    return CHARM_OK;
}

int charm_gmesh_leaves(const charm_gmesh_t *m, int32_t *leaves)
{
    if (!m || !leaves) { charm_set_error("gmesh_leaves: null argument"); return CHARM_ERR_ARG; }
    memcpy(leaves, m->leaf, (size_t) m->n * 5 * sizeof(int32_t));
    return CHARM_OK;
}

int charm_gmesh_geometry(const charm_gmesh_t *m, double *out)
{
    if (!m || !out) { charm_set_error("gmesh_geometry: null argument"); return CHARM_ERR_ARG; }
    memcpy(out, m->geo, (size_t) m->n * 23 * sizeof(double));
    return CHARM_OK;
}

int charm_gmesh_set_regions(charm_gmesh_t *m, const int32_t *tree_reg)
{
    if (!m || !tree_reg) { charm_set_error("gmesh_set_regions: null argument"); return CHARM_ERR_ARG; }
    free(m->treg);
    m->treg = (int32_t *) malloc((size_t) (m->ntree + 1) * sizeof(int32_t));
    if (!m->treg) { charm_set_error("gmesh_set_regions: out of memory"); return CHARM_ERR_MEM; }
    memcpy(m->treg, tree_reg, (size_t) m->ntree * sizeof(int32_t));
    return CHARM_OK;
}

const int32_t *charm_gmesh_regions(const charm_gmesh_t *m)
{
    return m ? m->treg : NULL;
}

/* each leaf's 8 corners through its tree's trilinear map, [n][8][3] in
 * z-order (corner x + 2 y + 4 z) */
int charm_gmesh_leaf_corners(const charm_gmesh_t *m, double *out)
{
    if (!m || !out) { charm_set_error("gmesh_leaf_corners: null argument"); return CHARM_ERR_ARG; }
    for (int64_t e = 0; e < m->n; e++) {
        const int32_t *q = m->leaf + 5 * e;
        const int64_t len = (int64_t) GM_ROOT >> q[4];
        for (int k = 0; k < 8; k++)
            gm_tree_point((const double (*)[3]) m->tv[q[0]], q[1] + (k & 1) * len, q[2] + ((k >> 1) & 1) * len,
                          q[3] + (k >> 2) * len, out + 3 * (8 * e + k));
    }
    return CHARM_OK;
}

int charm_gmesh_faces(const charm_gmesh_t *m, int32_t *faces)
{
    int64_t i;
    if (!m || !faces) { charm_set_error("gmesh_faces: null argument"); return CHARM_ERR_ARG; }
    for (i = 0; i < m->nf; i++) {
        faces[4 * i] = m->fe[i]; faces[4 * i + 1] = m->fe[m->nf + i];
        faces[4 * i + 2] = m->ff[i]; faces[4 * i + 3] = m->ff[m->nf + i];
    }
    return CHARM_OK;
}

/* device-side arrays (charm_gpu.hip uploads them) */
int charm_gmesh_arrays(const charm_gmesh_t *m, charm_gmesh_arrays_t *a)
{
    if (!m || !a) { charm_set_error("gmesh_arrays: null argument"); return CHARM_ERR_ARG; }
    a->n = m->n; a->nf = m->nf;
    a->lg = m->lg; a->fg = m->fg; a->fn = m->fn; a->fe = m->fe; a->ff = m->ff; a->lnb = m->lnb;
    return CHARM_OK;
}
