/*
 * or_model.c -- forest traversal, residual integrals, limiter and time
 * steppers of the oracle (FP64, TEST INFRASTRUCTURE ONLY).
 *
 * The face traversal is this file's own: a hash map of the leaves and a walk
 * over each leaf's upper faces that emits every face once, with the side and
 * hanging layout p4est_iterate hands to the reference's face callbacks
 * (side 0 = lower-coordinate quadrant with face 2a+1, side 1 = upper one with
 * face 2a; a hanging side lists its 4 small quadrants in face z-order; a
 * domain-boundary face has one side).  Periodic faces (not supported by the
 * reference, which evaluates neighbour bases at side-0 coordinates) evaluate
 * the non-owning side at the GP shifted by one period.
 */
#include "or_internal.h"
#ifdef _OPENMP
#include <omp.h>
/* element loops whose iterations touch only their own leaf */
#define OR_PARFOR _Pragma("omp parallel for schedule(static)")
#else
#define OR_PARFOR
#endif

/* ------------------------------------------------------------ hash map */
typedef struct { uint64_t *key; int64_t *val; uint64_t mask; } or_hash_t;

static uint64_t or_key(int level, int32_t x, int32_t y, int32_t z)
{
    return ((uint64_t)level << 57) | ((uint64_t)x << 38) | ((uint64_t)y << 19) | (uint64_t)z;
}

static uint64_t or_mix(uint64_t k)
{
    k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL; k ^= k >> 33;
    return k;
}

static int or_hash_init(or_hash_t *h, int64_t n)
{
    uint64_t cap = 16;
    while (cap < (uint64_t)(2 * n + 16)) cap <<= 1;
    h->key = (uint64_t *) malloc(cap * sizeof(uint64_t));
    h->val = (int64_t *) malloc(cap * sizeof(int64_t));
    if (!h->key || !h->val) return 1;
    memset(h->key, 0xff, cap * sizeof(uint64_t));
    h->mask = cap - 1;
    return 0;
}

static void or_hash_put(or_hash_t *h, uint64_t k, int64_t v)
{
    uint64_t i = or_mix(k) & h->mask;
    while (h->key[i] != ~0ULL && h->key[i] != k) i = (i + 1) & h->mask;
    h->key[i] = k; h->val[i] = v;
}

static int64_t or_hash_get(const or_hash_t *h, uint64_t k)
{
    uint64_t i = or_mix(k) & h->mask;
    while (h->key[i] != ~0ULL) {
        if (h->key[i] == k) return h->val[i];
        i = (i + 1) & h->mask;
    }
    return -1;
}

static int or_face_push(or_sim_t *s, int64_t *cap, const or_face_t *f)
{
    if (s->nf == *cap) {
        int64_t nc = *cap ? 2 * *cap : 1024;
        or_face_t *nf = (or_face_t *) realloc(s->f, nc * sizeof(or_face_t));
        if (!nf) return 1;
        s->f = nf; *cap = nc;
    }
    s->f[s->nf++] = *f;
    return 0;
}

static int or_periodic(const or_config_t *cfg, int a)
{
    return cfg->bnd_type[2*a] == OR_BND_PERIODIC;
}

static int or_build_faces(or_sim_t *s)
{
    or_hash_t h;
    int64_t e, cap = 0;
    int a, j;
    if (or_hash_init(&h, s->n)) return 1;
    for (e = 0; e < s->n; e++) {
        const int32_t *q = s->d[e].quad;
        or_hash_put(&h, or_key(q[3], q[0], q[1], q[2]), e);
    }
    for (e = 0; e < s->n; e++) {
        const int32_t *q = s->d[e].quad;
        int32_t len = 1 << (OR_MAXLEVEL - q[3]);
        for (a = 0; a < 3; a++) {
            int t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
            or_face_t f;
            int32_t p[3] = {q[0], q[1], q[2]};
            int64_t nb;
            memset(&f, 0, sizeof(f));
            f.wrap_axis = -1;
            if (q[a] == 0 && !or_periodic(&s->cfg, a)) {
                f.nsides = 1; f.face[0] = 2*a; f.quad[0][0] = (int) e; f.bnd = 2*a;
                if (or_face_push(s, &cap, &f)) goto fail;
            }
            p[a] += len;
            if (p[a] == OR_ROOT_LEN) {
                if (!or_periodic(&s->cfg, a)) {
                    memset(&f, 0, sizeof(f));
                    f.wrap_axis = -1;
                    f.nsides = 1; f.face[0] = 2*a+1; f.quad[0][0] = (int) e; f.bnd = 2*a+1;
                    if (or_face_push(s, &cap, &f)) goto fail;
                    continue;
                }
                p[a] = 0;
                f.wrap_axis = a;
            }
            f.nsides = 2; f.face[0] = 2*a+1; f.face[1] = 2*a;
            nb = or_hash_get(&h, or_key(q[3], p[0], p[1], p[2]));
            if (nb >= 0) {
                f.quad[0][0] = (int) e; f.quad[1][0] = (int) nb;
                if (or_face_push(s, &cap, &f)) goto fail;
                continue;
            }
            if (q[3] > 0) {
                int32_t lb = 2 * len, pb[3];
                for (j = 0; j < 3; j++) pb[j] = (p[j] / lb) * lb;
                nb = or_hash_get(&h, or_key(q[3]-1, pb[0], pb[1], pb[2]));
                if (nb >= 0) {
                    if (q[t1] == pb[t1] && q[t2] == pb[t2]) {
                        for (j = 0; j < 4; j++) {
                            int32_t c[3];
                            c[a] = q[a];
                            c[t1] = pb[t1] + (j & 1) * len;
                            c[t2] = pb[t2] + (j >> 1) * len;
                            int64_t sm = or_hash_get(&h, or_key(q[3], c[0], c[1], c[2]));
                            if (sm < 0) { snprintf(s->err, sizeof(s->err), "hanging face: missing small quad"); goto fail; }
                            f.quad[0][j] = (int) sm;
                        }
                        f.is_hanging[0] = 1;
                        f.quad[1][0] = (int) nb;
                        if (or_face_push(s, &cap, &f)) goto fail;
                    }
                    continue;
                }
            }
            for (j = 0; j < 4; j++) {
                int32_t c[3], hl = len / 2;
                c[a] = p[a];
                c[t1] = p[t1] + (j & 1) * hl;
                c[t2] = p[t2] + (j >> 1) * hl;
                int64_t sm = (q[3] < OR_MAXLEVEL) ? or_hash_get(&h, or_key(q[3]+1, c[0], c[1], c[2])) : -1;
                if (sm < 0) { snprintf(s->err, sizeof(s->err), "forest is not 2:1 face balanced"); goto fail; }
                f.quad[1][j] = (int) sm;
            }
            f.quad[0][0] = (int) e;
            f.is_hanging[1] = 1;
            if (or_face_push(s, &cap, &f)) goto fail;
        }
    }
    free(h.key); free(h.val);
    return 0;
fail:
    free(h.key); free(h.val);
    if (!s->err[0]) snprintf(s->err, sizeof(s->err), "face build failed");
    return 1;
}

/* ------------------------------------------------------------ lifecycle */
or_sim_t *or_create(const or_config_t *cfg, int64_t n_leaves, const int32_t *leaves)
{
    or_sim_t *s = (or_sim_t *) calloc(1, sizeof(or_sim_t));
    int64_t e;
    if (!s) return NULL;
    s->cfg = *cfg;
    s->n = n_leaves;
    s->nv = 4 + cfg->n_comp;
    s->d = (or_data_t *) calloc((size_t) n_leaves, sizeof(or_data_t));
    if (!s->d) { snprintf(s->err, sizeof(s->err), "out of memory"); return s; }
    OR_PARFOR
    for (e = 0; e < n_leaves; e++) {
        or_data_t *d = &s->d[e];
        memcpy(d->quad, &leaves[4*e], 4*sizeof(int32_t));
        or_geom_calc(d->quad, &d->g);
        d->grav[0] = cfg->grav[0]; d->grav[1] = cfg->grav[1]; d->grav[2] = cfg->grav[2];
    }
    or_build_faces(s);
    return s;
}

void or_leaf_vertices(const or_sim_t *s, const or_data_t *d, double v[8][3])
{
    if (s->tv) or_quad_vertices_tv((const double (*)[3]) (s->tv + 24 * (size_t) d->tree), d->quad, v);
    else or_quad_vertices(d->quad, v);
}

/* Multi-tree forest: every tree refined uniformly to `level` (leaves in
 * tree order, Morton order inside a tree, the order of p4est's global
 * numbering), geometry from each tree's vertices (src/charm_geom.c with
 * p4est_qcoord_to_vertex over the tree's own corners), faces: inside a tree
 * as the single-tree walk; on a tree face with a boundary slot
 * (tree_bnd >= 0) a boundary face; every other tree-face leaf face is paired
 * with the one leaf face elsewhere whose centre coincides with it (brute
 * force over those faces, tolerance 1e-9 of the face's size).  Side 0 of a
 * pair is the leaf with the smaller index.  Conforming forests only. */
or_sim_t *or_create_trees(const or_config_t *cfg, int64_t n_trees, const double *tree_vertices,
                          const int32_t *tree_bnd, int level)
{
    or_sim_t *s = (or_sim_t *) calloc(1, sizeof(or_sim_t));
    const int64_t nt = (int64_t) 1 << (3 * level), nside = (int64_t) 1 << level;
    const int32_t len = 1 << (OR_MAXLEVEL - level);
    int64_t e, cap = 0, nc = 0, i, j;
    int a;
    struct cand { int64_t e; int f; double x[3], tol; int used; } *cd = NULL;
    if (!s) return NULL;
    s->cfg = *cfg;
    s->n = n_trees * nt;
    s->nv = 4 + cfg->n_comp;
    s->d = (or_data_t *) calloc((size_t) s->n, sizeof(or_data_t));
    s->tv = (double *) malloc((size_t) n_trees * 24 * sizeof(double));
    cd = (struct cand *) malloc((size_t) (s->n * 6 + 1) * sizeof(*cd));
    if (!s->d || !s->tv || !cd || level < 0 || level > 8) {
        snprintf(s->err, sizeof(s->err), "out of memory or bad level");
        free(cd);
        return s;
    }
    memcpy(s->tv, tree_vertices, (size_t) n_trees * 24 * sizeof(double));
    for (e = 0; e < s->n; e++) {
        or_data_t *d = &s->d[e];
        const int64_t m = e % nt;
        int32_t c[3] = {0, 0, 0};
        int b;
        for (b = 0; b < level; b++)
            for (a = 0; a < 3; a++) c[a] |= (int32_t) (((m >> (3 * b + a)) & 1) << b);
        d->tree = (int32_t) (e / nt);
        d->quad[0] = c[0] * len; d->quad[1] = c[1] * len; d->quad[2] = c[2] * len; d->quad[3] = level;
        or_geom_calc_tv((const double (*)[3]) (s->tv + 24 * (size_t) d->tree), d->quad, &d->g);
        d->grav[0] = cfg->grav[0]; d->grav[1] = cfg->grav[1]; d->grav[2] = cfg->grav[2];
    }
    for (e = 0; e < s->n; e++) {
        const or_data_t *d = &s->d[e];
        const int64_t t = d->tree, base = t * nt;
        for (a = 0; a < 3; a++) {
            int side;
            for (side = 0; side < 2; side++) {
                const int f = 2 * a + side;
                const int32_t ci = d->quad[a] / len;
                or_face_t fr;
                memset(&fr, 0, sizeof(fr));
                fr.wrap_axis = -1;
                if ((side == 1 && ci + 1 < nside) || (side == 0 && ci > 0)) {
                    if (side == 1) {                    /* inside the tree, from the lower leaf */
                        int32_t p[3] = {d->quad[0] / len, d->quad[1] / len, d->quad[2] / len};
                        int64_t mm = 0;
                        int b;
                        p[a] += 1;
                        for (b = 0; b < level; b++)
                            for (j = 0; j < 3; j++) mm |= (int64_t) ((p[j] >> b) & 1) << (3 * b + j);
                        fr.nsides = 2; fr.face[0] = f; fr.face[1] = f - 1;
                        fr.quad[0][0] = (int) e; fr.quad[1][0] = (int) (base + mm);
                        if (or_face_push(s, &cap, &fr)) goto fail;
                    }
                    continue;
                }
                if (tree_bnd[6 * t + f] >= 0) {
                    fr.nsides = 1; fr.face[0] = f; fr.quad[0][0] = (int) e; fr.bnd = tree_bnd[6 * t + f];
                    if (or_face_push(s, &cap, &fr)) goto fail;
                    continue;
                }
                cd[nc].e = e; cd[nc].f = f; cd[nc].used = 0;
                memcpy(cd[nc].x, d->g.fc[f], sizeof(cd[nc].x));
                cd[nc].tol = 1e-9 * sqrt(d->g.area[f]);
                nc++;
            }
        }
    }
    for (i = 0; i < nc; i++) {
        int64_t hit = -1;
        if (cd[i].used) continue;
        for (j = i + 1; j < nc; j++) {
            double dd = 0.;
            int k;
            if (cd[j].used) continue;
            for (k = 0; k < 3; k++) dd = OR_MAX(dd, fabs(cd[i].x[k] - cd[j].x[k]));
            if (dd <= cd[i].tol) {
                if (hit >= 0) { snprintf(s->err, sizeof(s->err), "tree face: two matches"); goto fail; }
                hit = j;
            }
        }
        if (hit < 0) { snprintf(s->err, sizeof(s->err), "tree face without a partner or a boundary"); goto fail; }
        cd[i].used = cd[hit].used = 1;
        {
            or_face_t fr;
            memset(&fr, 0, sizeof(fr));
            fr.wrap_axis = -1;
            fr.nsides = 2;
            fr.face[0] = cd[i].f; fr.face[1] = cd[hit].f;          /* cd is in leaf order */
            fr.quad[0][0] = (int) cd[i].e; fr.quad[1][0] = (int) cd[hit].e;
            if (or_face_push(s, &cap, &fr)) goto fail;
        }
    }
    free(cd);
    return s;
fail:
    free(cd);
    if (!s->err[0]) snprintf(s->err, sizeof(s->err), "face build failed");
    return s;
}

void or_destroy(or_sim_t *s)
{
    if (!s) return;
    free(s->inc_ptr); free(s->inc); free(s->fq_off); free(s->fq); free(s->lq); free(s->tv);
    free(s->d); free(s->f); free(s);
}

int or_nvar(const or_sim_t *s) { return s->nv; }
int64_t or_n_faces(const or_sim_t *s) { return s->nf; }
const char *or_error(const or_sim_t *s) { return s->err[0] ? s->err : NULL; }

void or_set_state(or_sim_t *s, const double *u)
{
    int64_t e; int v, k;
    for (e = 0; e < s->n; e++)
        for (v = 0; v < s->nv; v++)
            for (k = 0; k < OR_BF; k++) s->d[e].c[v][k] = u[(e*s->nv + v)*OR_BF + k];
}

void or_get_state(const or_sim_t *s, double *u)
{
    int64_t e; int v, k;
    for (e = 0; e < s->n; e++)
        for (v = 0; v < s->nv; v++)
            for (k = 0; k < OR_BF; k++) u[(e*s->nv + v)*OR_BF + k] = s->d[e].c[v][k];
}

void or_faces(const or_sim_t *s, int32_t *out)
{
    int64_t i; int j;
    for (i = 0; i < s->nf; i++) {
        const or_face_t *f = &s->f[i];
        int32_t *o = &out[16*i];
        o[0] = f->nsides; o[1] = f->face[0]; o[2] = f->face[1];
        o[3] = f->is_hanging[0]; o[4] = f->is_hanging[1];
        for (j = 0; j < 4; j++) { o[5+j] = f->quad[0][j]; o[9+j] = f->quad[1][j]; }
        o[13] = f->bnd; o[14] = f->wrap_axis; o[15] = 0;
    }
}

void or_geom(const or_sim_t *s, int64_t leaf, double *out)
{
    const or_geom_t *g = &s->d[leaf].g;
    double *o = out;
    memcpy(o, g->n, sizeof(g->n)); o += 18;
    memcpy(o, g->fc, sizeof(g->fc)); o += 18;
    memcpy(o, g->face_gp, sizeof(g->face_gp)); o += 108;
    memcpy(o, g->face_gw, sizeof(g->face_gw)); o += 36;
    memcpy(o, g->face_gj, sizeof(g->face_gj)); o += 36;
    memcpy(o, g->quad_gp, sizeof(g->quad_gp)); o += 24;
    memcpy(o, g->quad_gw, sizeof(g->quad_gw)); o += 8;
    memcpy(o, g->quad_gj, sizeof(g->quad_gj)); o += 8;
    memcpy(o, g->area, sizeof(g->area)); o += 6;
    *o++ = g->volume;
    memcpy(o, g->c, sizeof(g->c)); o += 3;
    memcpy(o, g->dh, sizeof(g->dh)); o += 3;
    memcpy(o, g->a, sizeof(g->a)); o += 16;
    memcpy(o, g->a_inv, sizeof(g->a_inv)); o += 16;
}

/* ------------------------------------------------------------ fields */

/* charm_get_fields, src/charm_base_func.c:139-211 (loops bounded by the
 * real species count, as the reference's @todo intends) */
void or_get_fields(const or_sim_t *s, const or_data_t *d, const double x[3], or_cons_t *c)
{
    int v, i;
    double r[OR_REC_NV];
    for (v = 0; v < OR_FIELD_NV(s); v++) {
        r[v] = 0.;
        for (i = 0; i < OR_BF; i++) r[v] += d->c[v][i]*or_bf(&d->g, x, i);
    }
    c->ru = r[0]; c->rv = r[1]; c->rw = r[2]; c->re = r[3];
    for (i = 0; i < s->cfg.n_comp; i++) c->rc[i] = r[4+i];
    c->mat_id = d->mat_id;
}

/* charm_get_fields_avg, src/charm_base_func.c:214-240 */
void or_get_fields_avg(const or_sim_t *s, const or_data_t *d, or_cons_t *c)
{
    int igp, v, i;
    double r[OR_REC_NV];
    for (v = 0; v < OR_FIELD_NV(s); v++) r[v] = 0.;
    for (igp = 0; igp < OR_QGP; igp++) {
        const double *gx = d->g.quad_gp[igp];
        double gjw = d->g.quad_gj[igp]*d->g.quad_gw[igp]/d->g.volume;
        for (v = 0; v < OR_FIELD_NV(s); v++) {
            double f = 0.;
            for (i = 0; i < OR_BF; i++) f += d->c[v][i]*or_bf(&d->g, gx, i);
            r[v] += f * gjw;
        }
    }
    c->ru = r[0]; c->rv = r[1]; c->rw = r[2]; c->re = r[3];
    for (i = 0; i < s->cfg.n_comp; i++) c->rc[i] = r[4+i];
    c->mat_id = d->mat_id;
}

static void or_cons_get(const or_sim_t *s, const or_cons_t *c, double *u)
{
    int i;
    u[0] = c->ru; u[1] = c->rv; u[2] = c->rw; u[3] = c->re;
    for (i = 0; i < s->cfg.n_comp; i++) u[4+i] = c->rc[i];
}

/* outward normal of side 0's face: charm_face_get_normal plus the
 * orientation flip by (fc - c), src/models/euler/charm_model.c:332-344 */
void or_face_normal(const or_data_t *d, int face, double n[3])
{
    double l[3];
    int i;
    for (i = 0; i < 3; i++) { n[i] = d->g.n[face][i]; l[i] = d->g.fc[face][i] - d->g.c[i]; }
    if (or_scalar_prod(n, l) < 0)
        for (i = 0; i < 3; i++) n[i] *= -1.0;
}

void or_shift(const or_face_t *f, int side, int owner, const double x[3], double xs[3])
{
    xs[0] = x[0]; xs[1] = x[1]; xs[2] = x[2];
    if (f->wrap_axis >= 0 && side != owner)
        xs[f->wrap_axis] += (side == 1) ? -1.0 : 1.0;
}

void or_ns_diff(or_sim_t *s);

/* ----------------------------------------------------- convective terms */

/* volume integral, src/models/euler/charm_model.c:66-132
 * (= src/models/ns/charm_model_ns_conv.c:23-89) */
static void or_conv_volume(or_sim_t *s, or_data_t *d)
{
    int ibf, igp, cj, nc = s->cfg.n_comp;
    or_cons_t c; or_prim_t p;
    double fu, fv, fw, fe, gu, gv, gw, ge, hu, hv, hw, he, fc[OR_NC_MAX], gc[OR_NC_MAX], hc[OR_NC_MAX];
    for (ibf = 0; ibf < OR_BF; ibf++) {
        for (igp = 0; igp < OR_QGP; igp++) {
            const double *x = d->g.quad_gp[igp];
            double wj = d->g.quad_gj[igp] * d->g.quad_gw[igp];
            or_get_fields(s, d, x, &c);
            or_cons_to_prim(&s->cfg, &c, &p);
            fu = c.ru*p.u+p.p; fv = c.ru*p.v; fw = c.ru*p.w; fe = c.ru*p.e_tot+p.p*p.u;
            gu = c.rv*p.u; gv = c.rv*p.v+p.p; gw = c.rv*p.w; ge = c.rv*p.e_tot+p.p*p.v;
            hu = c.rw*p.u; hv = c.rw*p.v; hw = c.rw*p.w+p.p; he = c.rw*p.e_tot+p.p*p.w;
            for (cj = 0; cj < nc; ++cj) {
                fc[cj] = c.ru*p.c[cj]; gc[cj] = c.rv*p.c[cj]; hc[cj] = c.rw*p.c[cj];
            }
            double phi_x = or_bf_dx(&d->g, ibf) * d->g.quad_gj[igp] * d->g.quad_gw[igp];
            double phi_y = or_bf_dy(&d->g, ibf) * d->g.quad_gj[igp] * d->g.quad_gw[igp];
            double phi_z = or_bf_dz(&d->g, ibf) * d->g.quad_gj[igp] * d->g.quad_gw[igp];
            double phi = or_bf(&d->g, x, ibf) * d->g.quad_gj[igp] * d->g.quad_gw[igp];
            (void) wj;
            for (cj = 0; cj < nc; ++cj)
                d->integ[4+cj][ibf] -= (fc[cj]*phi_x+gc[cj]*phi_y+hc[cj]*phi_z);
            d->integ[0][ibf] -= ((fu*phi_x+gu*phi_y+hu*phi_z) + p.r*d->grav[0]*phi);
            d->integ[1][ibf] -= ((fv*phi_x+gv*phi_y+hv*phi_z) + p.r*d->grav[1]*phi);
            d->integ[2][ibf] -= ((fw*phi_x+gw*phi_y+hw*phi_z) + p.r*d->grav[2]*phi);
            d->integ[3][ibf] -= (fe*phi_x+ge*phi_y+he*phi_z);
        }
    }
}

static void or_accum(or_sim_t *s, or_data_t *d, const double *q, double bfv, int ibf)
{
    int v;
    for (v = 4; v < s->nv; v++) d->integ[v][ibf] += q[v] * bfv;
    for (v = 0; v < 4; v++) d->integ[v][ibf] += q[v] * bfv;
}

static int or_flux_checked(or_sim_t *s, const or_prim_t prim[2], const double n[3], double *q)
{
    if (or_flux_err(&s->cfg, s->cfg.flux, prim, n, q)) {
        snprintf(s->err, sizeof(s->err), "HLLC: inequality SL <= S* <= SR is FALSE");
        return 1;
    }
    return 0;
}

/* boundary face, src/models/euler/charm_model.c:140-212 */
static void or_conv_bnd(or_sim_t *s, const or_face_t *f)
{
    or_data_t *d = &s->d[f->quad[0][0]];
    int face = f->face[0], igp, ibf;
    double n[3], q[OR_NV_MAX];
    or_cons_t cons; or_prim_t prim[2];
    or_face_normal(d, face, n);
    for (igp = 0; igp < OR_FGP; igp++) {
        const double *x = d->g.face_gp[face][igp];
        double gw = d->g.face_gw[face][igp], gj = d->g.face_gj[face][igp];
        or_get_fields(s, d, x, &cons);
        or_cons_to_prim(&s->cfg, &cons, &prim[0]);
        or_bnd(&s->cfg, f->bnd, &prim[0], &prim[1], n);
        or_flux_checked(s, prim, n, q);
        for (ibf = 0; ibf < OR_BF; ibf++)
            or_accum(s, d, q, or_bf(&d->g, x, ibf) * gw * gj, ibf);
    }
}

/* interior face, conforming and hanging, src/models/euler/charm_model.c:
 * 215-372 */
static void or_conv_inner(or_sim_t *s, const or_face_t *f)
{
    int i, j, igp, ibf, h_side = -1;
    double n[3], q[OR_NV_MAX], xs[2][3];
    or_data_t *ud[2];
    or_cons_t cons[2]; or_prim_t prim[2];
    if (f->is_hanging[0] || f->is_hanging[1]) {
        for (j = 0; j < 4; j++) {
            for (i = 0; i < 2; i++) {
                if (f->is_hanging[i]) { ud[i] = &s->d[f->quad[i][j]]; h_side = i; }
                else ud[i] = &s->d[f->quad[i][0]];
            }
            or_face_normal(ud[0], f->face[0], n);
            for (igp = 0; igp < OR_FGP; igp++) {
                const double *x = ud[h_side]->g.face_gp[f->face[h_side]][igp];
                double gw = ud[h_side]->g.face_gw[f->face[h_side]][igp];
                double gj = ud[h_side]->g.face_gj[f->face[h_side]][igp];
                for (i = 0; i < 2; i++) {
                    or_shift(f, i, h_side, x, xs[i]);
                    or_get_fields(s, ud[i], xs[i], &cons[i]);
                    or_cons_to_prim(&s->cfg, &cons[i], &prim[i]);
                }
                or_flux_checked(s, prim, n, q);
                for (ibf = 0; ibf < OR_BF; ibf++)
                    for (i = 0; i < 2; i++)
                        or_accum(s, ud[i], q, (i ? -1. : 1.) * or_bf(&ud[i]->g, xs[i], ibf) * gw * gj, ibf);
            }
        }
    } else {
        for (i = 0; i < 2; i++) ud[i] = &s->d[f->quad[i][0]];
        or_face_normal(ud[0], f->face[0], n);
        for (igp = 0; igp < OR_FGP; igp++) {
            const double *x = ud[0]->g.face_gp[f->face[0]][igp];
            double gw = ud[0]->g.face_gw[f->face[0]][igp];
            double gj = ud[0]->g.face_gj[f->face[0]][igp];
            for (i = 0; i < 2; i++) {
                or_shift(f, i, 0, x, xs[i]);
                or_get_fields(s, ud[i], xs[i], &cons[i]);
                or_cons_to_prim(&s->cfg, &cons[i], &prim[i]);
            }
            or_flux_checked(s, prim, n, q);
            for (ibf = 0; ibf < OR_BF; ibf++)
                for (i = 0; i < 2; i++)
                    or_accum(s, ud[i], q, (i ? -1. : 1.) * or_bf(&ud[i]->g, xs[i], ibf) * gw * gj, ibf);
        }
    }
}

/* zero pass, src/models/euler/charm_model.c:436-450 */
static void or_zero(or_sim_t *s)
{
    int64_t e;
    OR_PARFOR
    for (e = 0; e < s->n; e++) memset(s->d[e].integ, 0, sizeof(s->d[e].integ));
}

#ifdef _OPENMP
/* ------------------------------------------- parallel convective pass
 * (OpenMP builds: the bench's CPU timing).  Phase 1 evaluates every face's
 * fluxes at its points in parallel into fq; phase 2 lets every leaf add its
 * faces' contributions itself, in face order, with the serial pass's
 * expressions: each accumulator sees the same sums in the same order (a leaf
 * that is both sides of one face, a one-leaf periodic box, is the only
 * exception). */
static int or_nsub(const or_face_t *f)
{
    return (f->nsides == 2 && (f->is_hanging[0] || f->is_hanging[1])) ? 4 : 1;
}

static int or_par_setup(or_sim_t *s)
{
    int64_t e, fi, k, *pos;
    int side, j;
    if (s->fq) return 0;
    s->inc_ptr = (int64_t *) calloc((size_t) s->n + 1, sizeof(int64_t));
    s->fq_off = (int64_t *) malloc(((size_t) s->nf + 1) * sizeof(int64_t));
    pos = (int64_t *) malloc(((size_t) s->n + 1) * sizeof(int64_t));
    if (!s->inc_ptr || !s->fq_off || !pos) { free(pos); return 1; }
    for (fi = 0; fi < s->nf; fi++) {
        const or_face_t *f = &s->f[fi];
        for (side = 0; side < f->nsides; side++) {
            if (f->is_hanging[side]) for (j = 0; j < 4; j++) s->inc_ptr[f->quad[side][j] + 1]++;
            else s->inc_ptr[f->quad[side][0] + 1]++;
        }
    }
    for (e = 0; e < s->n; e++) s->inc_ptr[e + 1] += s->inc_ptr[e];
    s->inc = (int64_t *) malloc((size_t) s->inc_ptr[s->n] * sizeof(int64_t));
    if (!s->inc) { free(pos); return 1; }
    memcpy(pos, s->inc_ptr, (size_t) s->n * sizeof(int64_t));
    for (fi = 0; fi < s->nf; fi++) {              /* ascending: each list in face order */
        const or_face_t *f = &s->f[fi];
        for (side = 0; side < f->nsides; side++) {
            if (f->is_hanging[side])
                for (j = 0; j < 4; j++) s->inc[pos[f->quad[side][j]]++] = fi << 4 | side << 3 | j;
            else s->inc[pos[f->quad[side][0]]++] = fi << 4 | side << 3 | 7;
        }
    }
    free(pos);
    for (k = 0, fi = 0; fi < s->nf; fi++) {
        s->fq_off[fi] = k;
        k += (int64_t) or_nsub(&s->f[fi]) * OR_FGP * s->nv;
    }
    s->fq_off[s->nf] = k;
    s->fq = (double *) malloc((size_t) k * sizeof(double));
    s->lq = (double *) malloc(((size_t) s->nf * 2 * OR_NV_MAX + 1) * sizeof(double));
    return s->fq == NULL || s->lq == NULL;
}

/* phase 1: the fluxes of or_conv_bnd / or_conv_inner at every face point */
static void or_face_fluxes(or_sim_t *s, int *bad)
{
    int64_t fi;
    #pragma omp parallel for schedule(dynamic, 256)
    for (fi = 0; fi < s->nf; fi++) {
        const or_face_t *f = &s->f[fi];
        double *out = s->fq + s->fq_off[fi], n[3], xs[2][3];
        or_cons_t cons[2]; or_prim_t prim[2];
        or_data_t *ud[2];
        int i, j, igp, h_side = -1;
        if (f->nsides != 2) {
            or_data_t *d = &s->d[f->quad[0][0]];
            or_face_normal(d, f->face[0], n);
            for (igp = 0; igp < OR_FGP; igp++) {
                or_get_fields(s, d, d->g.face_gp[f->face[0]][igp], &cons[0]);
                or_cons_to_prim(&s->cfg, &cons[0], &prim[0]);
                or_bnd(&s->cfg, f->bnd, &prim[0], &prim[1], n);
                if (or_flux_err(&s->cfg, s->cfg.flux, prim, n, out + igp * s->nv)) *bad = 1;
            }
        } else if (f->is_hanging[0] || f->is_hanging[1]) {
            for (j = 0; j < 4; j++) {
                for (i = 0; i < 2; i++) {
                    if (f->is_hanging[i]) { ud[i] = &s->d[f->quad[i][j]]; h_side = i; }
                    else ud[i] = &s->d[f->quad[i][0]];
                }
                or_face_normal(ud[0], f->face[0], n);
                for (igp = 0; igp < OR_FGP; igp++) {
                    const double *x = ud[h_side]->g.face_gp[f->face[h_side]][igp];
                    for (i = 0; i < 2; i++) {
                        or_shift(f, i, h_side, x, xs[i]);
                        or_get_fields(s, ud[i], xs[i], &cons[i]);
                        or_cons_to_prim(&s->cfg, &cons[i], &prim[i]);
                    }
                    if (or_flux_err(&s->cfg, s->cfg.flux, prim, n, out + (j * OR_FGP + igp) * s->nv)) *bad = 1;
                }
            }
        } else {
            for (i = 0; i < 2; i++) ud[i] = &s->d[f->quad[i][0]];
            or_face_normal(ud[0], f->face[0], n);
            for (igp = 0; igp < OR_FGP; igp++) {
                const double *x = ud[0]->g.face_gp[f->face[0]][igp];
                for (i = 0; i < 2; i++) {
                    or_shift(f, i, 0, x, xs[i]);
                    or_get_fields(s, ud[i], xs[i], &cons[i]);
                    or_cons_to_prim(&s->cfg, &cons[i], &prim[i]);
                }
                if (or_flux_err(&s->cfg, s->cfg.flux, prim, n, out + igp * s->nv)) *bad = 1;
            }
        }
    }
}

static void or_accum(or_sim_t *s, or_data_t *d, const double *q, double bfv, int ibf);

/* phase 2: each leaf adds its faces' contributions, in face order */
static void or_face_gather(or_sim_t *s)
{
    int64_t e;
    #pragma omp parallel for schedule(static)
    for (e = 0; e < s->n; e++) {
        or_data_t *d = &s->d[e];
        int64_t k;
        for (k = s->inc_ptr[e]; k < s->inc_ptr[e + 1]; k++) {
            const int64_t fi = s->inc[k] >> 4;
            const int side = (int) ((s->inc[k] >> 3) & 1), jj = (int) (s->inc[k] & 7);
            const or_face_t *f = &s->f[fi];
            const double *q0 = s->fq + s->fq_off[fi];
            double xs[3];
            int igp, ibf, j;
            if (f->nsides != 2) {
                const int face = f->face[0];
                for (igp = 0; igp < OR_FGP; igp++) {
                    const double *x = d->g.face_gp[face][igp];
                    double gw = d->g.face_gw[face][igp], gj = d->g.face_gj[face][igp];
                    for (ibf = 0; ibf < OR_BF; ibf++)
                        or_accum(s, d, q0 + igp * s->nv, or_bf(&d->g, x, ibf) * gw * gj, ibf);
                }
            } else if (f->is_hanging[0] || f->is_hanging[1]) {
                const int h_side = f->is_hanging[0] ? 0 : 1;
                for (j = (jj == 7 ? 0 : jj); j <= (jj == 7 ? 3 : jj); j++) {
                    const or_data_t *hs = &s->d[f->quad[h_side][j]];
                    for (igp = 0; igp < OR_FGP; igp++) {
                        const double *x = hs->g.face_gp[f->face[h_side]][igp];
                        double gw = hs->g.face_gw[f->face[h_side]][igp];
                        double gj = hs->g.face_gj[f->face[h_side]][igp];
                        or_shift(f, side, h_side, x, xs);
                        for (ibf = 0; ibf < OR_BF; ibf++)
                            or_accum(s, d, q0 + (j * OR_FGP + igp) * s->nv,
                                     (side ? -1. : 1.) * or_bf(&d->g, xs, ibf) * gw * gj, ibf);
                    }
                }
            } else {
                const or_data_t *d0 = &s->d[f->quad[0][0]];
                for (igp = 0; igp < OR_FGP; igp++) {
                    const double *x = d0->g.face_gp[f->face[0]][igp];
                    double gw = d0->g.face_gw[f->face[0]][igp], gj = d0->g.face_gj[f->face[0]][igp];
                    or_shift(f, side, 0, x, xs);
                    for (ibf = 0; ibf < OR_BF; ibf++)
                        or_accum(s, d, q0 + igp * s->nv, (side ? -1. : 1.) * or_bf(&d->g, xs, ibf) * gw * gj, ibf);
                }
            }
        }
    }
}

/* 0: done in parallel; nonzero: the serial pass runs instead */
static int or_conv_parallel(or_sim_t *s)
{
    int64_t e;
    int bad = 0;
    if (omp_get_max_threads() < 2 || or_par_setup(s)) return 1;
    #pragma omp parallel for schedule(static)
    for (e = 0; e < s->n; e++) or_conv_volume(s, &s->d[e]);
    or_face_fluxes(s, &bad);
    or_face_gather(s);
    if (bad) snprintf(s->err, sizeof(s->err), "HLLC: inequality SL <= S* <= SR is FALSE");
    return 0;
}
#endif

/* charm_timestep_conv, src/models/euler/charm_model.c:388-396 */
static void or_conv(or_sim_t *s)
{
    int64_t e, i;
#ifdef _OPENMP
    if (!or_conv_parallel(s)) return;
#endif
    for (e = 0; e < s->n; e++) or_conv_volume(s, &s->d[e]);
    for (i = 0; i < s->nf; i++) {
        if (s->f[i].nsides != 2) or_conv_bnd(s, &s->f[i]);
        else or_conv_inner(s, &s->f[i]);
    }
}

/* zero + viscous (when enabled) + convective pass of an NS stage:
 * src/models/ns/charm_model.c:246-252 without the chemistry */
void or_ns_residual(or_sim_t *s, double *out)
{
    int64_t e; int v, k;
    or_zero(s);
    or_ns_diff(s);
    or_conv(s);
    for (e = 0; e < s->n; e++)
        for (v = 0; v < s->nv; v++)
            for (k = 0; k < OR_BF; k++) out[(e*s->nv + v)*OR_BF + k] = s->d[e].integ[v][k];
}

/* the viscous integrals alone: [n][4][OR_BF] for ru rv rw re */
void or_ns_visc(or_sim_t *s, double *out)
{
    int64_t e; int v, k;
    or_zero(s);
    or_ns_diff(s);
    for (e = 0; e < s->n; e++)
        for (v = 0; v < 4; v++)
            for (k = 0; k < OR_BF; k++) out[(e*4 + v)*OR_BF + k] = s->d[e].integ[v][k];
}

void or_conv_residual(or_sim_t *s, double *out)
{
    int64_t e; int v, k;
    or_zero(s);
    or_conv(s);
    for (e = 0; e < s->n; e++)
        for (v = 0; v < s->nv; v++)
            for (k = 0; k < OR_BF; k++) out[(e*s->nv + v)*OR_BF + k] = s->d[e].integ[v][k];
}

/* c -= dt * A^-1 int, src/models/euler/charm_model.c:402-433 */
static void or_update(or_sim_t *s, double dt)
{
    int64_t e;
    OR_PARFOR
    for (e = 0; e < s->n; e++) {
        or_data_t *d = &s->d[e];
        double rhs[OR_NV_MAX][4];
        int v, i;
        for (v = 0; v < s->nv; v++) or_matr_vect_mult(d->g.a_inv, d->integ[v], rhs[v]);
        for (i = 0; i < OR_BF; i++)
            for (v = 0; v < s->nv; v++) d->c[v][i] -= dt * rhs[v][i];
    }
}

/* ------------------------------------------------ Barth-Jespersen limiter
 * src/limiters/charm_limiter_bj.c:11-332 */
static void or_lim_push(or_sim_t *s, or_data_t *d, const double *u)
{
    int v, i = d->l.count;
    if (i >= 7) { snprintf(s->err, sizeof(s->err), "limiter: more than 7 entries"); return; }
    for (v = 0; v < s->nv; v++) d->l.u[v][i] = u[v];
    d->l.count++;
}

/* the two values a face contributes to its leaves' neighbour lists,
 * src/limiters/charm_limiter_bj.c:30-238: the leaves of side i receive
 * u[1 - i]; a boundary face gives its leaf u[1], the mean pushed through the
 * boundary condition.  Conforming: the cell means; hanging: the big side's
 * centre value and the volume average of the four small centre values */
static void or_lim_face_values(or_sim_t *s, const or_face_t *f, double u[2][OR_NV_MAX])
{
    or_cons_t cons[2], cj;
    int i, j, v;
    if (f->nsides != 2) {
        or_data_t *d = &s->d[f->quad[0][0]];
        double n[3];
        or_prim_t prim[2];
        or_face_normal(d, f->face[0], n);
        or_get_fields_avg(s, d, &cons[0]);
        or_cons_to_prim(&s->cfg, &cons[0], &prim[0]);
        or_bnd(&s->cfg, f->bnd, &prim[0], &prim[1], n);
        or_prim_to_cons(&s->cfg, &cons[1], &prim[1]);
        or_cons_get(s, &cons[1], u[1]);
    } else if (f->is_hanging[0] || f->is_hanging[1]) {
        for (i = 0; i < 2; i++) {
            if (f->is_hanging[i]) {
                double svol = 0.;
                for (v = 0; v < s->nv; v++) u[i][v] = 0.;
                for (j = 0; j < 4; j++) {
                    or_data_t *dj = &s->d[f->quad[i][j]];
                    double vol = dj->g.volume, uj[OR_NV_MAX];
                    svol += vol;
                    or_get_fields(s, dj, dj->g.c, &cj);
                    or_cons_get(s, &cj, uj);
                    for (v = 0; v < s->nv; v++) u[i][v] += uj[v]*vol;
                }
                for (v = 0; v < s->nv; v++) u[i][v] /= svol;
            } else {
                or_data_t *di = &s->d[f->quad[i][0]];
                or_get_fields(s, di, di->g.c, &cons[i]);
                or_cons_get(s, &cons[i], u[i]);
            }
        }
        /* species entries are stored here too; the reference leaves them
         * unwritten on this branch (charm_limiter_bj.c:174-193) */
    } else {
        for (i = 0; i < 2; i++) {
            or_get_fields_avg(s, &s->d[f->quad[i][0]], &cons[i]);
            or_cons_get(s, &cons[i], u[i]);
        }
    }
}

void or_limiter(or_sim_t *s)
{
    int64_t e, fi;
    int i, j, v, k;
    double u[2][OR_NV_MAX];
    or_cons_t cons[2], cj;
    if (!s->cfg.limiter) return;
    OR_PARFOR
    for (e = 0; e < s->n; e++) {           /* :11-27 */
        or_cons_t ce;
        double ue[OR_NV_MAX];
        or_get_fields_avg(s, &s->d[e], &ce);
        or_cons_get(s, &ce, ue);
        s->d[e].l.count = 0;
        or_lim_push(s, &s->d[e], ue);
    }
#ifdef _OPENMP
    if (omp_get_max_threads() > 1 && !or_par_setup(s)) {
        /* the face values in parallel, then every leaf pushes its faces'
         * values in face order: the serial lists */
        OR_PARFOR
        for (fi = 0; fi < s->nf; fi++)
            or_lim_face_values(s, &s->f[fi], (double (*)[OR_NV_MAX]) (s->lq + (size_t) fi * 2 * OR_NV_MAX));
        OR_PARFOR
        for (e = 0; e < s->n; e++) {
            int64_t kk;
            for (kk = s->inc_ptr[e]; kk < s->inc_ptr[e + 1]; kk++) {
                const int64_t f = s->inc[kk] >> 4;
                const int side = (int) ((s->inc[kk] >> 3) & 1);
                const double *lu = s->lq + (size_t) f * 2 * OR_NV_MAX;
                or_lim_push(s, &s->d[e], lu + (s->f[f].nsides != 2 ? 1 : 1 - side) * OR_NV_MAX);
            }
        }
    } else
#endif
    for (fi = 0; fi < s->nf; fi++) {        /* :30-238 */
        const or_face_t *f = &s->f[fi];
        or_lim_face_values(s, f, u);
        if (f->nsides != 2) {
            or_lim_push(s, &s->d[f->quad[0][0]], u[1]);
        } else if (f->is_hanging[0] || f->is_hanging[1]) {
            int hs = f->is_hanging[0] ? 0 : 1, fs = 1 - hs;
            or_lim_push(s, &s->d[f->quad[fs][0]], u[hs]);
            for (j = 0; j < 4; j++) or_lim_push(s, &s->d[f->quad[hs][j]], u[fs]);
        } else {
            for (i = 0; i < 2; i++) {
                k = (i + 1) % 2;
                or_lim_push(s, &s->d[f->quad[k][0]], u[i]);
            }
        }
    }
    OR_PARFOR
    for (e = 0; e < s->n; e++) {            /* :241-315 */
        or_data_t *d = &s->d[e];
        double umin[OR_NV_MAX], umax[OR_NV_MAX], psi[OR_NV_MAX], vtx[8][3], f[OR_NV_MAX];
        or_cons_t cj;
        int i, v, k;
        if (d->l.count != 7) { snprintf(s->err, sizeof(s->err), "limiter: count != 7"); }
        for (v = 0; v < s->nv; v++) {
            umin[v] = umax[v] = d->l.u[v][0];
            for (i = 1; i < d->l.count; i++) {
                if (umin[v] > d->l.u[v][i]) umin[v] = d->l.u[v][i];
                if (umax[v] < d->l.u[v][i]) umax[v] = d->l.u[v][i];
            }
            psi[v] = 1.;
        }
        or_leaf_vertices(s, d, vtx);
#ifdef OR_MUTANT_BJ_FACE_CENTRES
        /* planted mistake (never set by the Makefile): the six face centres
         * instead of the eight vertices, for tests/test_oracle_mutants.py */
        for (i = 0; i < 6; i++) for (k = 0; k < 3; k++) vtx[i][k] = d->g.fc[i][k];
        for (i = 0; i < 6; i++) {
#else
        for (i = 0; i < 8; i++) {
#endif
            or_get_fields(s, d, vtx[i], &cj);
            or_cons_get(s, &cj, f);
            for (v = 0; v < s->nv; v++) {
                double u0 = d->l.u[v][0], pt = 1.;
                if (fabs(f[v]-u0) > OR_EPS) {
                    if (f[v] - u0 > 0) pt = OR_MIN(1, (umax[v] - u0) / (f[v] - u0));
                    else if (f[v] - u0 < 0) pt = OR_MIN(1, (umin[v] - u0) / (f[v] - u0));
                }
                if (pt < psi[v]) psi[v] = pt;
            }
        }
        for (k = 1; k < OR_BF; k++)
            for (v = 0; v < s->nv; v++) d->c[v][k] *= psi[v];
    }
}

/* ---------------------------------------------------------- steppers */

/* charm_model_euler_timestep_single body, src/models/euler/charm_model.c:
 * 519-535 (adapt/partition/write hooks are not part of the stepper) */
void or_euler_step(or_sim_t *s, double dt)
{
    or_zero(s);
    or_conv(s);
    or_update(s, dt);
    or_limiter(s);
}

void or_ns_chem(or_sim_t *s);
void or_ns_diff(or_sim_t *s);

static void or_ns_stage(or_sim_t *s, double dt)
{
    or_ns_chem(s);
    or_zero(s);
    or_ns_diff(s);
    or_conv(s);
    or_update(s, dt);
    or_limiter(s);
}

/* charm_model_ns_timestep_single with rk_0/1/2, src/models/ns/charm_model.c:
 * 119-184, 244-266 */
void or_ns_step(or_sim_t *s, double dt)
{
    int64_t e;
    s->dt = dt;                                   /* ctx->dt of the chemistry */
    OR_PARFOR
    for (e = 0; e < s->n; e++) memcpy(s->d[e].c_old, s->d[e].c, sizeof(s->d[e].c));
    or_ns_stage(s, dt);
    or_ns_stage(s, dt);
    OR_PARFOR
    for (e = 0; e < s->n; e++) {
        or_data_t *d = &s->d[e];
        int v, i;
        for (i = 0; i < OR_BF; i++)
            for (v = 0; v < OR_FIELD_NV(s); v++) {
                d->c[v][i] *= 0.25;
                d->c[v][i] += 0.75*d->c_old[v][i];
            }
    }
    or_ns_stage(s, dt);
    OR_PARFOR
    for (e = 0; e < s->n; e++) {
        or_data_t *d = &s->d[e];
        int v, i;
        for (i = 0; i < OR_BF; i++)
            for (v = 0; v < OR_FIELD_NV(s); v++) {
                d->c[v][i] *= 2.;
                d->c[v][i] /= 3.;
                d->c[v][i] += d->c_old[v][i] / 3.;
            }
    }
}

void or_run(or_sim_t *s, double dt, int nsteps)
{
    int i;
    for (i = 0; i < nsteps; i++) {
        if (s->cfg.model == 0) or_euler_step(s, dt);
        else or_ns_step(s, dt);
    }
}

/* ---------------------------------------------------------- diagnostics */
void or_conservation_sums(const or_sim_t *s, double *out)
{
    int64_t e; int v;
    for (v = 0; v < s->nv; v++) out[v] = 0.;
    for (e = 0; e < s->n; e++)
        for (v = 0; v < s->nv; v++) out[v] += s->d[e].g.volume * s->d[e].c[v][0];
}

/* the disabled dt branch of charm_model_euler_get_dt,
 * src/models/euler/charm_model.c:15-30 (mode 0: CFL*volume/(|u|+c));
 * mode 1 uses the edge length instead of the volume */
double or_cfl_dt(const or_sim_t *s, int mode)
{
    int64_t e;
    double dt = 1.e300;
    for (e = 0; e < s->n; e++) {
        const or_data_t *d = &s->d[e];
        or_cons_t c; or_prim_t p;
        or_get_fields(s, d, d->g.c, &c);
        or_cons_to_prim(&s->cfg, &c, &p);
        double len = mode == 0 ? d->g.volume : d->g.dh[0];
        double dl = s->cfg.cfl * len / (sqrt(p.u*p.u + p.v*p.v + p.w*p.w) + p.cz);
        if (dl < dt) dt = dl;
    }
    return dt;
}
