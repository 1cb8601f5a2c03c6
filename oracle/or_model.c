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
    for (e = 0; e < n_leaves; e++) {
        or_data_t *d = &s->d[e];
        memcpy(d->quad, &leaves[4*e], 4*sizeof(int32_t));
        or_geom_calc(d->quad, &d->g);
        d->grav[0] = cfg->grav[0]; d->grav[1] = cfg->grav[1]; d->grav[2] = cfg->grav[2];
    }
    or_build_faces(s);
    return s;
}

void or_destroy(or_sim_t *s)
{
    if (!s) return;
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
    double r[OR_NV_MAX];
    for (v = 0; v < s->nv; v++) {
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
    double r[OR_NV_MAX];
    for (v = 0; v < s->nv; v++) r[v] = 0.;
    for (igp = 0; igp < OR_QGP; igp++) {
        const double *gx = d->g.quad_gp[igp];
        double gjw = d->g.quad_gj[igp]*d->g.quad_gw[igp]/d->g.volume;
        for (v = 0; v < s->nv; v++) {
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
    for (e = 0; e < s->n; e++) memset(s->d[e].integ, 0, sizeof(s->d[e].integ));
}

/* charm_timestep_conv, src/models/euler/charm_model.c:388-396 */
static void or_conv(or_sim_t *s)
{
    int64_t e, i;
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
    int64_t e; int v, i;
    double rhs[OR_NV_MAX][4];
    for (e = 0; e < s->n; e++) {
        or_data_t *d = &s->d[e];
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

void or_limiter(or_sim_t *s)
{
    int64_t e, fi;
    int i, j, v, k;
    double u[2][OR_NV_MAX];
    or_cons_t cons[2], cj;
    if (!s->cfg.limiter) return;
    for (e = 0; e < s->n; e++) {           /* :11-27 */
        or_get_fields_avg(s, &s->d[e], &cons[0]);
        or_cons_get(s, &cons[0], u[0]);
        s->d[e].l.count = 0;
        or_lim_push(s, &s->d[e], u[0]);
    }
    for (fi = 0; fi < s->nf; fi++) {        /* :30-238 */
        const or_face_t *f = &s->f[fi];
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
            or_lim_push(s, d, u[1]);
        } else if (f->is_hanging[0] || f->is_hanging[1]) {
            int hs = f->is_hanging[0] ? 0 : 1, fs = 1 - hs;
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
            /* species entries are stored here too; the reference leaves
             * them unwritten on this branch (charm_limiter_bj.c:174-193) */
            or_lim_push(s, &s->d[f->quad[fs][0]], u[hs]);
            for (j = 0; j < 4; j++) or_lim_push(s, &s->d[f->quad[hs][j]], u[fs]);
        } else {
            for (i = 0; i < 2; i++) {
                or_get_fields_avg(s, &s->d[f->quad[i][0]], &cons[i]);
                or_cons_get(s, &cons[i], u[i]);
                k = (i + 1) % 2;
                or_lim_push(s, &s->d[f->quad[k][0]], u[i]);
            }
        }
    }
    for (e = 0; e < s->n; e++) {            /* :241-315 */
        or_data_t *d = &s->d[e];
        double umin[OR_NV_MAX], umax[OR_NV_MAX], psi[OR_NV_MAX], vtx[8][3], f[OR_NV_MAX];
        if (d->l.count != 7) { snprintf(s->err, sizeof(s->err), "limiter: count != 7"); }
        for (v = 0; v < s->nv; v++) {
            umin[v] = umax[v] = d->l.u[v][0];
            for (i = 1; i < d->l.count; i++) {
                if (umin[v] > d->l.u[v][i]) umin[v] = d->l.u[v][i];
                if (umax[v] < d->l.u[v][i]) umax[v] = d->l.u[v][i];
            }
            psi[v] = 1.;
        }
        or_quad_vertices(d->quad, vtx);
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
    int64_t e; int v, i;
    s->dt = dt;                                   /* ctx->dt of the chemistry */
    for (e = 0; e < s->n; e++) memcpy(s->d[e].c_old, s->d[e].c, sizeof(s->d[e].c));
    or_ns_stage(s, dt);
    or_ns_stage(s, dt);
    for (e = 0; e < s->n; e++) {
        or_data_t *d = &s->d[e];
        for (i = 0; i < OR_BF; i++)
            for (v = 0; v < s->nv; v++) {
                d->c[v][i] *= 0.25;
                d->c[v][i] += 0.75*d->c_old[v][i];
            }
    }
    or_ns_stage(s, dt);
    for (e = 0; e < s->n; e++) {
        or_data_t *d = &s->d[e];
        for (i = 0; i < OR_BF; i++)
            for (v = 0; v < s->nv; v++) {
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
