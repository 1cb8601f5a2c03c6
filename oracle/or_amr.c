/*
 * or_amr.c -- adaptation pieces of the oracle (FP64, TEST INFRASTRUCTURE
 * ONLY): the Green-Gauss density-gradient estimator, the refine / coarsen
 * criteria and the L2 projection between old and new leaves.
 */
#include "or_internal.h"

/* zero + face pass + volume pass of charm_adapt, src/charm_amr.c:350-599:
 * grad_u = (1/V) sum_faces sum_gp w J (rho_0 + rho_1)/2 n, then the error
 * estimate err2 = |grad|^2 / 12 V^(5/3) (:12-30).  rho_c is the density at
 * the leaf centre (the value the coarsening test uses, :116-118). */
void or_amr_indicator(or_sim_t *s, double *err2, double *rho_c)
{
    int64_t e, fi;
    int i, j, igp, d;
    for (e = 0; e < s->n; e++) s->d[e].grad_u[0] = s->d[e].grad_u[1] = s->d[e].grad_u[2] = 0.;
    for (fi = 0; fi < s->nf; fi++) {
        const or_face_t *f = &s->f[fi];
        double n[3], xs[2][3];
        or_cons_t cons[2];
        or_prim_t prim[2];
        if (f->nsides != 2) {                                   /* :362-423 */
            or_data_t *d = &s->d[f->quad[0][0]];
            int face = f->face[0];
            double l[3];
            for (i = 0; i < 3; i++) { n[i] = d->g.n[face][i]; l[i] = d->g.fc[face][i] - d->g.c[i]; }
            if (n[0]*l[0] + n[1]*l[1] + n[2]*l[2] < 0) for (i = 0; i < 3; i++) n[i] *= -1.0;
            for (igp = 0; igp < OR_FGP; igp++) {
                const double *x = d->g.face_gp[face][igp];
                double gw = d->g.face_gw[face][igp], gj = d->g.face_gj[face][igp];
                int v, k;
                double r[OR_NV_MAX];
                for (v = 0; v < s->nv; v++) {
                    r[v] = 0.;
                    for (k = 0; k < OR_BF; k++) r[v] += d->c[v][k] * or_bf(&d->g, x, k);
                }
                cons[0].ru = r[0]; cons[0].rv = r[1]; cons[0].rw = r[2]; cons[0].re = r[3];
                for (k = 0; k < s->cfg.n_comp; k++) cons[0].rc[k] = r[4+k];
                cons[0].mat_id = 0;
                or_cons_to_prim(&s->cfg, &cons[0], &prim[0]);
                or_bnd(&s->cfg, f->bnd, &prim[0], &prim[1], n);
                {
                    double qr = 0.5*(prim[0].r+prim[1].r), bfv = gw * gj;
                    int dd;
                    for (dd = 0; dd < 3; dd++) s->d[f->quad[0][0]].grad_u[dd] += qr * bfv * n[dd];
                }
            }
            continue;
        }
        {                                                       /* :426-549 */
            int nj = (f->is_hanging[0] || f->is_hanging[1]) ? 4 : 1;
            int h_side = f->is_hanging[0] ? 0 : (f->is_hanging[1] ? 1 : 0);
            for (j = 0; j < nj; j++) {
                or_data_t *ud[2];
                double l[3];
                for (i = 0; i < 2; i++) ud[i] = &s->d[f->is_hanging[i] ? f->quad[i][j] : f->quad[i][0]];
                for (i = 0; i < 3; i++) { n[i] = ud[0]->g.n[f->face[0]][i]; l[i] = ud[0]->g.fc[f->face[0]][i] - ud[0]->g.c[i]; }
                if (n[0]*l[0] + n[1]*l[1] + n[2]*l[2] < 0) for (i = 0; i < 3; i++) n[i] *= -1.0;
                for (igp = 0; igp < OR_FGP; igp++) {
                    const double *x = ud[h_side]->g.face_gp[f->face[h_side]][igp];
                    double gw = ud[h_side]->g.face_gw[f->face[h_side]][igp];
                    double gj = ud[h_side]->g.face_gj[f->face[h_side]][igp];
                    for (i = 0; i < 2; i++) {
                        int v, k;
                        double r[OR_NV_MAX];
                        xs[i][0] = x[0]; xs[i][1] = x[1]; xs[i][2] = x[2];
                        if (f->wrap_axis >= 0 && i != h_side) xs[i][f->wrap_axis] += (i == 1) ? -1.0 : 1.0;
                        for (v = 0; v < s->nv; v++) {
                            r[v] = 0.;
                            for (k = 0; k < OR_BF; k++) r[v] += ud[i]->c[v][k] * or_bf(&ud[i]->g, xs[i], k);
                        }
                        cons[i].ru = r[0]; cons[i].rv = r[1]; cons[i].rw = r[2]; cons[i].re = r[3];
                        for (k = 0; k < s->cfg.n_comp; k++) cons[i].rc[k] = r[4+k];
                        cons[i].mat_id = 0;
                        or_cons_to_prim(&s->cfg, &cons[i], &prim[i]);
                    }
                    for (d = 0; d < 3; d++)
                        for (i = 0; i < 2; i++)
                            ud[i]->grad_u[d] += (i ? -1. : 1.) * gw * gj * 0.5*(prim[0].r+prim[1].r) * n[d];
                }
            }
        }
    }
    for (e = 0; e < s->n; e++) {                                 /* :564-573, :12-30 */
        or_data_t *d = &s->d[e];
        double vol = d->g.volume, diff2 = 0.;
        or_cons_t c;
        or_prim_t p;
        int v, k;
        for (i = 0; i < 3; i++) d->grad_u[i] /= vol;
        for (i = 0; i < 3; i++) diff2 += d->grad_u[i] * d->grad_u[i];
        err2[e] = diff2 * (1. / 12.) * exp(log(vol)*5./3.);
        c.ru = d->c[0][0]; c.rv = d->c[1][0]; c.rw = d->c[2][0]; c.re = d->c[3][0];
        for (k = 0; k < s->cfg.n_comp; k++) c.rc[k] = d->c[4+k][0];
        c.mat_id = 0;
        (void) v;
        or_cons_to_prim(&s->cfg, &c, &p);
        rho_c[e] = p.r;
    }
}

/* refine / coarsen decision per leaf from the estimator:
 * refine (+1) if err2 > max_err^2 V and level < max_level
 * (charm_refine_err_estimate, src/charm_amr.c:32-48, with p4est_refine_ext's
 * maxlevel); a complete family of 8 leaves is coarsened (-1 on each member)
 * if no member refines, level > min_level, every member has
 * err2 <= max_err^2 V and sum(err2 + (mean rho - rho_c)^2 V) < max_err^2 8V
 * (charm_coarsen_err_estimate, :90-142). */
void or_amr_flags(const or_sim_t *s, const double *err2, const double *rho_c, double max_err,
                  int min_level, int max_level, int8_t *flags)
{
    int64_t e;
    double ge2 = max_err * max_err;
    for (e = 0; e < s->n; e++) {
        const or_data_t *d = &s->d[e];
        flags[e] = (err2[e] > ge2 * d->g.volume && d->quad[3] < max_level) ? 1 : 0;
    }
    for (e = 0; e + 8 <= s->n; e++) {
        const int32_t *q = s->d[e].quad;
        int l = q[3], c, ok;
        int32_t ph;
        if (l <= min_level || l == 0) continue;
        ph = OR_ROOT_LEN >> (l - 1);
        if (q[0] % ph || q[1] % ph || q[2] % ph) continue;
        ok = 1;
        for (c = 0; c < 8 && ok; c++) {
            const int32_t *r = s->d[e + c].quad;
            int32_t hc = ph / 2;
            ok = r[3] == l && r[0] == q[0] + (c & 1) * hc && r[1] == q[1] + ((c >> 1) & 1) * hc &&
                 r[2] == q[2] + ((c >> 2) & 1) * hc && flags[e + c] == 0;
        }
        if (!ok) continue;
        {
            double vol = s->d[e].g.volume, parentu = 0., tot = 0.;
            for (c = 0; c < 8; c++) parentu += rho_c[e + c] / 8;
            for (c = 0; c < 8 && ok; c++) {
                if (err2[e + c] > ge2 * vol) ok = 0;
                tot += err2[e + c];
                tot += (parentu - rho_c[e + c]) * (parentu - rho_c[e + c]) * vol;
            }
            if (ok && tot < ge2 * (vol * 8))
                for (c = 0; c < 8; c++) flags[e + c] = -1;
        }
        e += 7;
    }
}

/* refinement branch of charm_replace_quads (src/charm_amr.c:258-306):
 * c_new = A_new^-1 int_new phi_new phi_src c_src, by the new leaf's 8-point
 * rule */
static void refine_proj(int nv, const or_geom_t *gs, const double *cs, const or_geom_t *gn, double *out)
{
    double ar[4][4], rhs[4];
    int v, m, k, igp;
    for (m = 0; m < 4; m++) for (k = 0; k < 4; k++) {
        ar[m][k] = 0.;
        for (igp = 0; igp < OR_QGP; igp++)
            ar[m][k] += gn->quad_gw[igp]*gn->quad_gj[igp]
                      * or_bf(gn, gn->quad_gp[igp], m) * or_bf(gs, gn->quad_gp[igp], k);
    }
    for (v = 0; v < nv; v++) {
        or_matr_vect_mult(ar, &cs[v*4], rhs);
        or_matr_vect_mult((double (*)[4]) gn->a_inv, rhs, &out[v*4]);
    }
}

/* L2 projection of charm_replace_quads, src/charm_amr.c:206-307, applied in
 * the order the reference's adapt applies it (refine, coarsen, then the
 * balance's refinements, :601-608).  map[i] >= 0: the old leaf containing
 * new leaf i (same level: the coefficients are kept; finer: refinement
 * projection, applied once since projecting a P1 field to a sub-cell is
 * exact at any depth).  map[i] = -1 - f: new leaf i is the parent of the old
 * family f..f+7, coarsened by A_P^-1 sum_children int_child phi_P phi_child
 * c_child, or lies inside that parent after the balance refined it again
 * (the parent's coefficients are then projected onto it).  States are
 * [n][nv][4]. */
void or_amr_project(int nv, int64_t n_old, const int32_t *old_leaves, const double *u_old,
                    int64_t n_new, const int32_t *new_leaves, const int64_t *map, double *u_new)
{
    int64_t i;
    int v, m, k, igp, c;
    (void) n_old;
    for (i = 0; i < n_new; i++) {
        or_geom_t gn;
        or_geom_calc(&new_leaves[4*i], &gn);
        if (map[i] >= 0) {
            or_geom_t gs;
            if (old_leaves[4*map[i]+3] == new_leaves[4*i+3]) {
                memcpy(&u_new[i*nv*4], &u_old[map[i]*nv*4], (size_t) nv * 4 * sizeof(double));
                continue;
            }
            or_geom_calc(&old_leaves[4*map[i]], &gs);
            refine_proj(nv, &gs, &u_old[map[i]*nv*4], &gn, &u_new[i*nv*4]);
        } else {                                             /* coarsening, :219-256 */
            const int64_t f0 = -1 - map[i];
            const int32_t pl = old_leaves[4*f0+3] - 1, len = OR_ROOT_LEN >> pl;
            const int32_t pq[4] = {(old_leaves[4*f0] / len) * len, (old_leaves[4*f0+1] / len) * len,
                                   (old_leaves[4*f0+2] / len) * len, pl};
            or_geom_t gp;
            double rhs[OR_NV_MAX][4], cp[OR_NV_MAX * 4], ar[4][4];
            or_geom_calc(pq, &gp);
            for (v = 0; v < nv; v++) for (k = 0; k < 4; k++) rhs[v][k] = 0.;
            for (c = 0; c < 8; c++) {
                or_geom_t gc;
                const double *cc = &u_old[(f0 + c)*nv*4];
                double cr[4];
                or_geom_calc(&old_leaves[4*(f0 + c)], &gc);
                for (m = 0; m < 4; m++) for (k = 0; k < 4; k++) {
                    ar[m][k] = 0.;
                    for (igp = 0; igp < OR_QGP; igp++)
                        ar[m][k] += gc.quad_gw[igp]*gc.quad_gj[igp]
                                  * or_bf(&gp, gc.quad_gp[igp], m) * or_bf(&gc, gc.quad_gp[igp], k);
                }
                for (v = 0; v < nv; v++) {
                    or_matr_vect_mult(ar, &cc[v*4], cr);
                    for (k = 0; k < 4; k++) rhs[v][k] += cr[k];
                }
            }
            for (v = 0; v < nv; v++) or_matr_vect_mult(gp.a_inv, rhs[v], &cp[v*4]);
            if (new_leaves[4*i+3] == pl) memcpy(&u_new[i*nv*4], cp, (size_t) nv * 4 * sizeof(double));
            else refine_proj(nv, &gp, cp, &gn, &u_new[i*nv*4]);
        }
    }
}
