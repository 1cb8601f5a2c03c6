/*
 * or_ns.c -- Navier-Stokes viscous and chemistry terms of the oracle
 * (FP64, TEST INFRASTRUCTURE ONLY).
 *
 * Restates, callback by callback:
 *   multicomponent diffusion coefficients   src/models/ns/charm_model_ns_diff.c:18-90
 *   auxiliary tau / q (BR1 gradient pass)   src/models/ns/charm_model_ns_diff_grad.c:21-507
 *   viscous volume and surface integrals    src/models/ns/charm_model_ns_diff_integrals.c:21-408
 *   transport mu, k                          src/charm_globals.c:509-561
 *   tau / q traces                           src/charm_base_func.c:343-371
 *   chemistry                                src/models/ns/charm_model_ns_chem.c:18-202
 * on the oracle's own face traversal.  Periodic faces shift the far side's
 * evaluation point as in the convective pass.
 *
 * BR2 option (visc_br2_eta > 0; the reference has BR1 only).  Bassi and
 * Rebay's second scheme with the local lifting of each interior (sub)face,
 * written in the reference's tau / q projection.  The gradient pass's face
 * term at a point is Q(state) . n with the face-averaged coefficients (mu,
 * k, D); BR1 takes the averaged state.  With Q(own_i), the same term from
 * side i's own trace, side i's lifting of face f is
 *   l_i,f = A_i^-1 sum_points sgn_i [Q(avg) - Q(own_i)] phi_i wJ,
 * and tau_h = tau_own + sum_f l_f, where tau_own takes every interior face
 * term from the leaf's own trace (volume and boundary terms as BR1).  The
 * volume integrals keep tau_h; an interior face uses tau_own + eta l_f on
 * each side in the BR1 surface flux (averages of tau, tau u and q).  The
 * boundary terms (the no-slip wall's) are the reference's.
 */
#include "or_internal.h"

/* ------------------------------------------------------------ transport */

/* charm_get_visc_mu, src/charm_globals.c:515-538: mole-fraction weighted */
static double or_visc_mu(const or_sim_t *s, const or_data_t *d, const double x[3])
{
    or_cons_t c; or_prim_t p;
    double mu = 0., sm = 0.;
    int i;
    or_get_fields(s, d, x, &c);
    or_cons_to_prim(&s->cfg, &c, &p);
    for (i = 0; i < s->cfg.n_comp; i++) {
        double cm = p.c[i] / s->cfg.comp[i].m;
        sm += cm;
        mu += cm * or_comp_calc_ml(&s->cfg.comp[i], p.t);
    }
    return mu / sm;
}

/* charm_get_heat_k, src/charm_globals.c:541-561: mass-fraction weighted */
static double or_heat_k(const or_sim_t *s, const or_data_t *d, const double x[3])
{
    or_cons_t c; or_prim_t p;
    double kt = 0.;
    int i;
    or_get_fields(s, d, x, &c);
    or_cons_to_prim(&s->cfg, &c, &p);
    for (i = 0; i < s->cfg.n_comp; i++) kt += p.c[i] * or_comp_calc_kp(&s->cfg.comp[i], p.t);
    return kt;
}

/* charm_get_visc_tau / charm_get_heat_q, src/charm_base_func.c:343-371;
 * tau order xx yy zz xy xz yz */
static void or_tau_at(const or_data_t *d, const double x[3], double t[6])
{
    int i, k;
    for (i = 0; i < 6; i++) {
        t[i] = 0.;
        for (k = 0; k < OR_BF; k++) t[i] += d->tau[i][k] * or_bf(&d->g, x, k);
    }
}

static void or_q_at(const or_data_t *d, const double x[3], double q[3])
{
    int i, k;
    for (i = 0; i < 3; i++) {
        q[i] = 0.;
        for (k = 0; k < OR_BF; k++) q[i] += d->q[i][k] * or_bf(&d->g, x, k);
    }
}

/* value of one field (OR_BF coefficients) of leaf d at x */
static double or_coef_at(const or_data_t *d, const double c[OR_BF], const double x[3])
{
    double v = 0.;
    int k;
    for (k = 0; k < OR_BF; k++) v += c[k] * or_bf(&d->g, x, k);
    return v;
}

/* ------------------------------------------------------------ diffusion */

/* src/models/ns/charm_model_ns_diff.c:18-71 */
static void or_ns_diffusion(or_sim_t *s)
{
    const int nc = s->cfg.n_comp;
    int64_t e;
    int i, j;
    if (nc < 2) return;                                           /* :74-80 */
    for (e = 0; e < s->n; e++) {
        or_data_t *d = &s->d[e];
        or_cons_t cons; or_prim_t prim;
        double x[OR_NC_MAX], sm = 0., pabs;
        or_get_fields(s, d, d->g.c, &cons);
        or_cons_to_prim(&s->cfg, &cons, &prim);
        pabs = prim.p / 101325.0;
        memset(d->d, 0, sizeof(d->d));
        if (!s->cfg.use_diff) continue;
        for (i = 0; i < nc; i++) { x[i] = prim.c[i] / s->cfg.comp[i].m; sm += x[i]; }
        for (i = 0; i < nc; i++) x[i] /= sm;
        for (i = 0; i < nc; i++) {
            const or_comp_t *ci = &s->cfg.comp[i];
            double s_xd = 0.;
            if (fabs(prim.c[i] - 1.) <= OR_EPS) { d->d[i] = 0.; continue; }
            for (j = 0; j < nc; j++) {
                const or_comp_t *cj = &s->cfg.comp[j];
                double td = prim.t / sqrt(ci->ek * cj->ek);
                double wd = 1.06036 / pow(td, 0.1561) + 0.1930 / exp(0.47635 * td) +
                            1.03587 / exp(1.52996 * td) + 1.76474 / exp(3.89411 * td);
                double dij = (2.628e-7) * sqrt(0.5 * (1.0 / (1000.0 * ci->m) + 1.0 / (1000.0 * cj->m))) *
                             sqrt(prim.t * prim.t * prim.t) /
                             (pabs * pow((0.5 * (ci->sig + cj->sig)), 2.0) * wd);
                s_xd += i == j ? 0. : x[j] / dij;
            }
            d->d[i] = (1. - x[i]) / s_xd;
        }
    }
}

/* ------------------------------------------------ gradient (tau, q) pass */

static void or_grad_add(or_data_t *d, int ibf, const double qt6[6], const double qq[3], double bfv)
{
    int i;
    for (i = 0; i < 6; i++) d->int_tau[i][ibf] += qt6[i] * bfv;
    for (i = 0; i < 3; i++) d->int_q[i][ibf] -= qq[i] * bfv;
}

/* the same into the BR2 own-trace accumulators */
static void or_grad_add_own(or_data_t *d, int ibf, const double qt6[6], const double qq[3], double bfv)
{
    int i;
    for (i = 0; i < 6; i++) d->int_tau_own[i][ibf] += qt6[i] * bfv;
    for (i = 0; i < 3; i++) d->int_q_own[i][ibf] -= qq[i] * bfv;
}

/* src/models/ns/charm_model_ns_diff_grad.c:21-74 */
static void or_grad_volume(or_sim_t *s, or_data_t *d)
{
    const double t_ref = s->cfg.t_ref;
    int ibf, igp, i;
    for (ibf = 0; ibf < OR_BF; ibf++) {
        for (igp = 0; igp < OR_QGP; igp++) {
            const double *x = d->g.quad_gp[igp];
            double kt = or_heat_k(s, d, x), lambda = 0., mu = or_visc_mu(s, d, x);
            double lp = (lambda + 4.*mu/3.), lm = (lambda - 2.*mu/3.);
            double wj = d->g.quad_gj[igp] * d->g.quad_gw[igp];
            double phi_x = or_bf_dx(&d->g, ibf) * wj, phi_y = or_bf_dy(&d->g, ibf) * wj;
            double phi_z = or_bf_dz(&d->g, ibf) * wj;
            double tx, ty, tz;
            or_cons_t c; or_prim_t p;
            or_get_fields(s, d, x, &c);
            or_cons_to_prim(&s->cfg, &c, &p);
            d->int_tau[0][ibf] -= (lp*p.u*phi_x + lm*p.v*phi_y + lm*p.w*phi_z);
            d->int_tau[1][ibf] -= (lm*p.u*phi_x + lp*p.v*phi_y + lm*p.w*phi_z);
            d->int_tau[2][ibf] -= (lm*p.u*phi_x + lm*p.v*phi_y + lp*p.w*phi_z);
            d->int_tau[3][ibf] -= mu*(p.u*phi_y + p.v*phi_x);
            d->int_tau[4][ibf] -= mu*(p.w*phi_x + p.u*phi_z);
            d->int_tau[5][ibf] -= mu*(p.w*phi_y + p.v*phi_z);
            tx = kt*p.t*phi_x; ty = kt*p.t*phi_y; tz = kt*p.t*phi_z;
            for (i = 0; i < s->cfg.n_comp; i++) {
                double h = or_comp_calc_enthalpy(&s->cfg.comp[i], p.t, t_ref);
                tx += h*p.r*d->d[i]*p.c[i]*phi_x;
                ty += h*p.r*d->d[i]*p.c[i]*phi_y;
                tz += h*p.r*d->d[i]*p.c[i]*phi_z;
            }
            d->int_q[0][ibf] += tx; d->int_q[1][ibf] += ty; d->int_q[2][ibf] += tz;
        }
    }
}

/* face terms of the gradient pass at one point from the averaged state:
 * q6 = (lp fu n0 + lm fv n1 + lm fw n2, ..., mu (fu n1 + fv n0), ...),
 * qq = kt ft n + sum_i h_i(ft) fr d_i fc_i n */
static void or_grad_face_terms(const or_sim_t *s, double lp, double lm, double mu, double kt,
                               double fu, double fv, double fw, double ft, double fr,
                               const double *fc, const double *dm, const double n[3],
                               double q6[6], double qq[3])
{
    int i;
    q6[0] = lp*fu*n[0] + lm*fv*n[1] + lm*fw*n[2];
    q6[1] = lm*fu*n[0] + lp*fv*n[1] + lm*fw*n[2];
    q6[2] = lm*fu*n[0] + lm*fv*n[1] + lp*fw*n[2];
    q6[3] = mu*(fu*n[1] + fv*n[0]);
    q6[4] = mu*(fw*n[0] + fu*n[2]);
    q6[5] = mu*(fw*n[1] + fv*n[2]);
    qq[0] = kt*ft*n[0]; qq[1] = kt*ft*n[1]; qq[2] = kt*ft*n[2];
    for (i = 0; i < s->cfg.n_comp; i++) {
        double h = or_comp_calc_enthalpy(&s->cfg.comp[i], ft, s->cfg.t_ref);
        qq[0] += h*fr*dm[i]*fc[i]*n[0];
        qq[1] += h*fr*dm[i]*fc[i]*n[1];
        qq[2] += h*fr*dm[i]*fc[i]*n[2];
    }
}

/* src/models/ns/charm_model_ns_diff_grad.c:82-197 */
static void or_grad_bnd(or_sim_t *s, const or_face_t *f)
{
    or_data_t *d = &s->d[f->quad[0][0]];
    const int face = f->face[0];
    double n[3];
    int igp, ibf, i;
    or_face_normal(d, face, n);
    for (igp = 0; igp < OR_FGP; igp++) {
        const double *x = d->g.face_gp[face][igp];
        double gw = d->g.face_gw[face][igp], gj = d->g.face_gj[face][igp];
        double lambda = 0., mu = or_visc_mu(s, d, x), lp, lm, kt = or_heat_k(s, d, x);
        double fu, fv, fw, ft, fr, fc[OR_NC_MAX], q6[6], qq[3];
        or_cons_t cons; or_prim_t prim[2];
        lp = (lambda + 4.*mu/3.); lm = (lambda - 2.*mu/3.);
        or_get_fields(s, d, x, &cons);
        or_cons_to_prim(&s->cfg, &cons, &prim[0]);
        or_bnd(&s->cfg, f->bnd, &prim[0], &prim[1], n);
        if (s->cfg.bnd_type[f->bnd] == OR_BND_WALL_NO_SLIP) {
            fu = fv = fw = 0.;
            ft = s->cfg.bnd_param[f->bnd][0];
        } else {
            fu = (prim[0].u + prim[1].u)*0.5;
            fv = (prim[0].v + prim[1].v)*0.5;
            fw = (prim[0].w + prim[1].w)*0.5;
            ft = (prim[0].t + prim[1].t)*0.5;
        }
        fr = (prim[0].r + prim[1].r)*0.5;
        for (i = 0; i < s->cfg.n_comp; i++) fc[i] = (prim[0].c[i] + prim[1].c[i])*0.5;
        or_grad_face_terms(s, lp, lm, mu, kt, fu, fv, fw, ft, fr, fc, d->d, n, q6, qq);
        for (ibf = 0; ibf < OR_BF; ibf++) {
            or_grad_add(d, ibf, q6, qq, or_bf(&d->g, x, ibf) * gw * gj);
            if (s->cfg.visc_br2_eta > 0.) or_grad_add_own(d, ibf, q6, qq, or_bf(&d->g, x, ibf) * gw * gj);
        }
    }
}

/* one face point of an interior face, src/models/ns/charm_model_ns_diff_grad.c:
 * 253-330 (hanging) and 358-429 (conforming) */
/* one interior face point of the gradient pass: both traces, the averaged
 * coefficients and the face terms q6, qq from the averaged state */
typedef struct {
    or_prim_t prim[2];
    double    lp, lm, fmu, kt, dm[OR_NC_MAX];
    double    q6[6], qq[3];
} or_gpt_t;

static void or_grad_point_terms(const or_sim_t *s, or_data_t *ud[2], const double xs[2][3],
                                const double n[3], or_gpt_t *t)
{
    or_cons_t cons[2];
    double mu[2], lambda[2], kt = 0., flambda, fr, ft, fu, fv, fw, fc[OR_NC_MAX];
    int i;
    for (i = 0; i < 2; i++) {
        or_get_fields(s, ud[i], xs[i], &cons[i]);
        or_cons_to_prim(&s->cfg, &cons[i], &t->prim[i]);
        lambda[i] = 0.;
        mu[i] = or_visc_mu(s, ud[i], xs[i]);
        kt += or_heat_k(s, ud[i], xs[i]);
    }
    t->fmu = 0.5*(mu[0] + mu[1]);
    flambda = 0.5*(lambda[0] + lambda[1]);
    t->lp = (flambda + 4.*t->fmu/3.); t->lm = (flambda - 2.*t->fmu/3.);
    t->kt = kt * 0.5;
    fr = (t->prim[0].r + t->prim[1].r)*0.5;
    ft = (t->prim[0].t + t->prim[1].t)*0.5;
    fu = (t->prim[0].u + t->prim[1].u)*0.5;
    fv = (t->prim[0].v + t->prim[1].v)*0.5;
    fw = (t->prim[0].w + t->prim[1].w)*0.5;
    for (i = 0; i < s->cfg.n_comp; i++) {
        fc[i] = (t->prim[0].c[i] + t->prim[1].c[i])*0.5;
        t->dm[i] = (ud[0]->d[i] + ud[1]->d[i])*0.5;
    }
    or_grad_face_terms(s, t->lp, t->lm, t->fmu, t->kt, fu, fv, fw, ft, fr, fc, t->dm, n, t->q6, t->qq);
}

/* BR2: the face terms from side i's own trace, at the averaged coefficients */
static void or_grad_own_terms(const or_sim_t *s, const or_gpt_t *t, int i, const double n[3],
                              double q6[6], double qq[3])
{
    const or_prim_t *p = &t->prim[i];
    or_grad_face_terms(s, t->lp, t->lm, t->fmu, t->kt, p->u, p->v, p->w, p->t, p->r, p->c, t->dm, n, q6, qq);
}

static void or_grad_point(or_sim_t *s, or_data_t *ud[2], const double xs[2][3], const double n[3],
                          double wj)
{
    or_gpt_t t;
    double q6o[2][6], qqo[2][3];
    int i, ibf;
    const int br2 = s->cfg.visc_br2_eta > 0.;
    or_grad_point_terms(s, ud, xs, n, &t);
    if (br2)
        for (i = 0; i < 2; i++) or_grad_own_terms(s, &t, i, n, q6o[i], qqo[i]);
    for (ibf = 0; ibf < OR_BF; ibf++)
        for (i = 0; i < 2; i++) {
            const double bfv = (i ? -1. : 1.) * or_bf(&ud[i]->g, xs[i], ibf) * wj;
            or_grad_add(ud[i], ibf, t.q6, t.qq, bfv);
            if (br2) or_grad_add_own(ud[i], ibf, q6o[i], qqo[i], bfv);
        }
}

/* visits the points of an interior face (conforming, or the four sub-faces
 * of a hanging one, points on the small side) with fn */
typedef void (*or_point_fn)(or_sim_t *s, or_data_t *ud[2], const double xs[2][3], const double n[3],
                            double wj);

static void or_inner_points(or_sim_t *s, const or_face_t *f, or_point_fn fn)
{
    int i, j, igp, nj = (f->is_hanging[0] || f->is_hanging[1]) ? 4 : 1, h_side = 0;
    for (i = 0; i < 2; i++) if (f->is_hanging[i]) h_side = i;
    for (j = 0; j < nj; j++) {
        or_data_t *ud[2];
        double n[3], xs[2][3];
        for (i = 0; i < 2; i++) ud[i] = &s->d[f->is_hanging[i] ? f->quad[i][j] : f->quad[i][0]];
        or_face_normal(ud[0], f->face[0], n);
        for (igp = 0; igp < OR_FGP; igp++) {
            const double *x = ud[h_side]->g.face_gp[f->face[h_side]][igp];
            double wj = ud[h_side]->g.face_gw[f->face[h_side]][igp] * ud[h_side]->g.face_gj[f->face[h_side]][igp];
            for (i = 0; i < 2; i++) or_shift(f, i, h_side, x, xs[i]);
            fn(s, ud, (const double (*)[3]) xs, n, wj);
        }
    }
}

/* src/models/ns/charm_model_ns_diff_grad.c:485-507 */
static void or_ns_grad(or_sim_t *s)
{
    int64_t e, i;
    int k;
    for (e = 0; e < s->n; e++) {
        memset(s->d[e].int_tau, 0, sizeof(s->d[e].int_tau));
        memset(s->d[e].int_q, 0, sizeof(s->d[e].int_q));
    }
    for (e = 0; e < s->n; e++) or_grad_volume(s, &s->d[e]);
    if (s->cfg.visc_br2_eta > 0.)                      /* BR2: the volume terms are shared */
        for (e = 0; e < s->n; e++) {
            memcpy(s->d[e].int_tau_own, s->d[e].int_tau, sizeof(s->d[e].int_tau));
            memcpy(s->d[e].int_q_own, s->d[e].int_q, sizeof(s->d[e].int_q));
        }
    for (i = 0; i < s->nf; i++) {
        if (s->f[i].nsides != 2) or_grad_bnd(s, &s->f[i]);
        else or_inner_points(s, &s->f[i], or_grad_point);
    }
    for (e = 0; e < s->n; e++) {                                   /* :446-460 */
        or_data_t *d = &s->d[e];
        for (k = 0; k < 6; k++) or_matr_vect_mult(d->g.a_inv, d->int_tau[k], d->tau[k]);
        for (k = 0; k < 3; k++) or_matr_vect_mult(d->g.a_inv, d->int_q[k], d->q[k]);
        if (s->cfg.visc_br2_eta > 0.) {
            for (k = 0; k < 6; k++) or_matr_vect_mult(d->g.a_inv, d->int_tau_own[k], d->tau_own[k]);
            for (k = 0; k < 3; k++) or_matr_vect_mult(d->g.a_inv, d->int_q_own[k], d->q_own[k]);
        }
    }
}

/* ---------------------------------------------------- viscous integrals */

/* src/models/ns/charm_model_ns_diff_integrals.c:21-55 */
static void or_visc_volume(or_sim_t *s, or_data_t *d)
{
    int ibf, igp;
    for (ibf = 0; ibf < OR_BF; ibf++) {
        for (igp = 0; igp < OR_QGP; igp++) {
            const double *x = d->g.quad_gp[igp];
            double wj = d->g.quad_gj[igp] * d->g.quad_gw[igp], t[6];
            double phi_x = or_bf_dx(&d->g, ibf) * wj, phi_y = or_bf_dy(&d->g, ibf) * wj;
            double phi_z = or_bf_dz(&d->g, ibf) * wj, phi = or_bf(&d->g, x, ibf) * wj;
            or_cons_t c; or_prim_t p;
            or_get_fields(s, d, x, &c);
            or_cons_to_prim(&s->cfg, &c, &p);
            or_tau_at(d, x, t);
            /* t: xx yy zz xy xz yz */
            d->integ[0][ibf] += (t[0]*phi_x + t[3]*phi_y + t[4]*phi_z);
            d->integ[1][ibf] += (t[3]*phi_x + t[1]*phi_y + t[5]*phi_z);
            d->integ[2][ibf] += (t[4]*phi_x + t[5]*phi_y + t[2]*phi_z);
            d->integ[3][ibf] += (t[0]*p.u + t[3]*p.v + t[4]*p.w)*phi_x;
            d->integ[3][ibf] += (t[3]*p.u + t[1]*p.v + t[5]*p.w)*phi_y;
            d->integ[3][ibf] += (t[4]*p.u + t[5]*p.v + t[2]*p.w)*phi_z;
            d->integ[3][ibf] += d->chem_rhs*phi;
        }
    }
}

/* no-slip walls only: one point at the face centre with the cell-mean
 * state, src/models/ns/charm_model_ns_diff_integrals.c:61-134 */
static void or_visc_bnd(or_sim_t *s, const or_face_t *f)
{
    or_data_t *d = &s->d[f->quad[0][0]];
    const int face = f->face[0];
    double n[3], l[3], qt[3], vv[3], vt[3], un, ll, mu, qu, qv, qw, qe, gw = 1., gj;
    const double *x;
    or_cons_t cons; or_prim_t prim[2];
    int i, ibf;
    if (s->cfg.bnd_type[f->bnd] != OR_BND_WALL_NO_SLIP) return;
    or_face_normal(d, face, n);
    for (i = 0; i < 3; i++) l[i] = d->g.fc[face][i] - d->g.c[i];
    x = d->g.fc[face];
    gj = d->g.area[face];
    mu = or_visc_mu(s, d, x);
    or_get_fields_avg(s, d, &cons);
    or_cons_to_prim(&s->cfg, &cons, &prim[0]);
    or_q_at(d, x, qt);
    or_bnd(&s->cfg, f->bnd, &prim[0], &prim[1], n);
    vv[0] = prim[0].u; vv[1] = prim[0].v; vv[2] = prim[0].w;
    un = or_scalar_prod(vv, n);
    for (i = 0; i < 3; i++) vt[i] = vv[i] - un*n[i];
    ll = or_vect_length(l);
    qu = -mu*vt[0]/ll; qv = -mu*vt[1]/ll; qw = -mu*vt[2]/ll;
    qe = -qt[0]*n[0] - qt[1]*n[1] - qt[2]*n[2];
    for (ibf = 0; ibf < OR_BF; ibf++) {
        double bfv = or_bf(&d->g, x, ibf) * gw * gj;
        d->integ[0][ibf] -= qu * bfv;
        d->integ[1][ibf] -= qv * bfv;
        d->integ[2][ibf] -= qw * bfv;
        d->integ[3][ibf] -= qe * bfv;
    }
}

/* src/models/ns/charm_model_ns_diff_integrals.c:232-292 (hanging) and
 * 330-382 (conforming) */
/* the surface flux at one point from each side's stress t[i] and heat flux
 * qs[i] at its trace */
static void or_visc_flux(or_sim_t *s, or_data_t *ud[2], const double xs[2][3], const double n[3],
                         double wj, double ts[2][6], double qs[2][3])
{
    double ftau[6] = {0, 0, 0, 0, 0, 0}, fu = 0., fv = 0., fw = 0., ft = 0., qu, qv, qw, qe;
    int i, k, ibf;
    for (i = 0; i < 2; i++) {
        or_cons_t c; or_prim_t p;
        const double *t = ts[i], *qt = qs[i];
        or_get_fields(s, ud[i], xs[i], &c);
        or_cons_to_prim(&s->cfg, &c, &p);
        for (k = 0; k < 6; k++) ftau[k] += t[k];
        fu += p.u*t[0] + p.v*t[3] + p.w*t[4];
        fv += p.u*t[3] + p.v*t[1] + p.w*t[5];
        fw += p.u*t[4] + p.v*t[5] + p.w*t[2];
        ft += qt[0]*n[0] + qt[1]*n[1] + qt[2]*n[2];
    }
    for (k = 0; k < 6; k++) ftau[k] *= 0.5;
    fu *= 0.5; fv *= 0.5; fw *= 0.5; ft *= 0.5;
    qu = ftau[0]*n[0] + ftau[3]*n[1] + ftau[4]*n[2];
    qv = ftau[3]*n[0] + ftau[1]*n[1] + ftau[5]*n[2];
    qw = ftau[4]*n[0] + ftau[5]*n[1] + ftau[2]*n[2];
    qe = fu*n[0] + fv*n[1] + fw*n[2] - ft;
    for (ibf = 0; ibf < OR_BF; ibf++)
        for (i = 0; i < 2; i++) {
            double bfv = (i ? -1. : 1.) * or_bf(&ud[i]->g, xs[i], ibf) * wj;
            ud[i]->integ[0][ibf] -= qu * bfv;
            ud[i]->integ[1][ibf] -= qv * bfv;
            ud[i]->integ[2][ibf] -= qw * bfv;
            ud[i]->integ[3][ibf] -= qe * bfv;
        }
}

/* BR1: each side's auxiliary tau_h, q_h at its trace */
static void or_visc_point(or_sim_t *s, or_data_t *ud[2], const double xs[2][3], const double n[3],
                          double wj)
{
    double ts[2][6], qs[2][3];
    int i;
    for (i = 0; i < 2; i++) {
        or_q_at(ud[i], xs[i], qs[i]);
        or_tau_at(ud[i], xs[i], ts[i]);
    }
    or_visc_flux(s, ud, xs, n, wj, ts, qs);
}

/* BR2 surface term of one interior face, (sub)face by (sub)face: each side's
 * lifting l_i (9 fields: tau xx yy zz xy xz yz, q x y z) from the points of
 * the (sub)face, then the flux with tau_own_i + eta l_i at every point.
 * Two switches the Makefile never sets plant plausible mistakes for
 * tests/test_oracle_mutants.py: OR_MUTANT_BR2_TAU_H takes BR1's tau_h for
 * tau_own, OR_MUTANT_BR2_LIFT_SIGN builds the lifting from Q(own) - Q(avg). */
#ifdef OR_MUTANT_BR2_TAU_H
#define OR_BR2_TAU(D) ((D)->tau)
#define OR_BR2_Q(D) ((D)->q)
#else
#define OR_BR2_TAU(D) ((D)->tau_own)
#define OR_BR2_Q(D) ((D)->q_own)
#endif
#ifdef OR_MUTANT_BR2_LIFT_SIGN
#define OR_BR2_JUMP(AVG, OWN) ((OWN) - (AVG))
#else
#define OR_BR2_JUMP(AVG, OWN) ((AVG) - (OWN))
#endif

static void or_visc_face_br2(or_sim_t *s, const or_face_t *f)
{
    const double eta = s->cfg.visc_br2_eta;
    int i, j, k, igp, ibf, nj = (f->is_hanging[0] || f->is_hanging[1]) ? 4 : 1, h_side = 0;
    for (i = 0; i < 2; i++) if (f->is_hanging[i]) h_side = i;
    for (j = 0; j < nj; j++) {
        or_data_t *ud[2];
        double n[3], xs[OR_FGP][2][3], wj[OR_FGP], lint[2][9][OR_BF], lift[2][9][OR_BF];
        for (i = 0; i < 2; i++) ud[i] = &s->d[f->is_hanging[i] ? f->quad[i][j] : f->quad[i][0]];
        or_face_normal(ud[0], f->face[0], n);
        memset(lint, 0, sizeof(lint));
        for (igp = 0; igp < OR_FGP; igp++) {
            const double *x = ud[h_side]->g.face_gp[f->face[h_side]][igp];
            or_gpt_t t;
            wj[igp] = ud[h_side]->g.face_gw[f->face[h_side]][igp] * ud[h_side]->g.face_gj[f->face[h_side]][igp];
            for (i = 0; i < 2; i++) or_shift(f, i, h_side, x, xs[igp][i]);
            or_grad_point_terms(s, ud, (const double (*)[3]) xs[igp], n, &t);
            for (i = 0; i < 2; i++) {
                double q6o[6], qqo[3];
                or_grad_own_terms(s, &t, i, n, q6o, qqo);
                for (ibf = 0; ibf < OR_BF; ibf++) {
                    const double bfv = (i ? -1. : 1.) * or_bf(&ud[i]->g, xs[igp][i], ibf) * wj[igp];
                    for (k = 0; k < 6; k++) lint[i][k][ibf] += OR_BR2_JUMP(t.q6[k], q6o[k]) * bfv;
                    for (k = 0; k < 3; k++) lint[i][6 + k][ibf] -= OR_BR2_JUMP(t.qq[k], qqo[k]) * bfv;
                }
            }
        }
        for (i = 0; i < 2; i++)
            for (k = 0; k < 9; k++) or_matr_vect_mult(ud[i]->g.a_inv, lint[i][k], lift[i][k]);
        for (igp = 0; igp < OR_FGP; igp++) {
            double ts[2][6], qs[2][3];
            for (i = 0; i < 2; i++) {
                const double *xi = xs[igp][i];
                for (k = 0; k < 6; k++)
                    ts[i][k] = or_coef_at(ud[i], OR_BR2_TAU(ud[i])[k], xi) + eta * or_coef_at(ud[i], lift[i][k], xi);
                for (k = 0; k < 3; k++)
                    qs[i][k] = or_coef_at(ud[i], OR_BR2_Q(ud[i])[k], xi) + eta * or_coef_at(ud[i], lift[i][6 + k], xi);
            }
            or_visc_flux(s, ud, (const double (*)[3]) xs[igp], n, wj[igp], ts, qs);
        }
    }
}

/* src/models/ns/charm_model_ns_diff_integrals.c:399-408 */
static void or_ns_integrals(or_sim_t *s)
{
    int64_t e, i;
    for (e = 0; e < s->n; e++) or_visc_volume(s, &s->d[e]);
    for (i = 0; i < s->nf; i++) {
        if (s->f[i].nsides != 2) or_visc_bnd(s, &s->f[i]);
        else if (s->cfg.visc_br2_eta > 0.) or_visc_face_br2(s, &s->f[i]);
        else or_inner_points(s, &s->f[i], or_visc_point);
    }
}

/* charm_model_ns_timestep_diff, src/models/ns/charm_model_ns_diff.c:83-90 */
void or_ns_diff(or_sim_t *s)
{
    if (!s->cfg.use_visc) return;
    or_ns_diffusion(s);
    or_ns_grad(s);
    or_ns_integrals(s);
}

/* ------------------------------------------------------------ chemistry */

typedef struct {
    double k[4], w[4], phi[OR_NC_MAX], psi[OR_NC_MAX];
} or_chem_t;

/* reaction rates at the cell-mean temperature, src/models/ns/charm_model_ns_chem.c:
 * 31-40 (and :82-91) */
static void or_chem_rates(const or_config_t *cfg, double t, const double *c_, or_chem_t *w)
{
    int n, m, i;
    for (n = 0; n < cfg->n_react; n++) {
        const or_reaction_t *r = &cfg->react[n];
        w->k[n] = r->a*pow(t, r->n)*exp(-r->e/t/or_gR);
        w->w[n] = w->k[n];
        for (m = 0; m < 3; m++) {
            i = r->left[m];
            w->w[n] *= (i < 0 ? 1. : c_[i]);
        }
    }
}

/* _charm_model_ns_chem_rhs, src/models/ns/charm_model_ns_chem.c:18-63 */
static void or_chem_rhs(const or_config_t *cfg, double t, const double *c_, or_chem_t *w)
{
    int i, j, m, n, ic;
    or_chem_rates(cfg, t, c_, w);
    for (i = 0; i < cfg->n_comp; i++) {
        w->phi[i] = 0.;
        w->psi[i] = 0.;
        for (n = 0; n < cfg->n_react; n++) {
            const or_reaction_t *r = &cfg->react[n];
            for (m = 0; m < 3; m++) {
                if (i == r->left[m]) {
                    double ww = w->k[n];
                    for (j = 1; j < 3; j++) {
                        ic = r->left[(m+j) % 3];
                        ww *= (ic < 0 ? 1. : c_[ic]);
                    }
                    w->phi[i] += ww;
                }
                if (i == r->right[m]) w->psi[i] += w->w[n];
            }
        }
    }
}

/* _charm_model_ns_chem_rhs_save, src/models/ns/charm_model_ns_chem.c:66-108.
 * The reference subtracts chem_w[i] (species index) for a consumed species
 * (:97); with chem_parity that index is kept and reads 0 where the
 * reference reads past the array (i >= reactions); otherwise chem_w[n]. */
static double or_chem_energy(const or_config_t *cfg, double t, const double *c_)
{
    or_chem_t w;
    double rhs = 0.;
    int i, m, n;
    or_chem_rates(cfg, t, c_, &w);
    for (i = 0; i < cfg->n_comp; i++) {
        double ws = 0.;
        for (n = 0; n < cfg->n_react; n++) {
            const or_reaction_t *r = &cfg->react[n];
            for (m = 0; m < 3; m++) {
                if (i == r->left[m]) {
                    if (cfg->chem_parity) ws -= (i < cfg->n_react) ? w.w[i] : 0.;
                    else ws -= w.w[n];
                }
                if (i == r->right[m]) ws += w.w[n];
            }
        }
        /* the reference multiplies the molar rate (mol/m^3/s) by the mass
         * enthalpy (J/kg); chem_energy_molar takes the molar enthalpy
         * M_i h_i instead, which gives the source in J/m^3/s */
        rhs += ws*(cfg->chem_energy_molar ? cfg->comp[i].m : 1.)*or_comp_calc_enthalpy(&cfg->comp[i], t, cfg->t_ref);
    }
    return rhs;
}

/* _charm_model_ns_chem_iter_fn, src/models/ns/charm_model_ns_chem.c:111-156:
 * two semi-implicit passes on the cell-mean molar concentrations, energy
 * source, then rc_i <- [C_i M_i, 0, 0, 0] */
static void or_chem_cell(or_sim_t *s, or_data_t *d, double dt)
{
    const or_config_t *cfg = &s->cfg;
    const int nc = cfg->n_comp;
    or_cons_t cons; or_prim_t prim;
    double c[OR_NC_MAX], c_hat[OR_NC_MAX], c_[OR_NC_MAX];
    or_chem_t w;
    int i, j;
    or_get_fields_avg(s, d, &cons);
    or_cons_to_prim(cfg, &cons, &prim);
    for (i = 0; i < nc; i++) {
        c[i] = cons.rc[i] / cfg->comp[i].m;
        c_hat[i] = c[i];
    }
    for (j = 0; j < 2; j++) {
        for (i = 0; i < nc; i++) c_[i] = (c_hat[i] + c[i]) * 0.5;
        or_chem_rhs(cfg, prim.t, c_, &w);
        for (i = 0; i < nc; i++) {
            c_hat[i]  = (c[i] + dt * w.psi[i] * (1. + dt * w.phi[i] * 0.5));
            c_hat[i] /= (1. + dt * w.phi[i] + (dt * w.phi[i]) * (dt * w.phi[i]) * 0.5);
        }
    }
    for (i = 0; i < nc; i++) c_[i] = c_hat[i];
    d->chem_rhs = or_chem_energy(cfg, prim.t, c_);
    for (i = 0; i < nc; i++) {
        memset(d->c[4+i], 0, sizeof(d->c[4+i]));
        d->c[4+i][0] = c_hat[i] * cfg->comp[i].m;
    }
}

/* charm_model_ns_timestep_chem, src/models/ns/charm_model_ns_chem.c:165-202 */
void or_ns_chem(or_sim_t *s)
{
    int64_t e;
    if (s->cfg.n_react == 0) {
        for (e = 0; e < s->n; e++) s->d[e].chem_rhs = 0.;
        return;
    }
    for (e = 0; e < s->n; e++) or_chem_cell(s, &s->d[e], s->dt);
}

/* ------------------------------------------------------------ test hooks */

void or_ns_aux(or_sim_t *s, double *out)
{
    int64_t e;
    int k;
    or_ns_diffusion(s);
    or_ns_grad(s);
    for (e = 0; e < s->n; e++) {
        for (k = 0; k < 6; k++) memcpy(&out[(e*9 + k)*4], s->d[e].tau[k], 4*sizeof(double));
        for (k = 0; k < 3; k++) memcpy(&out[(e*9 + 6 + k)*4], s->d[e].q[k], 4*sizeof(double));
    }
}

/* BR2's tau_own, q_own in the layout of or_ns_aux (visc_br2_eta > 0) */
void or_ns_aux_own(or_sim_t *s, double *out)
{
    int64_t e;
    int k;
    or_ns_diffusion(s);
    or_ns_grad(s);
    for (e = 0; e < s->n; e++) {
        for (k = 0; k < 6; k++) memcpy(&out[(e*9 + k)*4], s->d[e].tau_own[k], 4*sizeof(double));
        for (k = 0; k < 3; k++) memcpy(&out[(e*9 + 6 + k)*4], s->d[e].q_own[k], 4*sizeof(double));
    }
}

void or_ns_diff_coeffs(or_sim_t *s, double *out)
{
    int64_t e;
    or_ns_diffusion(s);
    for (e = 0; e < s->n; e++) memcpy(&out[e*s->cfg.n_comp], s->d[e].d, (size_t) s->cfg.n_comp*sizeof(double));
}

void or_ns_chem_apply(or_sim_t *s, double dt, double *chem_rhs)
{
    int64_t e;
    s->dt = dt;
    or_ns_chem(s);
    if (chem_rhs) for (e = 0; e < s->n; e++) chem_rhs[e] = s->d[e].chem_rhs;
}
