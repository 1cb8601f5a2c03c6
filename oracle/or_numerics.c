/*
 * or_numerics.c -- pointwise numerics of the oracle (FP64).
 *
 * TEST INFRASTRUCTURE ONLY (see charm_oracle.h).  Each function restates the
 * reference function cited beside it; the arithmetic order of the reference
 * is kept wherever it fixes FP64 rounding.
 */
#include "or_internal.h"

/* gR, src/charm_eos.c:8 */
const double or_gR = 8.31446261815324;

double or_gas_r(void) { return or_gR; }
size_t or_prim_size(void) { return sizeof(or_prim_t); }
size_t or_cons_size(void) { return sizeof(or_cons_t); }
size_t or_config_size(void) { return sizeof(or_config_t); }

/* ---------------------------------------------------------------- vectors
 * src/charm_globals.c:28-44 */
double or_scalar_prod(const double a[3], const double b[3])
{
    return a[0]*b[0]+a[1]*b[1]+a[2]*b[2];
}

double or_vect_length(const double v[3])
{
    return sqrt(v[0]*v[0]+v[1]*v[1]+v[2]*v[2]);
}

void or_vect_prod(const double a[3], const double b[3], double r[3])
{
    r[0] =  a[1]*b[2]-a[2]*b[1];
    r[1] = -a[0]*b[2]+a[2]*b[0];
    r[2] =  a[0]*b[1]-a[1]*b[0];
}

/* 3x3 determinant in the reference's term order, src/charm_globals.c:244-256 */
double or_matr3_det(double a[3][3])
{
    double d = 0.;
    d += a[0][0]*a[1][1]*a[2][2];
    d += a[0][2]*a[1][0]*a[2][1];
    d += a[2][0]*a[0][1]*a[1][2];
    d -= a[0][2]*a[1][1]*a[2][0];
    d -= a[0][0]*a[1][2]*a[2][1];
    d -= a[0][1]*a[1][0]*a[2][2];
    return d;
}

/* Gauss-Jordan with partial pivoting (ties go to the later row, as the
 * reference's `<=` test does), src/charm_globals.c:285-394.  The final row
 * permutation of the reference acts on its scratch copy `a`, not on the
 * inverse, so it does not affect the result and is omitted. */
void or_matr_inv(const double a_src[4][4], double am[4][4])
{
    double a[4][4];
    int i, j, ni, nj, maxind;
    double fmaxval, tmp;
    for (i = 0; i < 4; i++)
        for (j = 0; j < 4; j++) {
            a[i][j] = a_src[i][j];
            am[i][j] = (i == j) ? 1.0 : 0.0;
        }
    for (i = 0; i < 4; i++) {
        maxind = i;
        fmaxval = fabs(a[i][i]);
        for (ni = i + 1; ni < 4; ni++) {
            if (fabs(fmaxval) <= fabs(a[ni][i])) {
                fmaxval = fabs(a[ni][i]);
                maxind = ni;
            }
        }
        if (i != maxind) {
            for (nj = 0; nj < 4; nj++) {
                tmp = a[i][nj]; a[i][nj] = a[maxind][nj]; a[maxind][nj] = tmp;
                tmp = am[i][nj]; am[i][nj] = am[maxind][nj]; am[maxind][nj] = tmp;
            }
        }
        double aii = a[i][i];
        for (j = 0; j < 4; j++) {
            a[i][j] = a[i][j] / aii;
            am[i][j] = am[i][j] / aii;
        }
        for (ni = 0; ni < 4; ni++) {
            if (ni != i) {
                double f = a[ni][i];
                for (nj = 0; nj < 4; nj++) {
                    a[ni][nj] = a[ni][nj] - f * a[i][nj];
                    am[ni][nj] = am[ni][nj] - f * am[i][nj];
                }
            }
        }
    }
}

/* src/charm_globals.c:397-406 */
void or_matr_vect_mult(double a[4][4], const double b[4], double r[4])
{
    int i, j;
    for (i = 0; i < 4; i++) {
        r[i] = 0.;
        for (j = 0; j < 4; j++) r[i] += a[i][j]*b[j];
    }
}

/* ----------------------------------------------------- component properties
 * src/charm_globals.c:603-694.  The polynomial branch of calc_cp is reached
 * only when ml_type == 1 (the reference compares ml_type with
 * COMP_CP_POLYNOM, charm_globals.c:613); this is restated as is. */
double or_comp_calc_cp(const or_comp_t *c, double t)
{
    int i;
    double res = 0., tt = 1.;
    if (c->cp_type == 0) return c->cp[0];
    else if (c->ml_type == 1) {
        for (i = 0; i < c->n_cp; i++) { res += tt*c->cp[i]; tt *= t; }
        return res;
    }
    return c->cp[0];
}

double or_comp_calc_cp_dt(const or_comp_t *c, double t)
{
    int i;
    double res = 0., tt = 1.;
    if (c->cp_type == OR_CP_POLYNOM) {
        for (i = 1; i < c->n_cp; i++) { res += tt*c->cp[i]*i; tt *= t; }
        return res;
    }
    return 0.;   /* the reference aborts here (charm_globals.c:644-646) */
}

double or_comp_calc_ml(const or_comp_t *c, double t)
{
    if (c->ml_type == OR_ML_CONST) return c->ml0;
    if (c->ml_type == OR_ML_SATHERLAND)
        return c->ml0 * sqrt(pow(t / c->t0, 3)) * (c->t0 + c->ts) / (t + c->ts);
    return c->ml0;
}

double or_comp_calc_kp(const or_comp_t *c, double t)
{
    if (c->kp_type == 0) return c->kp0;
    if (c->kp_type == 1)
        return c->kp0 * sqrt(pow(t / c->t0, 3)) * (c->t0 + c->ts) / (t + c->ts);
    return c->kp0;
}

double or_comp_calc_enthalpy(const or_comp_t *c, double t, double t_ref)
{
    int i;
    double h = c->h0, t1, t2;
    for (i = 0; i < c->n_cp; i++) {
        t2 = pow(t, i+1)/(i+1);
        t1 = pow(t_ref, (i+1))/(i+1);
        h += c->cp[i]*(t2-t1);
    }
    return h;
}

/* -------------------------------------------------------------------- EOS */

/* Newton solve of (cp(T) - R*sum c_i/M_i) T = e from T = 1 K,
 * src/charm_eos.c:12-54 */
static void or_calc_t(const or_config_t *cfg, or_prim_t *p)
{
    double tt = 1.0, e = p->e, m_ = 0.0, rm, cp, cp_dt, ft, ft_dt, tt1;
    int i, j;
    for (i = 0; i < cfg->n_comp; i++) m_ += p->c[i] / cfg->comp[i].m;
    rm = or_gR * m_;
    for (j = 0; j < 100; j++) {
        cp = 0.0; cp_dt = 0.0;
        for (i = 0; i < cfg->n_comp; i++) {
            cp    += p->c[i] * or_comp_calc_cp(&cfg->comp[i], tt);
            cp_dt += p->c[i] * or_comp_calc_cp_dt(&cfg->comp[i], tt);
        }
        ft    = (cp - rm) * tt - e;
        ft_dt = (cp - rm) + cp_dt * tt;
        tt1 = tt - ft / ft_dt;
        if (pow(tt1 - tt, 2) < OR_EPS) { tt = tt1; break; }
        tt = tt1;
    }
    p->t = tt;
}

/* src/charm_eos.c:57-96 */
static void or_eos_switch(or_prim_t *p, int flag)
{
    switch (flag) {
    case 0:
        if (p->r < OR_EPS) p->r = OR_EPS;
        p->p = p->r*p->e*(p->gam-1);
        p->cz = sqrt(p->gam*p->p/p->r);
        break;
    case 1:
        if (p->r < OR_EPS) p->r = OR_EPS;
        if (p->p < OR_EPS) p->p = OR_EPS;
        p->e = p->p/(p->r*(p->gam-1));
        p->t = p->e/p->cv;
        break;
    case 2:
        if (p->p < OR_EPS) p->p = OR_EPS;
        p->r = p->p*p->m/(p->t*or_gR);
        p->cz = sqrt(p->gam*p->p/p->r);
        break;
    case 3:
        if (p->p < OR_EPS) p->p = OR_EPS;
        p->r  = p->p*p->m/(p->t*or_gR);
        p->cz = sqrt(p->gam*p->p/p->r);
        p->e  = p->p/(p->r*(p->gam-1));
        break;
    case 4:
        if (p->r < OR_EPS) p->r = OR_EPS;
        p->p  = p->r*p->e*(p->gam-1);
        p->cz = sqrt(p->gam*p->p/p->r);
        p->t  = p->e/p->cv;
        break;
    default:
        break;
    }
}

/* ideal gas from component 0, src/charm_eos.c:98-118;
 * mixture, src/charm_eos.c:120-148 */
void or_eos(const or_config_t *cfg, or_prim_t *p, int flag)
{
    const or_comp_t *comp = &cfg->comp[0];
    double t;
    int i;
    if (comp->cp_type == OR_CP_POLYNOM && (flag == 1 || flag == 4)) {
        or_calc_t(cfg, p);
        t = p->t;
    } else {
        t = p->t;
    }
    if (cfg->eos == OR_EOS_IDEAL) {
        p->cp  = or_comp_calc_cp(comp, t);
        p->m   = comp->m;
        p->cv  = p->cp - or_gR/p->m;
        p->gam = p->cp/p->cv;
    } else {
        double M_ = 0.;
        p->cp = 0.;
        for (i = 0; i < cfg->n_comp; i++) {
            M_ += p->c[i]/cfg->comp[i].m;
            p->cp += p->c[i]*or_comp_calc_cp(&cfg->comp[i], t);
        }
        p->m   = 1./M_;
        p->cv  = p->cp - or_gR/p->m;
        p->gam = p->cp/p->cv;
    }
    or_eos_switch(p, flag);
}

/* src/charm_globals.c:183-208 */
void or_cons_to_prim(const or_config_t *cfg, const or_cons_t *c, or_prim_t *p)
{
    int i;
    p->mat_id = c->mat_id;
    p->r = 0.;
    for (i = 0; i < cfg->n_comp; i++) p->r += c->rc[i];
    p->u = c->ru/p->r;
    p->v = c->rv/p->r;
    p->w = c->rw/p->r;
    p->e_tot = c->re/p->r;
    p->e = p->e_tot-0.5*(p->u*p->u+p->v*p->v+p->w*p->w);
    for (i = 0; i < cfg->n_comp; i++) {
        p->c[i] = c->rc[i] / p->r;
        if (p->c[i] < 0.) p->c[i] = 0.;
        if (p->c[i] > 1.) p->c[i] = 1.;
    }
    or_eos(cfg, p, 4);
}

/* src/charm_globals.c:211-223 */
void or_prim_to_cons(const or_config_t *cfg, or_cons_t *c, const or_prim_t *p)
{
    int i;
    c->mat_id = p->mat_id;
    c->ru = p->r * p->u;
    c->rv = p->r * p->v;
    c->rw = p->r * p->w;
    c->re = p->r * (p->e + 0.5 * (p->u*p->u+p->v*p->v+p->w*p->w));
    for (i = 0; i < cfg->n_comp; i++) c->rc[i] = p->r * p->c[i];
}

/* ------------------------------------------------------------ boundaries
 * dispatcher src/charm_bnd_cond.c:21-33, functions :36-93 */
void or_bnd(const or_config_t *cfg, int face, const or_prim_t *in,
            or_prim_t *out, const double n[3])
{
    const double *param = cfg->bnd_param[face];
    int i;
    *out = *in;
    switch (cfg->bnd_type[face]) {
    case OR_BND_INLET:
        out->u = param[0]; out->v = param[1]; out->w = param[2];
        out->t = param[3]; out->p = param[4];
        break;
    case OR_BND_OUTLET:
        break;
    case OR_BND_WALL_SLIP: {
        double v[3] = {in->u, in->v, in->w};
        double svn = or_scalar_prod(v, n);
        double vv[3] = {n[0]*svn, n[1]*svn, n[2]*svn};
        for (i = 0; i < 3; i++) { v[i] -= vv[i]; v[i] -= vv[i]; }
        out->u = v[0]; out->v = v[1]; out->w = v[2];
        break;
    }
    case OR_BND_WALL_NO_SLIP: {
        double v[3] = {in->u, in->v, in->w};
        double svn = or_scalar_prod(v, n);
        double vv[3] = {n[0]*svn, n[1]*svn, n[2]*svn};
        for (i = 0; i < 3; i++) v[i] -= vv[i];
        out->u = v[0]; out->v = v[1]; out->w = v[2];
        out->t = param[0];
        break;
    }
    case OR_BND_MASS_FLOW: {
        /* src/charm_bnd_cond.c:101-146: param = (mass flux, static p,
         * total T, direction cosines, outflow p, composition).  The outflow
         * pressure param[6] is read and never written back: the reference
         * stores the last supersonic face's pressure there, which depends on
         * the face-visit order */
        double cs[3] = {param[3], param[4], param[5]};
        double fd = or_scalar_prod(n, cs);
        if (fd > 0.) {
            double vel = out->u*n[0] + out->v*n[1] + out->w*n[2];
            if (vel < out->cz) out->p = param[6];
        } else {
            double v2, k, vm;
            for (i = 0; i < cfg->n_comp; i++) out->c[i] = param[7+i];
            or_eos(cfg, out, 0);
            v2 = out->u*out->u + out->v*out->v + out->w*out->w;
            if (v2 > out->cz*out->cz) out->p = param[1];
            k = pow(param[0]*(out->cp - out->cv)/(out->p*fd), 2.) / (2.*out->cp);
            out->t = (-1. + sqrt(1. + 4.*k*param[2])) / (2.*k);
            vm = fabs(param[0] / (out->r*fd));
            out->u = cs[0]*vm; out->v = cs[1]*vm; out->w = cs[2]*vm;
        }
        break;
    }
    default:
        break;
    }
    out->mat_id = in->mat_id;
    or_eos(cfg, out, 3);
    out->e_tot = out->e+0.5*(out->u*out->u+out->v*out->v+out->w*out->w);
}

/* ----------------------------------------------------------------- fluxes */

/* Lax-Friedrichs, src/charm_fluxes.c:255-315 */
static void or_flux_lf(const or_config_t *cfg, const or_prim_t prim[2],
                       const double n[3], double *q)
{
    int i, j, fc = 4 + cfg->n_comp;
    double vn[2], ff[OR_NV_MAX][2], uu[OR_NV_MAX][2], alpha;
    for (i = 0; i < 2; i++) {
        vn[i] = prim[i].u*n[0]+prim[i].v*n[1]+prim[i].w*n[2];
        uu[0][i] = prim[i].r*prim[i].u;
        uu[1][i] = prim[i].r*prim[i].v;
        uu[2][i] = prim[i].r*prim[i].w;
        uu[3][i] = prim[i].r*prim[i].e_tot;
        ff[0][i] = prim[i].r*vn[i]*prim[i].u+prim[i].p*n[0];
        ff[1][i] = prim[i].r*vn[i]*prim[i].v+prim[i].p*n[1];
        ff[2][i] = prim[i].r*vn[i]*prim[i].w+prim[i].p*n[2];
        ff[3][i] = (prim[i].r*prim[i].e_tot+prim[i].p)*vn[i];
        for (j = 4; j < fc; j++) {
            uu[j][i] = prim[i].r*prim[i].c[j-4];
            ff[j][i] = prim[i].r*vn[i]*prim[i].c[j-4];
        }
    }
    alpha = OR_MAX(fabs(vn[0])+prim[0].cz, fabs(vn[1])+prim[1].cz);
    for (i = 0; i < fc; i++)
        q[i] = 0.5*(ff[i][1]+ff[i][0]-alpha*(uu[i][1]-uu[i][0]));
}

/* central, src/charm_fluxes.c:513-559 */
static void or_flux_cd(const or_config_t *cfg, const or_prim_t prim[2],
                       const double n[3], double *q)
{
    int i, j, fc = 4 + cfg->n_comp;
    double vn[2], ff[OR_NV_MAX][2];
    for (i = 0; i < 2; i++) {
        vn[i] = prim[i].u*n[0]+prim[i].v*n[1]+prim[i].w*n[2];
        ff[0][i] = prim[i].r*vn[i]*prim[i].u+prim[i].p*n[0];
        ff[1][i] = prim[i].r*vn[i]*prim[i].v+prim[i].p*n[1];
        ff[2][i] = prim[i].r*vn[i]*prim[i].w+prim[i].p*n[2];
        ff[3][i] = (prim[i].r*prim[i].e_tot+prim[i].p)*vn[i];
        for (j = 4; j < fc; j++) ff[j][i] = prim[i].r*prim[i].c[j-4]*vn[i];
    }
    for (i = 0; i < fc; i++) q[i] = 0.5*(ff[i][1]+ff[i][0]);
}

/* OR_MUTANT_HLLC_SIGN (never set by the Makefile) flips the sign of the
 * (S* - u_k) factor: a planted mistake that tests/test_oracle_mutants.py
 * builds on purpose to show that the star-region checks catch it */
#ifdef OR_MUTANT_HLLC_SIGN
#define OR_SS_MINUS_VK(SS, VK) ((VK)-(SS))
#else
#define OR_SS_MINUS_VK(SS, VK) ((SS)-(VK))
#endif
#define F_HLLC_U(UK, FK, SK, SS, PK, RK, VK) (((SS)*((SK)*(UK)-(FK)) + (SK)*( (PK)+(RK)*((SK)-(VK))*OR_SS_MINUS_VK(SS, VK) )) / ((SK)-(SS)))
#define F_HLLC_V(UK, FK, SK, SS, PK, RK, VK) (((SS)*((SK)*(UK)-(FK))) / ((SK)-(SS)))
#define F_HLLC_E(UK, FK, SK, SS, PK, RK, VK) (((SS)*((SK)*(UK)-(FK)) + (SK)*( (PK)+(RK)*((SK)-(VK))*OR_SS_MINUS_VK(SS, VK) )*(SS)) / ((SK)-(SS)))

/* x-split HLLC core, src/charm_fluxes.c:319-437.  Returns nonzero where the
 * reference aborts (wave ordering violated, i.e. NaN input, :354-357). */
static int or_hllc_x(const or_config_t *cfg, const or_prim_t p[2],
                     double *qu, double *qv, double *qw, double *qe, double *qc)
{
    int i, nc = cfg->n_comp;
    double sl, sr, p_star, s_star, p_pvrs, ql, qr, tmp;
    p_pvrs = 0.5*(p[0].p+p[1].p)-0.5*(p[1].u-p[0].u)*0.25*(p[0].r+p[1].r)*(p[0].cz+p[1].cz);
    p_star = (p_pvrs > 0.) ? p_pvrs : 0.;
    ql = (p_star <= p[0].p) ? 1 : sqrt(1.+(p[0].gam+1.)*(p_star/p[0].p-1.)/(2.*p[0].gam));
    qr = (p_star <= p[1].p) ? 1 : sqrt(1.+(p[1].gam+1.)*(p_star/p[1].p-1.)/(2.*p[1].gam));
    sl = p[0].u-p[0].cz*ql;
    sr = p[1].u+p[1].cz*qr;
    if (sl > sr) { tmp = sl; sl = sr; sr = tmp; }
    s_star  = p[1].p-p[0].p;
    s_star += p[0].r*p[0].u*(sl-p[0].u);
    s_star -= p[1].r*p[1].u*(sr-p[1].u);
    s_star /= (p[0].r*(sl-p[0].u)-p[1].r*(sr-p[1].u));
    if (s_star < sl) s_star = sl;
    if (s_star > sr) s_star = sr;
    if (!((sl <= s_star) && (s_star <= sr))) return 1;
    if (sl >= 0.) {
        *qu = p[0].r*p[0].u*p[0].u+p[0].p;
        *qv = p[0].r*p[0].v*p[0].u;
        *qw = p[0].r*p[0].w*p[0].u;
        *qe = (p[0].r*p[0].e_tot+p[0].p)*p[0].u;
        for (i = 0; i < nc; i++) qc[i] = p[0].r*p[0].u*p[0].c[i];
    } else if (sr <= 0.) {
        *qu = p[1].r*p[1].u*p[1].u+p[1].p;
        *qv = p[1].r*p[1].v*p[1].u;
        *qw = p[1].r*p[1].w*p[1].u;
        *qe = (p[1].r*p[1].e_tot+p[1].p)*p[1].u;
        for (i = 0; i < nc; i++) qc[i] = p[1].r*p[1].u*p[1].c[i];
    } else {
        int k = (s_star >= 0) ? 0 : 1;
        double sk = k ? sr : sl;
        const or_prim_t *P = &p[k];
        *qu = F_HLLC_U(P->r*P->u, P->r*P->u*P->u+P->p, sk, s_star, P->p, P->r, P->u);
        *qv = F_HLLC_V(P->r*P->v, P->r*P->u*P->v, sk, s_star, P->p, P->r, P->u);
        *qw = F_HLLC_V(P->r*P->w, P->r*P->u*P->w, sk, s_star, P->p, P->r, P->u);
        *qe = F_HLLC_E(P->r*P->e_tot, (P->r*P->e_tot+P->p)*P->u, sk, s_star, P->p, P->r, P->u);
        for (i = 0; i < nc; i++)
            qc[i] = F_HLLC_V(P->r*P->c[i], P->r*P->c[i]*P->u, sk, s_star, P->p, P->r, P->u);
    }
    return 0;
}

/* normal/tangent frame, src/charm_fluxes.c:458-488 */
static void or_frame(const double n[3], double nt[3][3])
{
    double ri;
    nt[0][0] = n[0]; nt[0][1] = n[1]; nt[0][2] = n[2];
    nt[1][0] = nt[1][1] = nt[1][2] = 0.;
    ri = sqrt(n[0]*n[0]+n[1]*n[1]);
    if (ri > OR_EPS) {
        nt[1][0] = -n[1]/ri; nt[1][1] = n[0]/ri; nt[1][2] = 0.;
    } else {
        ri = sqrt(n[1]*n[1]+n[2]*n[2]);
        if (ri > OR_EPS) {
            nt[1][0] = 0.; nt[1][1] = -n[2]/ri; nt[1][2] = n[1]/ri;
        } else {
            ri = sqrt(n[0]*n[0]+n[2]*n[2]);
            if (ri > OR_EPS) {
                nt[1][0] = -n[2]/ri; nt[1][1] = 0.; nt[1][2] = n[0]/ri;
            }
        }
    }
    or_vect_prod(nt[0], nt[1], nt[2]);
}

/* rotated HLLC, src/charm_fluxes.c:439-510 */
static int or_flux_hllc(const or_config_t *cfg, const or_prim_t prim[2],
                        const double n[3], double *q)
{
    int i, j, err;
    double nt[3][3], vv[3], vn[2][3], _qu = 0., _qv = 0., _qw = 0.;
    or_prim_t pr[2];
    or_frame(n, nt);
    for (i = 0; i < 2; i++) {
        vv[0] = prim[i].u; vv[1] = prim[i].v; vv[2] = prim[i].w;
        for (j = 0; j < 3; j++) vn[i][j] = or_scalar_prod(vv, nt[j]);
        pr[i] = prim[i];
        pr[i].u = vn[i][0]; pr[i].v = vn[i][1]; pr[i].w = vn[i][2];
    }
    err = or_hllc_x(cfg, pr, &_qu, &_qv, &_qw, &q[3], &q[4]);
    q[0] = _qu*nt[0][0] + _qv*nt[1][0] + _qw*nt[2][0];
    q[1] = _qu*nt[0][1] + _qv*nt[1][1] + _qw*nt[2][1];
    q[2] = _qu*nt[0][2] + _qv*nt[1][2] + _qw*nt[2][2];
    return err;
}

/* Exact Riemann solver: the Newton iteration on p* and the sampling at
 * xi = 0 of rim_orig, src/charm_fluxes.c:8-180, with the wave functions of
 * each side taking that side's gamma (gam_b == gam_e reduces to rim_orig,
 * which takes a single gamma).  *side receives 0 if the sampled state lies
 * left of the contact, 1 otherwise. */
static void or_rim_side(double *RI, double *EI, double *PI, double *UI, double *VI, double *WI,
                        int *side,
                        double RB, double PB, double UB, double VB, double WB,
                        double RE, double PE, double UE, double VE, double WE,
                        double gb, double ge)
{
    int step;
    double AB = gb - 1.0, AE = ge - 1.0;
    double DB = 2.0 / AB, DE = 2.0 / AE;
    double GB = sqrt(gb * AB), GE = sqrt(ge * AE);
    double HB = AB / 2.0, HE = AE / 2.0;
    double FB = 3.0 * gb - 1.0, FE = 3.0 * ge - 1.0;
    double OB = AB / (2.0 * gb), OE = AE / (2.0 * ge);
    double QB = gb + 1.0, QE = ge + 1.0;
    double PGB = QB / (2.0 * gb), PGE = QE / (2.0 * ge);
    double RGB = 4.0 * gb, RGE = 4.0 * ge;
    double SB = gb * AB, SE = ge * AE;
    double TB = QB / 2.0, TE = QE / 2.0;
    double RF = 0, RS = 0, EF = 0, ES = 0, SBL, SFL, SSL, SEL, D, FS1, F1, ZNB, PKB, ZFB,
           PKE, ZNE, F2, FS2, ZFE, DP, UBD, RUBD, UF = 0, UED, RUED, US = 0, PPE, PPB, P;
    double eps = OR_RIM_EPS;
    double CB = sqrt(gb * PB / RB);
    double CE = sqrt(ge * PE / RE);
    double EB = CB * CB / SB;
    double EE = CE * CE / SE;
    double RCB = RB * CB;
    double RCE = RE * CE;
    double DU = UB - UE;
    if (DU < -(2.0 * CB / AB + 2.0 * CE / AE)) {
        RF = 0.0; RS = 0.0; EF = 0.0; ES = 0.0;
        SBL = UB - CB;
        SFL = UB + 2.0 * CB / AB;
        SSL = UE - 2.0 * CE / AE;
        SEL = UE + CE;
    } else {
        P = (PB * RCE + PE * RCB + DU * RCB * RCE) / (RCB + RCE);
        step = 0;
        do {
            if (P < eps) P = eps;
            PPB = P / PB;
            if (PB <= P) {
                PKB = PGB * PPB + OB;
                ZNB = RCB * sqrt(PKB);
                F1 = (P - PB) / ZNB;
                FS1 = (QB * PPB + FB) / (RGB * ZNB * PKB);
            } else {
                ZFB = CB * exp(log(PPB) * OB);
                F1 = DB * (ZFB - CB);
                FS1 = ZFB / (gb * P);
            }
            PPE = P / PE;
            if (PE <= P) {
                PKE = PGE * PPE + OE;
                ZNE = RCE * sqrt(PKE);
                F2 = (P - PE) / ZNE;
                FS2 = (QE * PPE + FE) / (RGE * ZNE * PKE);
            } else {
                ZFE = CE * exp(log(PPE) * OE);
                F2 = DE * (ZFE - CE);
                FS2 = ZFE / (ge * P);
            }
            DP = (DU - F1 - F2) / (FS1 + FS2);
            P = P + DP;
        } while ((fabs(DU - F1 - F2) > eps) && (++step < OR_RIM_NEWTON_STEPS));

        PPB = P / PB;
        PPE = P / PE;
        ZFB = CB * exp(log(PPB) * OB);
        ZFE = CE * exp(log(PPE) * OE);
        if (PB <= P) {
            D = UB - sqrt((TB * P + HB * PB) / RB);
            UBD = UB - D;
            RUBD = RB * UBD;
            RF = RUBD * RUBD / (PB - P + RUBD * UBD);
            UF = D + RUBD / RF;
            EF = P / (AB * RF);
            SBL = D;
            SFL = D;
        } else {
            EF = ZFB * ZFB / SB;
            UF = UB + DB * (CB - ZFB);
            RF = P / (AB * EF);
            SBL = UB - CB;
            SFL = UF - ZFB;
        }
        if (PE <= P) {
            D = UE + sqrt((TE * P + HE * PE) / RE);
            UED = UE - D;
            RUED = RE * UED;
            RS = RUED * RUED / (PE - P + RUED * UED);
            US = D + RUED / RS;
            ES = P / (AE * RS);
            SEL = D;
            SSL = D;
        } else {
            ES = ZFE * ZFE / SE;
            US = UE - DE * (CE - ZFE);
            RS = P / (AE * ES);
            SSL = US + ZFE;
            SEL = UE + CE;
        }
    }
    if (SEL <= 0.0) {
        *RI = RE; *EI = EE; *UI = UE; *VI = VE; *WI = WE;
        *PI = AE*(*EI)*(*RI); *side = 1;
        return;
    }
    if (SBL >= 0.0) {
        *RI = RB; *EI = EB; *UI = UB; *VI = VB; *WI = WB;
        *PI = AB*(*EI)*(*RI); *side = 0;
        return;
    }
    if ((SSL >= 0.0) && (SFL <= 0.0)) {
        if (US >= 0.0) {
            *RI = RF; *EI = EF; *UI = UF; *VI = VB; *WI = WB;
            *PI = AB*(*EI)*(*RI); *side = 0;
        } else {
            *RI = RS; *EI = ES; *UI = US; *VI = VE; *WI = WE;
            *PI = AE*(*EI)*(*RI); *side = 1;
        }
        return;
    }
    if (SFL > 0.0) {
        *UI = (UB+DB*GB*sqrt(EB))/(1+DB);
        *VI = VB; *WI = WB;
        *EI = ((*UI)*(*UI))/SB;
        *RI = RB*exp(log((*EI)/EB)*(1/AB));
        *PI = AB*(*EI)*(*RI); *side = 0;
    } else {
        *UI = (UE-DE*GE*sqrt(EE))/(1+DE);
        *VI = VE; *WI = WE;
        *EI = ((*UI)*(*UI))/SE;
        *RI = RE*exp(log((*EI)/EE)*(1/AE));
        *PI = AE*(*EI)*(*RI); *side = 1;
    }
}

void or_rim(double *RI, double *EI, double *PI, double *UI, double *VI, double *WI,
            double RB, double PB, double UB, double VB, double WB,
            double RE, double PE, double UE, double VE, double WE,
            double gam_b, double gam_e)
{
    int side;
    or_rim_side(RI, EI, PI, UI, VI, WI, &side, RB, PB, UB, VB, WB,
                RE, PE, UE, VE, WE, gam_b, gam_e);
}

/* Godunov flux with the exact solver in the rotated frame, following the
 * disabled body of charm_calc_flux_godunov (src/charm_fluxes.c:184-252);
 * species ride on the sampled mass flux from the side of the contact the
 * sample lies on. */
static void or_flux_godunov(const or_config_t *cfg, const or_prim_t prim[2],
                            const double n[3], double *q)
{
    int i, j, side;
    double nt[3][3], vv[3], vn[2][3], ri, ei, pi, uu[3], uv[3], qr;
    or_frame(n, nt);
    for (i = 0; i < 2; i++) {
        vv[0] = prim[i].u; vv[1] = prim[i].v; vv[2] = prim[i].w;
        for (j = 0; j < 3; j++) vn[i][j] = or_scalar_prod(vv, nt[j]);
    }
    or_rim_side(&ri, &ei, &pi, &uu[0], &uu[1], &uu[2], &side,
                prim[0].r, prim[0].p, vn[0][0], vn[0][1], vn[0][2],
                prim[1].r, prim[1].p, vn[1][0], vn[1][1], vn[1][2],
                prim[0].gam, prim[1].gam);
    for (i = 0; i < 3; i++) {
        uv[i] = 0.;
        for (j = 0; j < 3; j++) uv[i] += nt[j][i]*uu[j];
    }
    qr = ri*uu[0];
    q[0] = qr*uv[0]+pi*n[0];
    q[1] = qr*uv[1]+pi*n[1];
    q[2] = qr*uv[2]+pi*n[2];
    q[3] = (ri*(ei+0.5*(uv[0]*uv[0]+uv[1]*uv[1]+uv[2]*uv[2]))+pi)*uu[0];
    for (i = 0; i < cfg->n_comp; i++) q[4+i] = qr*prim[side].c[i];
}

int or_flux_err(const or_config_t *cfg, int flux, const or_prim_t prim[2],
                const double n[3], double *q)
{
    switch (flux) {
    case OR_FLUX_LF:      or_flux_lf(cfg, prim, n, q); return 0;
    case OR_FLUX_HLLC:    return or_flux_hllc(cfg, prim, n, q);
    case OR_FLUX_GODUNOV: or_flux_godunov(cfg, prim, n, q); return 0;
    case OR_FLUX_CD:      or_flux_cd(cfg, prim, n, q); return 0;
    }
    return 1;
}

void or_flux(const or_config_t *cfg, int flux, const or_prim_t prim[2],
             const double n[3], double *q)
{
    (void)or_flux_err(cfg, flux, prim, n, q);
}

/* --------------------------------------------------------------- geometry
 * A tree has 8 vertices in z-order (the unit cube [0,1]^3 for the
 * single-tree forests).  Vertex coordinates follow p4est_qcoord_to_vertex
 * (p4est a9910fd, src/p4est_connectivity.c): trilinear blend of the 8 tree
 * vertices with weights w[1] = q / root_len, w[0] = 1 - w[1], summed z, y, x
 * outer to inner.  Called by the reference at src/charm_geom.c:26,36,196,
 * 267,283. */
const double or_unit_tree[8][3] = {
    {0,0,0},{1,0,0},{0,1,0},{1,1,0},{0,0,1},{1,0,1},{0,1,1},{1,1,1}};

static void or_qcoord_to_vertex(const double tv[8][3], int32_t x, int32_t y, int32_t z, double v[3])
{
    double wx[2], wy[2], wz[2], xf, yf;
    int xi, yi, zi, vi = 0;
    v[0] = v[1] = v[2] = 0.;
    wx[1] = (double) x / (double) OR_ROOT_LEN; wx[0] = 1. - wx[1];
    wy[1] = (double) y / (double) OR_ROOT_LEN; wy[0] = 1. - wy[1];
    wz[1] = (double) z / (double) OR_ROOT_LEN; wz[0] = 1. - wz[1];
    for (zi = 0; zi < 2; ++zi)
        for (yi = 0; yi < 2; ++yi) {
            yf = wz[zi] * wy[yi];
            for (xi = 0; xi < 2; ++xi) {
                xf = yf * wx[xi];
                v[0] += xf * tv[vi][0];
                v[1] += xf * tv[vi][1];
                v[2] += xf * tv[vi][2];
                vi++;
            }
        }
}

/* src/charm_geom.c:12-31 */
void or_quad_vertices_tv(const double tv[8][3], const int32_t q[4], double v[8][3])
{
    int32_t l = 1 << (OR_MAXLEVEL - q[3]);
    int i = 0, ix, iy, iz;
    for (iz = 0; iz < 2; iz++)
        for (iy = 0; iy < 2; iy++)
            for (ix = 0; ix < 2; ix++)
                or_qcoord_to_vertex(tv, q[0] + ix*l, q[1] + iy*l, q[2] + iz*l, v[i++]);
}

/* src/charm_geom.c:40-128 */
static void or_quad_gp_at_point(double vtx[8][3], const double ref[3], double gp[3], double *gj)
{
    int ix, iy, iz, vi;
    double wx[2], wy[2], wz[2], xf, yf, jac[3][3];
    wx[0] = 1.-ref[0]; wx[1] = 1.+ref[0];
    wy[0] = 1.-ref[1]; wy[1] = 1.+ref[1];
    wz[0] = 1.-ref[2]; wz[1] = 1.+ref[2];
    gp[0] = gp[1] = gp[2] = 0.;
    vi = 0;
    for (iz = 0; iz < 2; iz++) for (iy = 0; iy < 2; iy++) {
        yf = wz[iz]*wy[iy];
        for (ix = 0; ix < 2; ix++) {
            xf = yf*wx[ix]/8.;
            gp[0] += xf*vtx[vi][0]; gp[1] += xf*vtx[vi][1]; gp[2] += xf*vtx[vi][2];
            vi++;
        }
    }
    wx[0] = -1.; wx[1] = 1.;
    jac[0][0] = jac[1][0] = jac[2][0] = 0.;
    vi = 0;
    for (iz = 0; iz < 2; iz++) for (iy = 0; iy < 2; iy++) {
        yf = wz[iz]*wy[iy];
        for (ix = 0; ix < 2; ix++) {
            xf = yf*wx[ix]/8.;
            jac[0][0] += xf*vtx[vi][0]; jac[1][0] += xf*vtx[vi][1]; jac[2][0] += xf*vtx[vi][2];
            vi++;
        }
    }
    wx[0] = 1.-ref[0]; wx[1] = 1.+ref[0];
    wy[0] = -1.; wy[1] = 1.;
    jac[0][1] = jac[1][1] = jac[2][1] = 0.;
    vi = 0;
    for (iz = 0; iz < 2; iz++) for (iy = 0; iy < 2; iy++) {
        yf = wz[iz]*wy[iy];
        for (ix = 0; ix < 2; ix++) {
            xf = yf*wx[ix]/8.;
            jac[0][1] += xf*vtx[vi][0]; jac[1][1] += xf*vtx[vi][1]; jac[2][1] += xf*vtx[vi][2];
            vi++;
        }
    }
    wy[0] = 1.-ref[1]; wy[1] = 1.+ref[1];
    wz[0] = -1.; wz[1] = 1.;
    jac[0][2] = jac[1][2] = jac[2][2] = 0.;
    vi = 0;
    for (iz = 0; iz < 2; iz++) for (iy = 0; iy < 2; iy++) {
        yf = wz[iz]*wy[iy];
        for (ix = 0; ix < 2; ix++) {
            xf = yf*wx[ix]/8.;
            jac[0][2] += xf*vtx[vi][0]; jac[1][2] += xf*vtx[vi][1]; jac[2][2] += xf*vtx[vi][2];
            vi++;
        }
    }
    *gj = or_matr3_det(jac);
}

static const int or_ftv[6][4] = {{0,2,4,6},{1,3,5,7},{0,1,4,5},{2,3,6,7},{0,1,2,3},{4,5,6,7}};

/* base functions, src/charm_base_func.c:8-125 (P1 part) */
double or_bf(const or_geom_t *g, const double x[3], int k)
{
    switch (k) {
    case 0: return 1.;
    case 1: return (x[0]-g->c[0])/g->dh[0];
    case 2: return (x[1]-g->c[1])/g->dh[1];
    case 3: return (x[2]-g->c[2])/g->dh[2];
    }
    return 0;
}
double or_bf_dx(const or_geom_t *g, int k) { return k == 1 ? 1./g->dh[0] : 0.; }
double or_bf_dy(const or_geom_t *g, int k) { return k == 2 ? 1./g->dh[1] : 0.; }
double or_bf_dz(const or_geom_t *g, int k) { return k == 3 ? 1./g->dh[2] : 0.; }

/* charm_geom_quad_calc and helpers, src/charm_geom.c:130-211,255-268,
 * 443-530 */
void or_geom_calc_tv(const double tv[8][3], const int32_t q[4], or_geom_t *g)
{
    int32_t l = 1 << (OR_MAXLEVEL - q[3]), l2 = l / 2;
    double v[8][3], x[4][3], e[2][3], nl;
    int i, j, f, ig, t;
    const double s3 = 1./sqrt(3.);
    const double tq[8][3] = {
        {-s3,-s3,-s3},{ s3,-s3,-s3},{-s3, s3,-s3},{ s3, s3,-s3},
        {-s3,-s3, s3},{ s3,-s3, s3},{-s3, s3, s3},{ s3, s3, s3}};
    const int32_t fc[6][3] = {{0,l2,l2},{l,l2,l2},{l2,0,l2},{l2,l,l2},{l2,l2,0},{l2,l2,l}};

    for (f = 0; f < 6; f++) {
        or_qcoord_to_vertex(tv, q[0]+fc[f][0], q[1]+fc[f][1], q[2]+fc[f][2], g->fc[f]);
        for (i = 0; i < 4; i++)
            or_qcoord_to_vertex(tv, q[0] + l*(or_ftv[f][i]%2), q[1] + l*((or_ftv[f][i]/2)%2),
                                q[2] + l*(or_ftv[f][i]/4), x[i]);
        for (i = 0; i < 3; i++) { e[0][i] = x[1][i]-x[0][i]; e[1][i] = x[2][i]-x[0][i]; }
        or_vect_prod(e[0], e[1], g->n[f]);
        nl = or_vect_length(g->n[f]);
        for (i = 0; i < 3; i++) g->n[f][i] /= nl;
        /* 6-point rule on triangles (v0,v1,v2), (v1,v2,v3): :443-494 */
        {
            double a = 2.0/3.0, b = 1.0/6.0, pt[3][3];
            for (i = 0; i < OR_FGP; i++) g->face_gw[f][i] = 1./6.;
            for (t = 0; t < 2; t++) {
                for (j = 0; j < 3; j++) for (i = 0; i < 3; i++) pt[j][i] = x[t+j][i];
                double a1 = pt[0][0]-pt[2][0], a2 = pt[1][0]-pt[2][0], a3 = pt[2][0];
                double b1 = pt[0][1]-pt[2][1], b2 = pt[1][1]-pt[2][1], b3 = pt[2][1];
                double c1 = pt[0][2]-pt[2][2], c2 = pt[1][2]-pt[2][2], c3 = pt[2][2];
                double E = a1*a1+b1*b1+c1*c1, F = a2*a2+b2*b2+c2*c2, G = a1*a2+b1*b2+c1*c2;
                g->face_gj[f][3*t+0] = g->face_gj[f][3*t+1] = g->face_gj[f][3*t+2] = sqrt(E*F-G*G);
                g->face_gp[f][3*t+0][0] = a*a1+b*a2+a3;
                g->face_gp[f][3*t+0][1] = a*b1+b*b2+b3;
                g->face_gp[f][3*t+0][2] = a*c1+b*c2+c3;
                g->face_gp[f][3*t+1][0] = b*a1+a*a2+a3;
                g->face_gp[f][3*t+1][1] = b*b1+a*b2+b3;
                g->face_gp[f][3*t+1][2] = b*c1+a*c2+c3;
                g->face_gp[f][3*t+2][0] = b*a1+b*a2+a3;
                g->face_gp[f][3*t+2][1] = b*b1+b*b2+b3;
                g->face_gp[f][3*t+2][2] = b*c1+b*c2+c3;
            }
        }
        g->area[f] = 0.;
        for (j = 0; j < OR_FGP; j++) g->area[f] += g->face_gw[f][j]*g->face_gj[f][j];
    }
    or_qcoord_to_vertex(tv, q[0]+l2, q[1]+l2, q[2]+l2, g->c);
    or_quad_vertices_tv(tv, q, v);
    for (i = 0; i < OR_QGP; i++) {
        or_quad_gp_at_point(v, tq[i], g->quad_gp[i], &g->quad_gj[i]);
        g->quad_gw[i] = 1.;
    }
    for (j = 0; j < 3; j++) {
        double vmin = v[0][j], vmax = v[0][j];
        for (i = 1; i < 8; i++) {
            if (vmin > v[i][j]) vmin = v[i][j];
            if (vmax < v[i][j]) vmax = v[i][j];
        }
        g->dh[j] = vmax - vmin;
    }
    g->volume = 0.;
    for (i = 0; i < OR_QGP; i++) g->volume += g->quad_gw[i]*g->quad_gj[i];
    for (i = 0; i < 4; i++) for (j = 0; j < 4; j++) {
        g->a[i][j] = 0.;
        for (ig = 0; ig < OR_QGP; ig++)
            g->a[i][j] += g->quad_gw[ig]*g->quad_gj[ig]
                        * or_bf(g, g->quad_gp[ig], i) * or_bf(g, g->quad_gp[ig], j);
    }
    or_matr_inv(g->a, g->a_inv);
}

/* the single-tree forests: the unit cube */
void or_quad_vertices(const int32_t q[4], double v[8][3])
{
    or_quad_vertices_tv(or_unit_tree, q, v);
}

void or_geom_calc(const int32_t q[4], or_geom_t *g)
{
    or_geom_calc_tv(or_unit_tree, q, g);
}
