/*
 * charm_pointwise.c -- the reference's pointwise host entry points over a
 * forest's configuration: EOS functions (charm_mat_eos_ideal / _mix,
 * src/charm_eos.c:12-148), boundary-condition functions and their
 * dispatcher (charm_bnd_cond, charm_bnd_cond_fn_*, src/charm_bnd_cond.c:
 * 21-146), and the conservative <-> primitive conversions
 * (charm_param_cons_to_prim / _prim_to_cons / charm_prim_cpy,
 * src/charm_globals.c:183-240).
 *
 * They keep the reference's names, arguments and FP64 arithmetic for
 * callers that evaluate single states (drivers, boundary setup, output).
 * The device path does not call them: the stage kernels carry their own
 * restatements of the same formulas.
 */
#include <math.h>
#include <string.h>
#include "charm_internal.h"

#define GAS_R 8.31446261815324          /* src/charm_eos.c:8 */
#define CHARM_EPS 1.e-12

/* charm_comp_calc_cp (src/charm_globals.c:603-625): the power series when
 * cp_type is polynomial and ml_type 1, else cp[0] */
double charm_comp_cp(const charm_comp_desc_t *q, double t)
{
    double r = 0.0, tt = 1.0;
    int k;
    if (!(q->cp_type != 0 && q->ml_type == 1)) return q->cp[0];
    for (k = 0; k < q->n_cp; k++) { r += tt * q->cp[k]; tt *= t; }
    return r;
}

/* charm_comp_calc_cp_dt (src/charm_globals.c:628-649) */
double charm_comp_cp_dt(const charm_comp_desc_t *q, double t)
{
    double r = 0.0, tt = 1.0;
    int k;
    for (k = 1; k < q->n_cp; k++) { r += tt * q->cp[k] * k; tt *= t; }
    return r;
}

/* _calc_t (src/charm_eos.c:12-54): Newton on (cp(T) - R sum c_i / M_i) T = e
 * from T = 1 K, at most 100 steps, stop when (dT)^2 < 1e-12 */
double charm_calc_t(const charm_config_t *cfg, const double *c, double e)
{
    double tt = 1.0, m_ = 0.0, rm;
    int i, j;
    for (i = 0; i < cfg->n_comp; i++) m_ += c[i] / cfg->comp[i].m;
    rm = GAS_R * m_;
    for (j = 0; j < 100; j++) {
        double cp = 0.0, cp_dt = 0.0, ft, ft_dt, tt1;
        for (i = 0; i < cfg->n_comp; i++) {
            cp += c[i] * charm_comp_cp(&cfg->comp[i], tt);
            cp_dt += c[i] * charm_comp_cp_dt(&cfg->comp[i], tt);
        }
        ft = (cp - rm) * tt - e;
        ft_dt = (cp - rm) + cp_dt * tt;
        tt1 = tt - ft / ft_dt;
        if ((tt1 - tt) * (tt1 - tt) < CHARM_EPS) { tt = tt1; break; }
        tt = tt1;
    }
    return tt;
}

/* _charm_mat_eos_switch (src/charm_eos.c:57-96) */
static void eos_switch(charm_prim_t *p, int flag)
{
    switch (flag) {
    case 0:                                   /* p = p(r, e) */
        if (p->r < CHARM_EPS) p->r = CHARM_EPS;
        p->p = p->r * p->e * (p->gam - 1);
        p->cz = sqrt(p->gam * p->p / p->r);
        break;
    case 1:                                   /* e = e(r, p) */
        if (p->r < CHARM_EPS) p->r = CHARM_EPS;
        if (p->p < CHARM_EPS) p->p = CHARM_EPS;
        p->e = p->p / (p->r * (p->gam - 1));
        p->t = p->e / p->cv;
        break;
    case 2:                                   /* r = r(T, p) */
        if (p->p < CHARM_EPS) p->p = CHARM_EPS;
        p->r = p->p * p->m / (p->t * GAS_R);
        p->cz = sqrt(p->gam * p->p / p->r);
        break;
    case 3:                                   /* (T, p) -> (r, cz, e) */
        if (p->p < CHARM_EPS) p->p = CHARM_EPS;
        p->r = p->p * p->m / (p->t * GAS_R);
        p->cz = sqrt(p->gam * p->p / p->r);
        p->e = p->p / (p->r * (p->gam - 1));
        break;
    case 4:                                   /* (r, e) -> (p, cz, T) */
        if (p->r < CHARM_EPS) p->r = CHARM_EPS;
        p->p = p->r * p->e * (p->gam - 1);
        p->cz = sqrt(p->gam * p->p / p->r);
        p->t = p->e / p->cv;
        break;
    default:
        break;
    }
}

static const charm_config_t *forest_cfg(charm_forest_t *p4est, int flag)
{
    charm_ctx_t *c = charm_ctx_of(p4est);
    if (!c) return NULL;
    if (flag < 0 || flag > 4) {
        charm_set_error("eos: flag %d outside 0..4", flag);
        c->err = CHARM_ERR_ARG;
        return NULL;
    }
    return &c->cfg;
}

void charm_eos_ideal_cfg(const charm_config_t *cfg, charm_prim_t *p, int flag)
{
    const charm_comp_desc_t *q = &cfg->comp[0];
    double t = p->t;
    if (q->cp_type != 0 && (flag == 1 || flag == 4)) {
        p->t = charm_calc_t(cfg, p->c, p->e);
        t = p->t;
    }
    p->cp = charm_comp_cp(q, t);
    p->m = q->m;
    p->cv = p->cp - GAS_R / p->m;
    p->gam = p->cp / p->cv;
    eos_switch(p, flag);
}

void charm_eos_mix_cfg(const charm_config_t *cfg, charm_prim_t *p, int flag)
{
    double t = p->t, M_ = 0.0;
    int i;
    if (cfg->comp[0].cp_type != 0 && (flag == 1 || flag == 4)) {
        p->t = charm_calc_t(cfg, p->c, p->e);
        t = p->t;
    }
    p->cp = 0.0;
    for (i = 0; i < cfg->n_comp; i++) {
        M_ += p->c[i] / cfg->comp[i].m;
        p->cp += p->c[i] * charm_comp_cp(&cfg->comp[i], t);
    }
    p->m = 1.0 / M_;
    p->cv = p->cp - GAS_R / p->m;
    p->gam = p->cp / p->cv;
    eos_switch(p, flag);
}

/* src/charm_eos.c:98-118 */
void charm_mat_eos_ideal(charm_forest_t *p4est, charm_prim_t *p, int flag)
{
    const charm_config_t *cfg = forest_cfg(p4est, flag);
    if (cfg && p) charm_eos_ideal_cfg(cfg, p, flag);
}

/* src/charm_eos.c:120-148 */
void charm_mat_eos_mix(charm_forest_t *p4est, charm_prim_t *p, int flag)
{
    const charm_config_t *cfg = forest_cfg(p4est, flag);
    if (cfg && p) charm_eos_mix_cfg(cfg, p, flag);
}

/* src/charm_eos.c:150-154: the reference's table EOS stops the program with
 * "TABLE EOS IS NOT RELEASED"; here the error is recorded on the forest's
 * context (or as the last error) and p is left as it is */
void charm_mat_eos_table(charm_forest_t *p4est, charm_prim_t *p, int flag)
{
    charm_ctx_t *c = p4est ? charm_ctx_of(p4est) : NULL;
    (void) p; (void) flag;
    charm_set_error("eos: TABLE EOS IS NOT RELEASED");
    if (c) c->err = CHARM_ERR_UNSUPPORTED;
}

/* src/charm_eos.c:10 */
charm_real_t charm_eos_get_r(void)
{
    return GAS_R;
}

/* src/charm_bnd_cond.c:9-19 over the names of src/charm_globals.c:13-19:
 * the type whose name matches, CHARM_BND_UNKNOWN otherwise */
int charm_bnd_type_by_name(const char *name)
{
    static const char *const names[] = {"BOUND_INLET", "BOUND_OUTLET", "BOUND_WALL_SLIP",
                                        "BOUND_WALL_NO_SLIP", "BOUND_MASS_FLOW"};
    int i;
    for (i = 0; name && i < (int) (sizeof(names) / sizeof(names[0])); i++)
        if (!strcmp(names[i], name)) return i;
    return CHARM_BND_UNKNOWN;
}

/* src/charm_globals.c:226-240 */
void charm_prim_cpy(charm_prim_t *dest, charm_prim_t *src)
{
    if (dest != src) memcpy(dest, src, sizeof(*dest));
}

/* src/charm_bnd_cond.c:36-47 */
void charm_bnd_cond_fn_inlet(charm_prim_t *par_in, charm_prim_t *par_out, int8_t face, charm_real_t *param,
                             charm_real_t n[3])
{
    (void) face; (void) n;
    charm_prim_cpy(par_out, par_in);
    par_out->u = param[0];
    par_out->v = param[1];
    par_out->w = param[2];
    par_out->t = param[3];
    par_out->p = param[4];
}

/* src/charm_bnd_cond.c:50-53 */
void charm_bnd_cond_fn_outlet(charm_prim_t *par_in, charm_prim_t *par_out, int8_t face, charm_real_t *param,
                              charm_real_t n[3])
{
    (void) face; (void) param; (void) n;
    charm_prim_cpy(par_out, par_in);
}

/* src/charm_bnd_cond.c:56-75: v - 2 (v . n) n, subtracted in two steps */
void charm_bnd_cond_fn_wall_slip(charm_prim_t *par_in, charm_prim_t *par_out, int8_t face, charm_real_t *param,
                                 charm_real_t n[3])
{
    double v[3], svn, vv[3];
    int i;
    (void) face; (void) param;
    v[0] = par_in->u; v[1] = par_in->v; v[2] = par_in->w;
    charm_prim_cpy(par_out, par_in);
    svn = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
    for (i = 0; i < 3; i++) vv[i] = n[i] * svn;
    for (i = 0; i < 3; i++) { v[i] -= vv[i]; v[i] -= vv[i]; }
    par_out->u = v[0]; par_out->v = v[1]; par_out->w = v[2];
}

/* src/charm_bnd_cond.c:79-98: tangential velocity kept, wall temperature */
void charm_bnd_cond_fn_wall_no_slip(charm_prim_t *par_in, charm_prim_t *par_out, int8_t face,
                                    charm_real_t *param, charm_real_t n[3])
{
    double v[3], svn, vv[3];
    int i;
    (void) face;
    v[0] = par_in->u; v[1] = par_in->v; v[2] = par_in->w;
    charm_prim_cpy(par_out, par_in);
    svn = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
    for (i = 0; i < 3; i++) vv[i] = n[i] * svn;
    for (i = 0; i < 3; i++) v[i] -= vv[i];
    par_out->u = v[0]; par_out->v = v[1]; par_out->w = v[2];
    par_out->t = param[0];
}

/* the configuration of the boundary being evaluated: set by charm_bnd_cond
 * before it calls a BC function (the reference's glob_mat and
 * charm_get_p4est() globals, src/charm_bnd_cond.c:7,28) */
static __thread const charm_config_t *g_bnd_cfg;

/* src/charm_bnd_cond.c:101-146.  param = (mass flux, static p, total T,
 * direction cosines, outflow p, composition).  As in the reference, a
 * supersonic outflow face stores its pressure in param[6], which later
 * subsonic faces read (charm_bnd_cond passes the context's own parameters);
 * the device kernels keep param[6] fixed instead. */
void charm_bnd_cond_fn_mass_flow(charm_prim_t *par_in, charm_prim_t *par_out, int8_t face,
                                 charm_real_t *param, charm_real_t n[3])
{
    const charm_config_t *cfg = g_bnd_cfg;
    double cs[3], fd, vel, v2, k;
    int ic;
    (void) face;
    charm_prim_cpy(par_out, par_in);
    if (!cfg) {
        charm_set_error("bnd_cond_fn_mass_flow: no configuration (call through charm_bnd_cond)");
        return;
    }
    cs[0] = param[3]; cs[1] = param[4]; cs[2] = param[5];
    fd = n[0] * cs[0] + n[1] * cs[1] + n[2] * cs[2];
    if (fd > 0.) {
        vel = par_out->u * n[0] + par_out->v * n[1] + par_out->w * n[2];
        if (vel < par_out->cz) par_out->p = param[6];
        else param[6] = par_out->p;
    } else {
        for (ic = 0; ic < cfg->n_comp; ic++) par_out->c[ic] = param[7 + ic];
        if (cfg->eos == CHARM_EOS_IDEAL) charm_eos_ideal_cfg(cfg, par_out, 0);
        else charm_eos_mix_cfg(cfg, par_out, 0);
        v2 = par_out->u * par_out->u + par_out->v * par_out->v + par_out->w * par_out->w;
        if (v2 > par_out->cz * par_out->cz) par_out->p = param[1];
        k = pow(param[0] * (par_out->cp - par_out->cv) / (par_out->p * fd), 2.) / (2. * par_out->cp);
        par_out->t = (-1 + sqrt(1.0 + 4.0 * k * param[2])) / (2. * k);
        vel = fabs(param[0] / (par_out->r * fd));
        par_out->u = cs[0] * vel;
        par_out->v = cs[1] * vel;
        par_out->w = cs[2] * vel;
    }
}

/* src/charm_bnd_cond.c:21-33: the BC function of the face's boundary, then
 * EOS flag 3 (T, p) -> (r, cz, e) and e_tot */
void charm_bnd_cond(charm_forest_t *p4est, int treeid, int8_t face, charm_prim_t *par_in,
                    charm_prim_t *par_out, charm_real_t n[3])
{
    charm_ctx_t *c = charm_ctx_of(p4est);
    const charm_config_t *cfg;
    double *param;
    (void) treeid;
    if (!c) return;
    cfg = &c->cfg;
    if (face < 0 || face > 5 || cfg->bnd_type[(int) face] == CHARM_BND_PERIODIC) {
        charm_set_error("bnd_cond: face %d has no boundary condition", (int) face);
        c->err = CHARM_ERR_ARG;
        return;
    }
    /* the boundary's own parameters (bnd->params, which the mass-flow
     * condition updates) */
    param = c->cfg.bnd_param[(int) face];
    g_bnd_cfg = cfg;
    switch (cfg->bnd_type[(int) face]) {
    case CHARM_BND_INLET: charm_bnd_cond_fn_inlet(par_in, par_out, face, param, n); break;
    case CHARM_BND_OUTLET: charm_bnd_cond_fn_outlet(par_in, par_out, face, param, n); break;
    case CHARM_BND_WALL_SLIP: charm_bnd_cond_fn_wall_slip(par_in, par_out, face, param, n); break;
    case CHARM_BND_WALL_NO_SLIP: charm_bnd_cond_fn_wall_no_slip(par_in, par_out, face, param, n); break;
    case CHARM_BND_MASS_FLOW: charm_bnd_cond_fn_mass_flow(par_in, par_out, face, param, n); break;
    default:
        g_bnd_cfg = NULL;
        charm_set_error("bnd_cond: unknown boundary type %d", cfg->bnd_type[(int) face]);
        c->err = CHARM_ERR_ARG;
        return;
    }
    g_bnd_cfg = NULL;
    par_out->mat_id = par_in->mat_id;
    if (cfg->eos == CHARM_EOS_IDEAL) charm_eos_ideal_cfg(cfg, par_out, 3);
    else charm_eos_mix_cfg(cfg, par_out, 3);
    par_out->e_tot = par_out->e + 0.5 * (par_out->u * par_out->u + par_out->v * par_out->v +
                                         par_out->w * par_out->w);
}

/* src/charm_globals.c:183-208: EOS flag 4 of the configured material */
void charm_param_cons_to_prim(charm_forest_t *p4est, charm_prim_t *p, charm_cons_t *c)
{
    charm_ctx_t *x = charm_ctx_of(p4est);
    const charm_config_t *cfg;
    int i;
    if (!x || !p || !c) return;
    cfg = &x->cfg;
    p->mat_id = c->mat_id;
    p->r = 0.0;
    for (i = 0; i < cfg->n_comp; i++) p->r += c->rc[i];
    p->u = c->ru / p->r;
    p->v = c->rv / p->r;
    p->w = c->rw / p->r;
    p->e_tot = c->re / p->r;
    p->e = p->e_tot - 0.5 * (p->u * p->u + p->v * p->v + p->w * p->w);
    for (i = 0; i < cfg->n_comp; i++) {
        p->c[i] = c->rc[i] / p->r;
        if (p->c[i] < 0.) p->c[i] = 0.;
        if (p->c[i] > 1.) p->c[i] = 1.;
    }
    if (cfg->eos == CHARM_EOS_IDEAL) charm_eos_ideal_cfg(cfg, p, 4);
    else charm_eos_mix_cfg(cfg, p, 4);
}

/* src/charm_globals.c:211-223 */
void charm_param_prim_to_cons(charm_forest_t *p4est, charm_cons_t *c, charm_prim_t *p)
{
    charm_ctx_t *x = charm_ctx_of(p4est);
    int i;
    if (!x || !p || !c) return;
    c->mat_id = p->mat_id;
    c->ru = p->r * p->u;
    c->rv = p->r * p->v;
    c->rw = p->r * p->w;
    c->re = p->r * (p->e + 0.5 * (p->u * p->u + p->v * p->v + p->w * p->w));
    for (i = 0; i < x->cfg.n_comp; i++) c->rc[i] = p->r * p->c[i];
}
