/*
 * charm_model.c -- the reference's plugin entry points over the device
 * context: charm_model_{euler,ns}_timestep_single, charm_model_*_get_dt,
 * charm_timesteps, charm_limiter[_bj], charm_calc_flux_*.
 *
 * The forest's user pointer is a charm_ctx_t, as p4est->user_pointer is the
 * reference's charm_ctx_t (src/charm_3d.c:36-37).  Single rank: the ghost
 * arguments are accepted and passed through untouched.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "charm_internal.h"

int charm_ctx_create(charm_forest_t *f, const charm_config_t *cfg, charm_ctx_t **out)
{
    charm_ctx_t *c;
    charm_mesh_t *m = NULL;
    int ce = cfg ? cfg->chunk_elems : 0, cf = cfg ? cfg->chunk_faces : 0, rc;
    if (!f || !cfg || !out) { charm_set_error("ctx_create: null argument"); return CHARM_ERR_ARG; }
    c = (charm_ctx_t *) calloc(1, sizeof(*c));
    if (!c) { charm_set_error("ctx_create: out of memory"); return CHARM_ERR_MEM; }
    c->cfg = *cfg;
    c->forest = f;
    if (ce <= 0 || cf <= 0) {
        if ((rc = charm_gpu_recommended_chunks(cfg, &ce, &cf))) { free(c); return rc; }
    }
    if ((rc = charm_gpu_create(cfg, &c->gpu)) ||
        (rc = charm_forest_flatten(f, ce, cf, &m)) ||
        (rc = charm_gpu_upload_mesh(c->gpu, m))) {
        charm_mesh_destroy(m);
        charm_gpu_destroy(c->gpu);
        free(c);
        return rc;
    }
    charm_mesh_destroy(m);
    c->chunk_elems = ce;
    c->chunk_faces = cf;
    charm_forest_set_user_pointer(f, c);
    *out = c;
    return CHARM_OK;
}

int charm_ctx_destroy(charm_ctx_t *c)
{
    if (!c) return CHARM_OK;
    if (c->forest && charm_forest_user_pointer(c->forest) == c)
        charm_forest_set_user_pointer(c->forest, NULL);
    charm_gpu_destroy(c->gpu);
    free(c);
    return CHARM_OK;
}

charm_gpu_ctx_t *charm_ctx_gpu(charm_ctx_t *c) { return c ? c->gpu : NULL; }

int charm_ctx_error(charm_ctx_t *c)
{
    int e;
    if (!c) { charm_set_error("null context"); return CHARM_ERR_ARG; }
    e = c->err;
    c->err = CHARM_OK;
    return e;
}

charm_ctx_t *charm_ctx_of(charm_forest_t *p4est)
{
    charm_ctx_t *c = (charm_ctx_t *) charm_forest_user_pointer(p4est);
    if (!c) charm_set_error("forest has no charm_ctx_t user pointer");
    return c;
}

/* get_dt returns TAU: the CFL branch after the early return is disabled in
 * the reference (src/models/euler/charm_model.c:37-54,
 * src/models/ns/charm_model.c:43-60); charm_gpu_get_dt exposes it */
charm_real_t charm_model_euler_get_dt(charm_forest_t *p4est)
{
    charm_ctx_t *c = charm_ctx_of(p4est);
    return c ? c->cfg.dt : 0.;
}

charm_real_t charm_model_ns_get_dt(charm_forest_t *p4est)
{
    charm_ctx_t *c = charm_ctx_of(p4est);
    return c ? c->cfg.dt : 0.;
}

static void timestep(charm_forest_t *p4est, int model, charm_real_t *dt)
{
    charm_ctx_t *c = charm_ctx_of(p4est);
    int rc;
    if (!c || !dt) return;
    /* adapt when REFINE_PERIOD divides a nonzero step number, then take dt
     * (src/models/euler/charm_model.c:468-490, src/models/ns/charm_model.c:199-215) */
    if (c->cfg.refine_period > 0 && c->timestep > 0 && c->timestep % c->cfg.refine_period == 0) {
        charm_adapt(p4est, NULL, NULL);
        if (c->err) return;
    }
    *dt = model == CHARM_MODEL_EULER ? charm_model_euler_get_dt(p4est) : charm_model_ns_get_dt(p4est);
    /* write out the solution before the step (src/models/euler/charm_model.c:507-510) */
    if (c->cfg.write_period > 0 && c->timestep % c->cfg.write_period == 0) {
        charm_write_solution(p4est);
        if (c->err) return;
    }
    rc = charm_gpu_step(c->gpu, model, *dt, 1);
    if (rc) c->err = rc;
}

/* src/models/euler/charm_model.c:454-540 (single rank: no repartition) */
void charm_model_euler_timestep_single(charm_forest_t *p4est, charm_real_t *dt,
                                       charm_ghost_t **ghost, charm_data_t **ghost_data)
{
    (void) ghost; (void) ghost_data;
    timestep(p4est, CHARM_MODEL_EULER, dt);
}

/* src/models/ns/charm_model.c:187-273 */
void charm_model_ns_timestep_single(charm_forest_t *p4est, charm_real_t *dt,
                                    charm_ghost_t **ghost, charm_data_t **ghost_data)
{
    (void) ghost; (void) ghost_data;
    timestep(p4est, CHARM_MODEL_NS, dt);
}

/* src/charm_timesteps.c:23-51: the loop advances t by dt until t >= TMAX
 * and logs every LOG_OUTPUT_STEP steps */
void charm_timesteps(charm_forest_t *p4est)
{
    charm_ctx_t *c = charm_ctx_of(p4est);
    charm_real_t dt;
    if (!c) return;
    dt = c->cfg.model == CHARM_MODEL_EULER ? charm_model_euler_get_dt(p4est) : charm_model_ns_get_dt(p4est);
    for (c->t = 0., c->timestep = 0; c->t < c->cfg.tmax; c->t += dt, c->timestep++) {
        if (c->cfg.model == CHARM_MODEL_EULER) charm_model_euler_timestep_single(p4est, &dt, NULL, NULL);
        else charm_model_ns_timestep_single(p4est, &dt, NULL, NULL);
        if (c->err) return;
        if (c->cfg.log_period > 0 && c->timestep % c->cfg.log_period == 0) {
            double ms = 0.;
            charm_gpu_last_timing(c->gpu, &ms, NULL);
            printf("STEP = %8d\tTIME = %16.8e\tTAU = %16.8e\tELAPSED_TIME = %16.8e\n",
                   c->timestep, c->t, dt, ms * 1e-3);
        }
    }
}

int charm_gmesh_timesteps(charm_gpu_ctx_t *g, const charm_gmesh_t *m, const charm_config_t *cfg)
{
    /* the reference's loop, t += dt while t < TMAX, with dt from get_dt (TAU,
     * src/models/euler/charm_model.c:30-38); the solution is written before
     * every step FILE_OUTPUT_STEP divides (src/models/euler/charm_model.c:
     * 507-510); runs of steps between outputs and log lines go to the device
     * in one call */
    const double dt = cfg ? cfg->dt : 0.0;
    double t = 0.0;
    int64_t n = 0, step = 0;
    double *u = NULL;
    int rc = CHARM_OK;
    if (!g || !m || !cfg || !(dt > 0.0)) { charm_set_error("gmesh_timesteps: bad arguments"); return CHARM_ERR_ARG; }
    charm_gmesh_info(m, &n, NULL);
    if (cfg->write_period > 0) {
        u = (double *) malloc((size_t) (n + 1) * 4 * (4 + cfg->n_comp) * sizeof(double));
        if (!u) { charm_set_error("gmesh_timesteps: out of memory"); return CHARM_ERR_MEM; }
    }
    while (t < cfg->tmax) {
        int64_t k = 0;
        double ms = 0.0;
        if (cfg->write_period > 0 && step % cfg->write_period == 0) {
            char name[512];
            snprintf(name, sizeof(name), "%.480s_%08lld", cfg->out_prefix[0] ? cfg->out_prefix : "charm_3d",
                     (long long) step);
            if ((rc = charm_gpu_get_state(g, u, n)) || (rc = charm_gmesh_write_vtk(name, m, u, cfg))) break;
        }
        /* steps up to the next output or log line, or to TMAX */
        do {
            t += dt;
            k++;
        } while (t < cfg->tmax && !(cfg->write_period > 0 && (step + k) % cfg->write_period == 0) &&
                 !(cfg->log_period > 0 && (step + k - 1) % cfg->log_period == 0));
        if ((rc = charm_gpu_step(g, cfg->model, dt, (int) k))) break;
        charm_gpu_last_timing(g, &ms, NULL);
        /* the reference logs after step s when LOG_OUTPUT_STEP divides s */
        if (cfg->log_period > 0 && (step + k - 1) % cfg->log_period == 0)
            printf("STEP = %8lld\tTIME = %16.8e\tTAU = %16.8e\tELAPSED_TIME = %16.8e\n", (long long) (step + k - 1),
                   t - dt, dt, ms * 1e-3 / (double) k);
        step += k;
    }
    free(u);
    return rc;
}

/* src/limiters/charm_limiter.c:9-16 */
void charm_limiter(charm_forest_t *p4est, charm_ghost_t *ghost, charm_data_t *ghost_data)
{
    charm_ctx_t *c = charm_ctx_of(p4est);
    if (c && c->cfg.limiter) charm_limiter_bj(p4est, ghost, ghost_data);
}

void charm_limiter_bj(charm_forest_t *p4est, charm_ghost_t *ghost, charm_data_t *ghost_data)
{
    charm_ctx_t *c = charm_ctx_of(p4est);
    int rc;
    (void) ghost; (void) ghost_data;
    if (!c) return;
    rc = charm_gpu_limiter_bj(c->gpu);
    if (rc) c->err = rc;
}

/* pointwise fluxes through the device flux kernel (one point) with the
 * reference's argument roles (include/charm_fluxes.h:19-32) */
static void flux_point(charm_forest_t *p4est, int kind, charm_prim_t prim[2], charm_real_t *qu,
                       charm_real_t *qv, charm_real_t *qw, charm_real_t *qe, charm_real_t qc[],
                       charm_real_t n[3])
{
    charm_ctx_t *c = charm_ctx_of(p4est);
    double row[2][5 + CHARM_MAX_COMP], q[4 + CHARM_MAX_COMP];
    int s, i, nc;
    if (!c) return;
    nc = c->cfg.n_comp;
    for (s = 0; s < 2; s++) {
        row[s][0] = prim[s].r; row[s][1] = prim[s].u; row[s][2] = prim[s].v;
        row[s][3] = prim[s].w; row[s][4] = prim[s].p;
        for (i = 0; i < nc; i++) row[s][5 + i] = prim[s].c[i];
    }
    if (charm_gpu_flux_batch(c->gpu, kind, 1, row[0], row[1], n, q)) { c->err = CHARM_ERR_HIP; return; }
    *qu = q[0]; *qv = q[1]; *qw = q[2]; *qe = q[3];
    for (i = 0; i < nc; i++) qc[i] = q[4 + i];
}

void charm_calc_flux_lf(charm_forest_t *p4est, charm_prim_t prim[2], charm_real_t *qu,
                        charm_real_t *qv, charm_real_t *qw, charm_real_t *qe,
                        charm_real_t qc[], charm_real_t n[3])
{
    flux_point(p4est, CHARM_FLUX_LF, prim, qu, qv, qw, qe, qc, n);
}

void charm_calc_flux_hllc(charm_forest_t *p4est, charm_prim_t prim[2], charm_real_t *qu,
                          charm_real_t *qv, charm_real_t *qw, charm_real_t *qe,
                          charm_real_t qc[], charm_real_t n[3])
{
    flux_point(p4est, CHARM_FLUX_HLLC, prim, qu, qv, qw, qe, qc, n);
}

void charm_calc_flux_godunov(charm_forest_t *p4est, charm_prim_t prim[2], charm_real_t *qu,
                             charm_real_t *qv, charm_real_t *qw, charm_real_t *qe,
                             charm_real_t qc[], charm_real_t n[3])
{
    flux_point(p4est, CHARM_FLUX_GODUNOV, prim, qu, qv, qw, qe, qc, n);
}

void charm_calc_flux_cd(charm_forest_t *p4est, charm_prim_t prim[2], charm_real_t *qu,
                        charm_real_t *qv, charm_real_t *qw, charm_real_t *qe,
                        charm_real_t qc[], charm_real_t n[3])
{
    flux_point(p4est, CHARM_FLUX_CD, prim, qu, qv, qw, qe, qc, n);
}
