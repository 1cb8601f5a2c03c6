/*
 * charm_amr.c -- host side of the mesh adaptation: refine / coarsen
 * decisions from the device estimator and the adapt driver charm_adapt
 * (estimator -> flags -> refine / coarsen / 2:1 face balance -> flatten ->
 * upload with FP64 projection of the state).
 *
 * One difference from the reference, on purpose: p4est_coarsen_ext runs
 * after p4est_refine_ext on the refined forest, so it also visits the
 * families that refinement has just created, whose gradient field is never
 * computed (src/charm_amr.c:579-608 calls replace_quads, which does not set
 * par.a.grad_u).  Here the decisions are taken on the forest before
 * refinement, so a family is coarsened only if all eight members existed
 * before this adaptation; freshly refined leaves are kept.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "charm_internal.h"

int charm_amr_flags(int64_t n, const int32_t *leaves, const double *err2, const double *rho_c,
                    double max_err, int min_level, int max_level, int8_t *flags)
{
    const double ge2 = max_err * max_err;
    int64_t i;
    if (n < 0 || (n > 0 && (!leaves || !err2 || !rho_c || !flags))) {
        charm_set_error("amr_flags: bad arguments");
        return CHARM_ERR_ARG;
    }
    /* charm_refine_err_estimate (src/charm_amr.c:32-48) under
     * p4est_refine_ext's maxlevel (src/charm_amr.c:601-603) */
    for (i = 0; i < n; i++) {
        const int l = leaves[4*i + 3];
        const double vol = ldexp(1.0, -3 * l);
        flags[i] = (int8_t) ((err2[i] > ge2 * vol && l < max_level) ? 1 : 0);
    }
    /* charm_coarsen_err_estimate (src/charm_amr.c:90-142) on each complete
     * family: eight consecutive leaves of one level whose first member sits
     * at the parent's anchor (p4est_coarsen_ext, callbackorphans = 0) */
    for (i = 0; i + 8 <= n; i++) {
        const int32_t *q = &leaves[4*i];
        const int l = q[3];
        int32_t plen, h;
        int c, ok = 1;
        double vol, parentu = 0., tot = 0.;
        if (l <= min_level || l < 1) continue;
        plen = CHARM_ROOT_LEN >> (l - 1);
        h = plen >> 1;
        if (q[0] % plen || q[1] % plen || q[2] % plen) continue;
        for (c = 0; c < 8 && ok; c++) {
            const int32_t *r = &leaves[4*(i + c)];
            ok = r[3] == l && flags[i + c] == 0 &&
                 r[0] == q[0] + (c & 1) * h && r[1] == q[1] + ((c >> 1) & 1) * h &&
                 r[2] == q[2] + ((c >> 2) & 1) * h;
        }
        if (!ok) continue;
        vol = ldexp(1.0, -3 * l);
        for (c = 0; c < 8; c++) parentu += rho_c[i + c] / 8;
        for (c = 0; c < 8 && ok; c++) {
            const double d = parentu - rho_c[i + c];
            if (err2[i + c] > ge2 * vol) ok = 0;
            tot += err2[i + c];
            tot += d * d * vol;
        }
        if (ok && tot < ge2 * (vol * 8))
            for (c = 0; c < 8; c++) flags[i + c] = -1;
        i += 7;
    }
    return CHARM_OK;
}

int charm_amr_adjoint_weights(int64_t n, int nv, const int32_t *leaves, const double *grad, double *w)
{
    double mean = 0.0;
    int64_t i;
    if (n < 0 || nv < 1 || (n > 0 && (!leaves || !grad || !w))) {
        charm_set_error("amr_adjoint_weights: bad arguments");
        return CHARM_ERR_ARG;
    }
    for (i = 0; i < n; i++) {
        const double vol = ldexp(1.0, -3 * leaves[4*i + 3]);
        double s = 0.0;
        for (int v = 0; v < nv; v++) s += fabs(grad[((size_t) i * nv + v) * 4]);
        w[i] = s / vol;
        mean += w[i];
    }
    if (n > 0) mean /= (double) n;
    if (!(mean > 0.0)) {
        charm_set_error("amr_adjoint_weights: the gradient of the cell means is zero");
        return CHARM_ERR_ARG;
    }
    for (i = 0; i < n; i++) w[i] /= mean;
    return CHARM_OK;
}

/* src/charm_amr.c:575-610; weight (or NULL) scales the estimator per leaf */
static int adapt(charm_ctx_t *c, const double *weight)
{
    const int64_t n = charm_forest_num_leaves(c->forest);
    double *err2 = (double *) malloc((size_t) (n ? n : 1) * 2 * sizeof(double));
    int32_t *leaves = (int32_t *) malloc((size_t) (n ? n : 1) * 4 * sizeof(int32_t));
    int8_t *flags = (int8_t *) malloc((size_t) (n ? n : 1));
    int64_t *map = NULL, n_map = 0;
    charm_mesh_t *m = NULL;
    int rc;
    if (!err2 || !leaves || !flags) {
        free(err2); free(leaves); free(flags);
        charm_set_error("adapt: out of memory");
        return CHARM_ERR_MEM;
    }
    if (!(rc = charm_gpu_amr_indicator(c->gpu, err2, err2 + n, n)) && weight)
        for (int64_t i = 0; i < n; i++) err2[i] *= weight[i];
    if (!rc &&
        !(rc = charm_forest_get_leaves(c->forest, leaves)) &&
        !(rc = charm_amr_flags(n, leaves, err2, err2 + n, c->cfg.max_err, c->cfg.min_level,
                               c->cfg.max_level, flags)) &&
        !(rc = charm_forest_adapt(c->forest, flags, c->cfg.max_level, c->cfg.min_level, &map, &n_map)) &&
        !(rc = charm_forest_flatten(c->forest, c->chunk_elems, c->chunk_faces, &m)))
        rc = charm_gpu_remesh(c->gpu, m, map, n_map);
    charm_mesh_destroy(m);
    free(map); free(err2); free(leaves); free(flags);
    return rc;
}

void charm_adapt(charm_forest_t *p4est, charm_ghost_t *ghost, charm_data_t *ghost_data)
{
    charm_ctx_t *c = charm_ctx_of(p4est);
    int rc;
    (void) ghost; (void) ghost_data;
    if (!c) return;
    rc = adapt(c, NULL);
    if (rc) c->err = rc;
}

/* src/charm_amr.c:309-347: the reference refines recursively by
 * charm_refine_init_err_estimate, which reads the gradient par.a.grad_u that
 * nothing has computed before the first charm_adapt (:50-83 against
 * :581-599), so its initial mesh is undefined.  Here the call records
 * CHARM_ERR_UNSUPPORTED with that reason; an initial mesh is built by
 * charm_adapt passes on a projected initial state. */
void charm_adapt_init(charm_forest_t *p4est)
{
    charm_ctx_t *c = p4est ? charm_ctx_of(p4est) : NULL;
    charm_set_error("adapt_init: the reference's initial refinement reads an estimator gradient it never "
                    "computes; refine with charm_adapt on a projected initial state");
    if (c) c->err = CHARM_ERR_UNSUPPORTED;
}

int charm_adapt_weighted(charm_forest_t *p4est, const double *weight, int64_t n)
{
    charm_ctx_t *c = charm_ctx_of(p4est);
    if (!c || !weight || n != charm_forest_num_leaves(p4est)) {
        charm_set_error("adapt_weighted: no context, no weights or %lld weights for %lld leaves",
                        (long long) n, (long long) (p4est ? charm_forest_num_leaves(p4est) : 0));
        return CHARM_ERR_ARG;
    }
    return adapt(c, weight);
}
