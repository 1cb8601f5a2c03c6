/*
 * or_ns.c -- Navier-Stokes viscous and chemistry terms of the oracle
 * (FP64, TEST INFRASTRUCTURE ONLY).
 *
 * With use_visc = 0 and no reactions (the configuration of the Euler-type
 * runs), the reference's NS stage reduces to: chem_rhs = 0
 * (src/models/ns/charm_model_ns_chem.c:159-176) and no viscous pass
 * (src/models/ns/charm_model_ns_diff.c:86).
 */
#include "or_internal.h"

void or_ns_chem(or_sim_t *s)
{
    int64_t e;
    if (s->cfg.n_react == 0) {
        for (e = 0; e < s->n; e++) s->d[e].chem_rhs = 0.;
        return;
    }
    snprintf(s->err, sizeof(s->err), "chemistry not available in this oracle build");
}

void or_ns_diff(or_sim_t *s)
{
    if (!s->cfg.use_visc) return;
    snprintf(s->err, sizeof(s->err), "viscous terms not available in this oracle build");
}
