/*
 * charm_main.c -- the charm_3d program over this path (src/charm_3d.c:11-50
 * of the reference): read task.yaml from the working directory (or the
 * path given as the first argument), build the unit-cube forest at
 * MIN_LEVEL, set the initial condition, write the initial solution, run
 * charm_timesteps to TMAX and exit.  Exit status 0 on success, 1 with the
 * error message on stderr otherwise.
 */
#include <stdio.h>
#include <stdlib.h>
#include "charm_gpu.h"

static int fail(const char *what)
{
    fprintf(stderr, "charm_3d: %s: %s\n", what, charm_gpu_last_error());
    return 1;
}

int main(int argc, char **argv)
{
    const char *path = argc > 1 ? argv[1] : "task.yaml";
    static charm_task_t task;
    charm_forest_t *forest = NULL;
    charm_ctx_t *ctx = NULL;
    double *u = NULL;
    int64_t n;
    int rc = 1;
    if (charm_task_load_yaml(path, &task)) return fail(path);
    if (charm_task_forest(&task, &forest)) return fail("mesh");
    if (charm_ctx_create(forest, &task.cfg, &ctx)) { fail("context"); goto done; }
    n = charm_forest_num_leaves(forest);
    u = (double *) malloc((size_t) n * 4 * (4 + task.cfg.n_comp) * sizeof(double));
    if (!u) { fprintf(stderr, "charm_3d: out of memory\n"); goto done; }
    if (charm_task_initial_state(&task, forest, u, n) || charm_gpu_set_state(charm_ctx_gpu(ctx), u, n)) {
        fail("initial condition");
        goto done;
    }
    if (task.cfg.write_period > 0) {
        charm_write_solution(forest);
        if (charm_ctx_error(ctx)) { fail("output"); goto done; }
    }
    charm_timesteps(forest);
    if (charm_ctx_error(ctx)) { fail("time steps"); goto done; }
    rc = 0;
done:
    free(u);
    charm_ctx_destroy(ctx);
    charm_forest_destroy(forest);
    return rc;
}
