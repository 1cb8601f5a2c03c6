/*
 * charm_main.c -- the charm_3d program over this path (src/charm_3d.c:11-50
 * of the reference): read task.yaml from the working directory (or the
 * path given as the first argument), build the mesh at MIN_LEVEL (the
 * unit-cube forest, or the multi-tree forest over the task's Gmsh mesh, read
 * relative to the task file), set the initial condition, write the initial
 * solution, run the time steps to TMAX and exit.  Exit status 0 on success,
 * 1 with the error message on stderr otherwise.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "charm_gpu.h"

static int fail(const char *what)
{
    fprintf(stderr, "charm_3d: %s: %s\n", what, charm_gpu_last_error());
    return 1;
}

/* a Gmsh task: the multi-tree mesh on one context, charm_gmesh_timesteps */
static int run_gmsh(charm_task_t *task, const char *path)
{
    char dir[512] = ".";
    const char *slash = strrchr(path, '/');
    charm_gmesh_t *gm = NULL;
    charm_gpu_ctx_t *g = NULL;
    double *u = NULL;
    int64_t n = 0;
    int rc = 1;
    if (slash) snprintf(dir, sizeof(dir), "%.*s", (int) (slash - path), path);
    if (charm_task_gmesh(task, dir, &gm)) return fail("mesh");
    if (task->cfg.refine_period > 0 && task->cfg.max_level > task->cfg.min_level)
        fprintf(stderr, "charm_3d: warning: adaptation across trees is not built; the Gmsh mesh runs at "
                        "MIN_LEVEL %d (REFINE_PERIOD %d, MAX_LEVEL %d not acted on)\n",
                task->cfg.min_level, task->cfg.refine_period, task->cfg.max_level);
    charm_gmesh_info(gm, &n, NULL);
    if (charm_gpu_create(&task->cfg, &g) || charm_gpu_upload_gmesh(g, gm)) { fail("context"); goto done; }
    u = (double *) malloc((size_t) n * 4 * (4 + task->cfg.n_comp) * sizeof(double));
    if (!u) { fprintf(stderr, "charm_3d: out of memory\n"); goto done; }
    if (charm_task_initial_state_gmesh(task, gm, u, n) || charm_gpu_set_state(g, u, n)) {
        fail("initial condition");
        goto done;
    }
    if (charm_gmesh_timesteps(g, gm, &task->cfg)) { fail("time steps"); goto done; }
    rc = 0;
done:
    free(u);
    charm_gpu_destroy(g);
    charm_gmesh_destroy(gm);
    return rc;
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
    if (!strcmp(task.mesh_files_type, "gmsh_msh")) return run_gmsh(&task, path);
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
