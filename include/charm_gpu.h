/*
 * charm_gpu.h -- C-ABI of the MI355X-native CHARM_3D DG residual and
 * time-step path (libcharm3d.so).
 *
 * Three layers, all plain C types (no torch, no C++ in the signatures):
 *
 *  1. charm_forest_*  the host linear octree that replaces p4est for this
 *                     path (uniform fill, 2:1 face balance, flattening);
 *  2. charm_gpu_*     device context: mesh upload after every adapt/balance,
 *                     state transfer, per-operator launches, diagnostics;
 *  3. the reference's own entry points (charm_model_*_timestep_single,
 *     charm_model_*_get_dt, charm_timesteps, charm_limiter[_bj],
 *     charm_calc_flux_*) with p4est_t* replaced by charm_forest_t*.
 *
 * Every charm_forest_* / charm_gpu_* function returns CHARM_OK (0) or a
 * negative error code and never aborts; charm_gpu_last_error() returns the
 * message of the last failure on the calling thread.  The reference-named
 * functions keep the reference's void/real return types; on failure they
 * record the error (charm_gpu_last_error) instead of calling charm_abort
 * (reference: src/charm_globals.c:459-470).
 *
 * Threading: one host thread per context; every context owns one HIP stream
 * (default priority).  All device work of a context is ordered on it.
 */
#ifndef CHARM_GPU_H
#define CHARM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double charm_real_t;       /* include/charm_globals.h:16 */

enum {
    CHARM_OK = 0,
    CHARM_ERR_ARG = -1,
    CHARM_ERR_HIP = -2,
    CHARM_ERR_MEM = -3,
    CHARM_ERR_MESH = -4,
    CHARM_ERR_STATE = -5,      /* NaN / non-positive pressure detected */
    CHARM_ERR_UNSUPPORTED = -6
};

/* precision of the bulk arithmetic: FP64 throughout, FP32 throughout, or
 * mixed (FP64 state, update, RK combine and limiter; FP32 fluxes and
 * integrals; constant-cp Euler terms) */
enum { CHARM_FP64 = 0, CHARM_FP32 = 1, CHARM_FP32_MIX = 2 };
/* FLUX_TYPE (src/charm_init_yaml.cpp:352-379) */
enum { CHARM_FLUX_LF = 0, CHARM_FLUX_HLLC = 1, CHARM_FLUX_GODUNOV = 2, CHARM_FLUX_CD = 3 };
/* boundary types (include/charm_globals.h:432-439) plus periodic links */
enum { CHARM_BND_INLET = 0, CHARM_BND_OUTLET = 1, CHARM_BND_WALL_SLIP = 2,
       CHARM_BND_WALL_NO_SLIP = 3, CHARM_BND_MASS_FLOW = 4, CHARM_BND_UNKNOWN = 5,
       CHARM_BND_PERIODIC = 100 };
/* EOS (src/charm_eos.c:98-154) */
enum { CHARM_EOS_IDEAL = 0, CHARM_EOS_MIX = 1 };
/* MODEL name (src/charm_init_yaml.cpp:394-413) */
enum { CHARM_MODEL_EULER = 0, CHARM_MODEL_NS = 1 };

#define CHARM_MAX_COMP 8

/* one species, the fields of charm_comp_t used by this path
 * (include/charm_globals.h:40-56) */
typedef struct {
    double m;                  /* molar mass */
    int    cp_type;            /* 0 const, 1 polynomial */
    int    ml_type;            /* 0 const, 1 Sutherland */
    int    n_cp;
    double cp[8];              /* cp = sum cp[i] T^i */
    double ml0, kp0, t0, ts, h0;
    int    kp_type;            /* 0 const, 1 Sutherland */
    double sig, ek;            /* Lennard-Jones sigma (Angstrom), eps / k (K): diffusion */
} charm_comp_desc_t;

/* one reaction (include/charm_globals.h:102-110): species indices (-1:
 * none) and k = a T^n exp(-e / (R T)) */
typedef struct {
    int    left[3], right[3];
    double a, e, n;
} charm_reaction_desc_t;

/* the task.yaml keys this path consumes (src/charm_init_yaml.cpp:333-426) */
typedef struct {
    int    precision;          /* CHARM_FP64 / CHARM_FP32 */
    int    model;              /* CHARM_MODEL_* */
    int    flux;               /* CHARM_FLUX_* */
    int    limiter;            /* 0 NONE, 1 BJ */
    int    eos;                /* CHARM_EOS_* */
    int    n_comp;
    charm_comp_desc_t comp[CHARM_MAX_COMP];
    int    bnd_type[6];        /* per cube face -x,+x,-y,+y,-z,+z */
    double bnd_param[6][16];
    double grav[3];
    double cfl;                /* CFL */
    double dt;                 /* TAU */
    double tmax;               /* TMAX */
    int    log_period;         /* LOG_OUTPUT_STEP (0: no log) */
    int    chunk_elems;        /* elements per workgroup chunk (0: default) */
    int    chunk_faces;        /* max faces per chunk (0: default) */
    /* adaptation (src/charm_init_yaml.cpp:383-387) */
    double max_err;            /* MAX_ERROR */
    int    min_level;          /* MIN_LEVEL */
    int    max_level;          /* MAX_LEVEL */
    int    refine_period;      /* REFINE_PERIOD (0: no adaptation) */
    /* Navier-Stokes model (src/charm_init_yaml.cpp:394-426) */
    int    use_visc;           /* VISCOSITY */
    int    use_diff;           /* DIFFUSION */
    double t_ref;              /* T_REF of the enthalpy */
    int    n_react;
    charm_reaction_desc_t react[4];
    int    chem_parity;        /* 1: the reference's chem_w[species] index quirk
                                  (src/models/ns/charm_model_ns_chem.c:97) */
    int    chem_energy_molar;  /* energy source of the chemistry: 0 the reference's
                                  sum_i w_i h_i, molar rate (mol/m^3/s) times mass
                                  enthalpy (J/kg) (src/models/ns/charm_model_ns_chem.c:
                                  66-108); 1 sum_i w_i M_i h_i (J/m^3/s), the rate times
                                  the molar enthalpy */
    double visc_br2_eta;       /* viscous flux on interior faces: 0 the reference's BR1
                                  (src/models/ns/charm_model_ns_diff_integrals.c:202-382);
                                  > 0 BR2 with this penalty factor: each side's stress is
                                  its own-trace auxiliary field plus eta times the face's
                                  local lifting (no counterpart in the reference) */
    /* output (src/charm_init_yaml.cpp:388, include/charm_globals.h:294) */
    int    write_period;       /* FILE_OUTPUT_STEP: VTK files every that many steps (0: none) */
    char   comp_name[CHARM_MAX_COMP][32];   /* components' names: cell data C_<name> */
    char   out_prefix[256];    /* file prefix ("" : "charm_3d", CHARM_STRING) */
} charm_config_t;

/* ------------------------------------------------------------ forest */
typedef struct charm_forest charm_forest_t;   /* replaces p4est_t */
typedef struct charm_mesh   charm_mesh_t;     /* flattened leaves + faces */

/* uniform fill at `level` of the unit cube (p4est_new_ext(..., min_level,
 * fill_uniform = 1, ...), src/charm_3d.c:36-37) */
int charm_forest_new_uniform(int level, const int periodic[3], charm_forest_t **out);
/* forest from explicit leaves (x, y, z, level) in root units (2^19) */
int charm_forest_new_leaves(int64_t n, const int32_t *leaves, const int periodic[3],
                            charm_forest_t **out);
int charm_forest_destroy(charm_forest_t *f);
int64_t charm_forest_num_leaves(const charm_forest_t *f);
int charm_forest_get_leaves(const charm_forest_t *f, int32_t *out /* [n][4] */);
/* refine (+1) / coarsen (-1, whole families only) by per-leaf flags, then
 * 2:1 face balance (p4est_refine_ext / coarsen_ext / balance_ext(FACE),
 * src/charm_amr.c:601-608).  `map` (may be NULL) receives for every new leaf
 * the index of the old leaf it came from (refined children: the parent;
 * coarsened parent: -1 - first child). */
int charm_forest_adapt(charm_forest_t *f, const int8_t *flags, int max_level, int min_level,
                       int64_t **map, int64_t *n_map);
int charm_forest_balance(charm_forest_t *f);
void *charm_forest_user_pointer(const charm_forest_t *f);
void charm_forest_set_user_pointer(charm_forest_t *f, void *p);

/* flatten the leaves into the element/face arrays the kernels consume */
int charm_forest_flatten(const charm_forest_t *f, int chunk_elems, int chunk_faces,
                         charm_mesh_t **out);
int charm_mesh_destroy(charm_mesh_t *m);
int64_t charm_mesh_num_faces(const charm_mesh_t *m);
int64_t charm_mesh_num_chunks(const charm_mesh_t *m);
/* face records over all chunks: a face whose leaves lie in two chunks is
 * held (and evaluated) by both; ratio to charm_mesh_num_faces = the stage
 * kernel's face-work overhead */
int64_t charm_mesh_num_chunk_faces(const charm_mesh_t *m);
/* the workgroup chunks: ch_ebeg / ch_fbeg [nch + 1] leaf and face-record
 * ranges, cf4 [ncf][4] the chunk-local face records (sorted by axis, then
 * kind, within a chunk), slot_ptr [n + 1] and slots [slot_ptr[n]] the
 * leaves' face incidences as (chunk-local face << 1 | side) */
int charm_mesh_get_chunks(const charm_mesh_t *m, int32_t *ch_ebeg, int32_t *ch_fbeg, int32_t *cf4,
                          int32_t *slot_ptr, uint16_t *slots);
/* faces: [nf][4] = (e0, e1 (or -1-domain_face), info, 0) with info =
 * dir0 | kind << 3 | subface << 5; kind 0 conforming, 1 side 0 hanging
 * (small), 2 side 1 hanging (small), 3 boundary */
int charm_mesh_get_faces(const charm_mesh_t *m, int32_t *out);
/* limiter neighbour table: [n][6] */
int charm_mesh_get_neighbors(const charm_mesh_t *m, int32_t *out);

/* FP64 Taylor-basis mass matrix of a level-`level` cube by the 8-point
 * Gauss rule and its Gauss-Jordan inverse (charm_geom_quad_calc,
 * src/charm_geom.c:496-530; charm_matr_inv, src/charm_globals.c:285-394) */
int charm_level_mass_inverse(int level, double a[4][4], double ainv[4][4]);

/* ------------------------------------------------------------ device */
typedef struct charm_gpu_ctx charm_gpu_ctx_t;

const char *charm_gpu_last_error(void);
int charm_gpu_device_count(int *n);
int charm_gpu_create(const charm_config_t *cfg, charm_gpu_ctx_t **out);
int charm_gpu_destroy(charm_gpu_ctx_t *ctx);
/* call after every adapt/balance: uploads the flattened mesh */
int charm_gpu_upload_mesh(charm_gpu_ctx_t *ctx, const charm_mesh_t *mesh);
/* state as FP64 [n][V][4] on the host (V = 4 + n_comp, order ru rv rw re
 * rc[0..]); converted to the context precision on the device */
int charm_gpu_set_state(charm_gpu_ctx_t *ctx, const double *u, int64_t n);
int charm_gpu_get_state(charm_gpu_ctx_t *ctx, double *u, int64_t n);
/* int_* integrals after the zero + convective pass (the value the
 * reference's update consumes, src/models/euler/charm_model.c:519-525),
 * FP64 [n][V][4] */
int charm_gpu_residual(charm_gpu_ctx_t *ctx, double *out, int64_t n);
/* Barth-Jespersen limiter on the current state */
int charm_gpu_limiter_bj(charm_gpu_ctx_t *ctx);
/* whole time steps of the configured model with a fixed dt */
int charm_gpu_run(charm_gpu_ctx_t *ctx, double dt, int nsteps);
/* as charm_gpu_run with an explicit model (CHARM_MODEL_EULER: forward Euler
 * + limiter; CHARM_MODEL_NS: SSP-RK3 with the limiter inside each stage) */
int charm_gpu_step(charm_gpu_ctx_t *ctx, int model, double dt, int nsteps);
/* chunk sizes (elements, faces) whose LDS footprint fits a workgroup for
 * this configuration */
int charm_gpu_recommended_chunks(const charm_config_t *cfg, int *chunk_elems, int *chunk_faces);
/* average device time of `reps` back-to-back launches of the stage kernel
 * and of the limiter kernel on the current state (HIP events, ms) */
int charm_gpu_time_kernels(charm_gpu_ctx_t *ctx, double dt, int reps, double *stage_ms,
                           double *limiter_ms);
/* average device time of a plain copy of the state (read + write bytes in
 * *bytes): the bandwidth reference for the kernels above */
int charm_gpu_time_copy(charm_gpu_ctx_t *ctx, int reps, double *copy_ms, double *bytes);
/* CFL time step: mode 0 = CFL*volume/(|u|+c) (the reference's disabled
 * branch, src/models/euler/charm_model.c:15-30), mode 1 = CFL*h/(|u|+c) */
int charm_gpu_get_dt(charm_gpu_ctx_t *ctx, int mode, double *dt);
/* FP64 sums_e V_e U[e][v][0] */
int charm_gpu_conservation_sums(charm_gpu_ctx_t *ctx, double *out);
/* Allocation-free variant of charm_gpu_conservation_sums (no counterpart in
 * the reference; for callers that cannot allocate: code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The V FP64 sums
 * of the current state go to out_dev (device memory, V doubles); the block
 * partials live in a static array of the library, one per GPU (1.2 MB, held
 * until the process ends), which limits the mesh to 4,194,304 leaves (larger:
 * CHARM_ERR_ARG).  The work is only enqueued on the context's stream: no
 * allocation, memset, copy or synchronisation; wait on the stream (e.g.
 * charm_gpu_synchronize) before reading out_dev.  Calls that overlap on one
 * GPU race on the shared array. */
int charm_gpu_conservation_sums_noalloc(charm_gpu_ctx_t *ctx, double *out_dev);
/* 1 if a NaN / non-positive pressure was met since the last call */
int charm_gpu_state_flag(charm_gpu_ctx_t *ctx, int *flag);
int charm_gpu_synchronize(charm_gpu_ctx_t *ctx);
/* device time of the last charm_gpu_run, split per kernel family (ms) */
int charm_gpu_last_timing(charm_gpu_ctx_t *ctx, double *stage_ms, double *limiter_ms);

/* adaptation, device side.  amr_indicator: per leaf the FP64 interpolation
 * error estimate err2 = |grad rho|^2 / 12 V^(5/3) of the Green-Gauss density
 * gradient and the cell-mean density rho_c (src/charm_amr.c:12-30,350-599).
 * remesh: uploads the mesh of the adapted forest and projects the state
 * onto it in FP64 (charm_replace_quads, src/charm_amr.c:206-307); `map` is
 * the old->new map of charm_forest_adapt and is checked against the leaves
 * of the previous mesh before any kernel runs. */
int charm_gpu_amr_indicator(charm_gpu_ctx_t *ctx, double *err2, double *rho_c, int64_t n);
int charm_gpu_remesh(charm_gpu_ctx_t *ctx, const charm_mesh_t *mesh, const int64_t *map, int64_t n_map);

/* Navier-Stokes pieces on the current state, for checks: ns_aux = the BR1
 * auxiliary stress and heat flux, FP64 [n][9][4] (tau xx yy zz xy xz yz,
 * q x y z; src/models/ns/charm_model_ns_diff_grad.c:485-507); ns_chem = the
 * chemistry applied to the state with step dt and its energy source [n]
 * (src/models/ns/charm_model_ns_chem.c:111-202). */
int charm_gpu_ns_aux(charm_gpu_ctx_t *ctx, double *aux, int64_t n);
int charm_gpu_ns_chem(charm_gpu_ctx_t *ctx, double dt, double *chem_rhs, int64_t n);

/* Discrete adjoint (no counterpart in the reference): nsteps SSP-RK3 steps
 * (Euler terms, no limiter, constant cp, 1-2 species, LF / HLLC / CD) from
 * the current state with a fixed dt; J = sum_{n=1..N} dt Q(u^n), Q(u) = int
 * of the pressure over the leaves whose centre lies in [lo, hi].
 * charm_gpu_objective runs the forward pass and returns J.
 * charm_gpu_adjoint also returns dJ/du^0 (FP64 [n][V][4], may be NULL) and
 * dJ/d(inlet u, T, p) of boundary face theta_face (-1: none) by reverse
 * passes over checkpoints stored every ckpt steps (each segment's stage
 * states are recomputed, except the last segment's, which the forward pass
 * keeps), and the device times of the forward and the
 * reverse pass.  ckpt = 0 keeps every stage state of the forward pass
 * instead (3 nsteps + 1 states of device memory, no recompute).  Both leave
 * the state at u^N. */
int charm_gpu_objective(charm_gpu_ctx_t *ctx, double dt, int nsteps, const double lo[3],
                        const double hi[3], double *J);
/* The stage VJP the adjoint sweeps with, on its own (no counterpart in the
 * reference): for the forward-Euler stage S(u) = u - dt A^-1 int(u) (the
 * Euler model's step, charm_gpu_step with CHARM_MODEL_EULER, no limiter),
 * out = S_u^T x = x + (d int / d u)^T (-dt A^-1 x): the transposed volume,
 * interior-face, hanging-face and boundary-face linearisations and the mass
 * matrix solve.  u, x, out FP64 [n][V][4] on the host; the context's state
 * is unchanged.  Same configurations as the adjoint. */
int charm_gpu_stage_vjp(charm_gpu_ctx_t *ctx, double dt, const double *u, const double *x, double *out,
                        int64_t n);
int charm_gpu_adjoint(charm_gpu_ctx_t *ctx, double dt, int nsteps, int ckpt, const double lo[3],
                      const double hi[3], int theta_face, double *J, double *grad_u0,
                      double *grad_theta, double *fwd_ms, double *adj_ms);

/* refine (+1) / coarsen (-1 on all eight members of a family) / keep (0)
 * decisions from the estimator, with the criteria of
 * charm_refine_err_estimate / charm_coarsen_err_estimate
 * (src/charm_amr.c:32-48,90-142) and p4est_refine_ext's max_level cap.
 * leaves [n][4] in Morton order (charm_forest_get_leaves). */
int charm_amr_flags(int64_t n, const int32_t *leaves, const double *err2, const double *rho_c,
                    double max_err, int min_level, int max_level, int8_t *flags);

/* Adjoint-weighted refinement (no counterpart in the reference, which has
 * no adjoint; for drivers that refine toward an output functional).
 * charm_amr_adjoint_weights: from dJ/du0 of charm_gpu_adjoint (grad
 * [n][nv][4], FP64), w_i = lambda_i / mean(lambda) with lambda_i =
 * sum_v |dJ/dU_{i,v,0}| / V_i, the sensitivity of J to leaf i's cell means
 * per unit volume.  charm_adapt_weighted: charm_adapt with the estimator
 * err2_i multiplied by weight_i in both the refine and the coarsen
 * criteria (weight [n], n = the forest's leaves). */
int charm_amr_adjoint_weights(int64_t n, int nv, const int32_t *leaves, const double *grad, double *w);
int charm_adapt_weighted(charm_forest_t *p4est, const double *weight, int64_t n);

/* pointwise flux on the device for a batch of point pairs: prim rows are
 * (r, u, v, w, p, c[0..n_comp-1]); normals [n][3]; out [n][4 + n_comp] */
int charm_gpu_flux_batch(charm_gpu_ctx_t *ctx, int flux, int64_t n, const double *prim_l,
                         const double *prim_r, const double *normals, double *out);
/* its vector-Jacobian product (no counterpart in the reference): for qbar
 * [n][4 + n_comp], prim_l_bar / prim_r_bar [n][5 + n_comp] = d(qbar . q) /
 * d(prim rows), FP64 on dual numbers; LF, HLLC and CD, constant cp, 1-2
 * species */
int charm_gpu_flux_vjp_batch(charm_gpu_ctx_t *ctx, int flux, int64_t n, const double *prim_l,
                             const double *prim_r, const double *normals, const double *qbar,
                             double *prim_l_bar, double *prim_r_bar);

/* No counterpart in the reference; for callers that check the adjoint's
 * central differences, whose objective has kinks where a perturbation moves
 * a face point across a branch of the HLLC flux.  For two FP64 constant-cp
 * contexts on the same octree mesh: at every interior face point the branch
 * hllc_x takes (the shock factor on either side, the wave-speed swap, the
 * upwind and star regions, the S* clamps, the PVRS clamp) for each context's
 * current state; points = the interior face points, flips = those whose
 * branch differs between the two states, per_bit [9] (may be NULL) = those
 * where each part differs: 0 p* > p_L, 1 p* > p_R, 2 the wave speeds
 * swapped, 3 S_L >= 0, 4 S_R <= 0, 5 S* >= 0, 6 S* clamped to S_L, 7 to S_R,
 * 8 p_pvrs <= 0. */
int charm_gpu_hllc_branch_flips(charm_gpu_ctx_t *a, charm_gpu_ctx_t *b, int64_t *points, int64_t *flips,
                                int64_t *per_bit);

/* VTK output (src/charm_output.c:10-126): <name>_0000.vtu (one
 * UnstructuredGrid piece, raw appended data) and <name>.pvtu; points are the
 * shared leaf corners, cells hexahedra with cell data treeid, level,
 * mpirank and the cell-centre fields R P E E_TOT U V W T C_<species> of the
 * state u (FP64 [n][V][4]) by cons_to_prim with EOS flag 4 */
int charm_write_vtk(const char *name, int64_t n, const int32_t *leaves, const double *u,
                    const charm_config_t *cfg);
/* ------------------------------------------------------------ task.yaml */
/* one initial-condition region (charm_reg_t, include/charm_globals.h) */
typedef struct {
    int    id, cell_type, mat_id;
    char   name[64];
    double v[3], t, p, grav[3];
    double c[CHARM_MAX_COMP];          /* mass fractions by component index */
} charm_region_desc_t;
/* what charm_init_context_yaml reads (src/charm_init_yaml.cpp:333-426) */
typedef struct {
    charm_config_t cfg;                /* control, components, materials, model, reactions */
    int    repartition_period;         /* REPARTITION_PERIOD (single rank: unused) */
    char   bnd_name[6][64];            /* boundaries by cube side */
    int    n_regions;
    charm_region_desc_t region[16];
    char   mesh_files_type[32];        /* mesh.files_type */
    char   mesh_name[256];             /* mesh.name */
    int    mesh_periodic[3];           /* unit_cube: mesh.periodic [x, y, z] */
} charm_task_t;
/* Parse a task.yaml (the reference's keys; yaml-cpp's role is taken by a
 * parser of the YAML subset task files use).  Boundaries of a unit-cube mesh
 * are bound to its sides by name: LEFT -x, RIGHT +x, FRONT -y, BACK +y,
 * BOTTOM -z, TOP +z; a Gmsh mesh's boundaries (at most 6) take the
 * configuration's slots 0-5 in the order of the file.  This build's own keys: control.PRECISION (FP64, FP32, FP32_MIX;
 * default FP64), mesh.files_type unit_cube with mesh.periodic.  Errors
 * (with the reference's messages) are returned, never fatal. */
int charm_task_load_yaml(const char *path, charm_task_t *task);
/* the uniform forest at MIN_LEVEL (src/charm_3d.c:36-37) of a unit-cube
 * task; a Gmsh task's mesh is charm_task_gmesh (CHARM_ERR_UNSUPPORTED here) */
int charm_task_forest(const charm_task_t *task, charm_forest_t **out);
/* charm_init_initial_condition (src/charm_init.c:12-73): the region of cell
 * type 0 through EOS flag 3 into coefficient 0, slopes zero; u FP64 [n][V][4] */
int charm_task_initial_state(const charm_task_t *task, const charm_forest_t *f, double *u, int64_t n);
/* ------------------------------------------------- Gmsh connectivity
 * charm_conn_reader_msh (src/charm_connectivity.c:315-519) without p4est: a
 * Gmsh 2.x ASCII mesh of hexahedra (type 5) and boundary quadrangles (type
 * 3) as multi-tree connectivity.  Tree corners are in z-order; tree_to_face
 * holds the neighbour's face + 6 * orientation as p4est's connectivity
 * (orientation: the face corner of the higher face number that meets corner
 * 0 of the lower one); a boundary face links to itself.  The kernels of this
 * library run on the axis-aligned unit-cube octree: a connectivity is read,
 * checked and bound to a task's boundaries and regions, not time-stepped. */
typedef struct charm_conn charm_conn_t;
typedef struct {
    int  dim, id;                      /* $PhysicalNames: dimension, physical id */
    char name[64];
} charm_conn_patch_t;
int  charm_conn_read_msh(const char *path, charm_conn_t **out);
void charm_conn_destroy(charm_conn_t *c);
int  charm_conn_info(const charm_conn_t *c, int64_t *num_vertices, int64_t *num_trees, int *num_patches);
/* any output may be NULL: vertices [nv][3], tree_to_vertex [nt][8],
 * tree_to_tree [nt][6], tree_to_face [nt][6], tree_to_patch [nt][6] (physical
 * id of the boundary quadrangle on the face, -1 none), tree_region [nt]
 * (physical id of the hexahedron) */
int  charm_conn_get(const charm_conn_t *c, double *vertices, int32_t *tree_to_vertex, int32_t *tree_to_tree,
                    int8_t *tree_to_face, int32_t *tree_to_patch, int32_t *tree_region);
int  charm_conn_patch(const charm_conn_t *c, int i, charm_conn_patch_t *p);
/* patches and regions by name against a task (charm_conn_bnd_find_by_id /
 * charm_conn_reg_find_by_id, src/charm_connectivity.c:178-230):
 * tree_bnd [nt][6] = the task's boundary slot (0..5) or -1 for faces without
 * a patch, tree_reg [nt] = index into task->region; a name the task lacks is
 * CHARM_ERR_MESH with the reference's message */
int  charm_conn_bind_task(const charm_conn_t *c, const charm_task_t *task, int32_t *tree_bnd, int32_t *tree_reg);
/* ------------------------------------------------- multi-tree meshes
 * The forest the reference builds over a connectivity, refined uniformly
 * (p4est_new_ext at MIN_LEVEL, src/charm_3d.c:36-37), with every leaf's
 * trilinear geometry (charm_geom_quad_calc, src/charm_geom.c:12-172,496-530)
 * in FP64 on the host.  Conforming: no refinement across trees.
 * vertices [nv][3], tree_to_vertex [nt][8] (z-order corners), tree_bnd
 * [nt][6] = the boundary slot (0..5, an index into charm_config_t's
 * bnd_type / bnd_param) of a tree face, or -1 where it meets another tree;
 * level 0..6.  Leaves are numbered tree by tree, Morton order inside a tree. */
typedef struct charm_gmesh charm_gmesh_t;
int  charm_gmesh_build(int64_t n_vertices, const double *vertices, int64_t n_trees, const int32_t *tree_to_vertex,
                       const int32_t *tree_bnd, int level, charm_gmesh_t **out);
void charm_gmesh_destroy(charm_gmesh_t *m);
int  charm_gmesh_info(const charm_gmesh_t *m, int64_t *n_leaves, int64_t *n_faces);
/* per leaf: tree, x, y, z, level (tree coordinates in units of 2^-19) */
int  charm_gmesh_leaves(const charm_gmesh_t *m, int32_t *leaves);
/* per leaf 23 doubles: centre[3], dh[3], volume, A^-1[4][4] */
int  charm_gmesh_geometry(const charm_gmesh_t *m, double *out);
/* per face: side-0 leaf, side-1 leaf or -1 - boundary slot, side-0 face, side-1 face (-1 on a boundary) */
int  charm_gmesh_faces(const charm_gmesh_t *m, int32_t *faces);
/* the context runs on the mesh from now on (replacing a charm_gpu_upload_mesh
 * octree): the Euler and SSP-RK3 models with constant cp, no viscosity and no
 * chemistry, every flux, BJ, every boundary type.  The state layout is the
 * octree's, [n][V][4], in this mesh's leaf order. */
int  charm_gpu_upload_gmesh(charm_gpu_ctx_t *c, const charm_gmesh_t *m);
/* each leaf's 8 corners through its tree's trilinear map, [n][8][3] in
 * z-order (corner x + 2 y + 4 z) */
int  charm_gmesh_leaf_corners(const charm_gmesh_t *m, double *corners);
/* VTK output of a multi-tree mesh in charm_write_vtk's format: each leaf a
 * hexahedron of its own 8 corners, cell data treeid, level, mpirank and the
 * cell-centre fields of u (FP64 [n][V][4]) */
int  charm_gmesh_write_vtk(const char *name, const charm_gmesh_t *m, const double *u, const charm_config_t *cfg);
/* A Gmsh task's mesh (files_type gmsh_msh; src/charm_3d.c:32-37): mesh.name
 * (relative to dir unless absolute) read by charm_conn_read_msh, its patches
 * and regions bound to the task's boundary and region names
 * (charm_conn_bind_task), refined uniformly to MIN_LEVEL.  The mesh keeps
 * each tree's region for charm_task_initial_state_gmesh. */
int  charm_task_gmesh(const charm_task_t *task, const char *dir, charm_gmesh_t **out);
/* charm_init_initial_condition (src/charm_init.c:12-73) on such a mesh:
 * each leaf takes the region of its tree; u FP64 [n][V][4] */
int  charm_task_initial_state_gmesh(const charm_task_t *task, const charm_gmesh_t *m, double *u, int64_t n);
/* charm_timesteps (src/charm_timesteps.c:23-51) on a context that runs a
 * multi-tree mesh: dt = TAU (the models' get_dt), t += dt while t < TMAX,
 * the solution written before every step FILE_OUTPUT_STEP divides
 * (<out_prefix>_<step %08d>, charm_gmesh_write_vtk), a log line after every
 * step LOG_OUTPUT_STEP divides.  Adaptation across trees is not built:
 * REFINE_PERIOD is not acted on. */
int  charm_gmesh_timesteps(charm_gpu_ctx_t *g, const charm_gmesh_t *m, const charm_config_t *cfg);
/* ------------------------------------------------ reference entry points
 * (include/charm_models.h:15-23, charm_limiter.h:13-15, charm_fluxes.h:19-32,
 * charm_amr.h:9-10)
 * The forest's user pointer is a charm_ctx_t built by charm_ctx_create. */
typedef struct charm_ctx charm_ctx_t;
typedef struct charm_ghost charm_ghost_t;     /* p4est_ghost_t: single rank, unused */
typedef struct charm_data  charm_data_t;      /* ghost records: unused */

/* the reference's prim record (include/charm_globals.h:59-76) */
typedef struct {
    charm_real_t r, u, v, w, e, e_tot, p, t, cz, gam, cp, cv, m;
    int          mat_id;
    charm_real_t c[128];
} charm_prim_t;

/* the reference's conservative record (include/charm_globals.h:80-87) */
typedef struct {
    charm_real_t ru, rv, rw, re;
    charm_real_t rc[128];
    int          mat_id;
} charm_cons_t;
int  charm_ctx_create(charm_forest_t *f, const charm_config_t *cfg, charm_ctx_t **out);
int  charm_ctx_destroy(charm_ctx_t *c);
charm_gpu_ctx_t *charm_ctx_gpu(charm_ctx_t *c);
/* status of the last failed reference-named call on this context (they
 * return void like the reference's); reading it clears it.  The message is
 * charm_gpu_last_error(). */
int charm_ctx_error(charm_ctx_t *c);

void         charm_model_euler_timestep_single(charm_forest_t *p4est, charm_real_t *dt,
                                               charm_ghost_t **ghost, charm_data_t **ghost_data);
charm_real_t charm_model_euler_get_dt(charm_forest_t *p4est);
void         charm_model_ns_timestep_single(charm_forest_t *p4est, charm_real_t *dt,
                                            charm_ghost_t **ghost, charm_data_t **ghost_data);
charm_real_t charm_model_ns_get_dt(charm_forest_t *p4est);
/* runs steps until t >= TMAX like src/charm_timesteps.c:23-51 */
void charm_timesteps(charm_forest_t *p4est);
void charm_limiter(charm_forest_t *p4est, charm_ghost_t *ghost, charm_data_t *ghost_data);
/* one adaptation of the forest and the state: estimator, flags, refine /
 * coarsen / 2:1 face balance, re-flattening and FP64 projection
 * (include/charm_amr.h:9-10, src/charm_amr.c:575-610).  The time-step entry
 * points call it when REFINE_PERIOD divides a nonzero step number. */
void charm_adapt(charm_forest_t *p4est, charm_ghost_t *ghost, charm_data_t *ghost_data);
/* the reference's initial refinement (include/charm_amr.h:9,
 * src/charm_amr.c:309-347) reads an estimator gradient that is never
 * computed before it (:50-83 against :581-599); this records
 * CHARM_ERR_UNSUPPORTED with that reason.  Build the initial mesh with
 * charm_adapt passes on a projected initial state. */
void charm_adapt_init(charm_forest_t *p4est);
void charm_limiter_bj(charm_forest_t *p4est, charm_ghost_t *ghost, charm_data_t *ghost_data);
/* Pointwise host functions with the reference's names and arguments, over
 * the forest's configuration, FP64 (the device path does not call them; its
 * kernels restate the same formulas):
 *  EOS functions, eos_fn (include/charm_globals.h:122; src/charm_eos.c:
 *  57-148): flag 0 p(r,e), 1 e(r,p), 2 r(T,p), 3 (T,p) -> (r,cz,e),
 *  4 (r,e) -> (p,cz,T); a polynomial cp of component 0 takes T from the
 *  Newton solve for flags 1 and 4 (include/charm_eos.h:14-18) */
void charm_mat_eos_ideal(charm_forest_t *p4est, charm_prim_t *p, int flag);
void charm_mat_eos_mix(charm_forest_t *p4est, charm_prim_t *p, int flag);
/* the table EOS (include/charm_eos.h:18, src/charm_eos.c:150-154), which the
 * reference leaves unreleased and aborts on: records CHARM_ERR_UNSUPPORTED
 * ("TABLE EOS IS NOT RELEASED") and leaves p unchanged */
void charm_mat_eos_table(charm_forest_t *p4est, charm_prim_t *p, int flag);
/* the gas constant 8.31446261815324 (include/charm_eos.h:20, src/charm_eos.c:10) */
charm_real_t charm_eos_get_r(void);
/* BOUND_INLET .. BOUND_MASS_FLOW by name -> CHARM_BND_INLET .. _MASS_FLOW,
 * anything else CHARM_BND_UNKNOWN (include/charm_bnd_cond.h:33,
 * src/charm_bnd_cond.c:9-19) */
int charm_bnd_type_by_name(const char *name);
/* boundary-condition functions, bnd_fn (include/charm_globals.h:237;
 * include/charm_bnd_cond.h:14-33, src/charm_bnd_cond.c:36-146); the
 * mass-flow one records CHARM_ERR_UNSUPPORTED's message and passes the
 * state through */
void charm_bnd_cond_fn_inlet(charm_prim_t *par_in, charm_prim_t *par_out, int8_t face, charm_real_t *param,
                             charm_real_t n[3]);
void charm_bnd_cond_fn_outlet(charm_prim_t *par_in, charm_prim_t *par_out, int8_t face, charm_real_t *param,
                              charm_real_t n[3]);
void charm_bnd_cond_fn_wall_slip(charm_prim_t *par_in, charm_prim_t *par_out, int8_t face, charm_real_t *param,
                                 charm_real_t n[3]);
void charm_bnd_cond_fn_wall_no_slip(charm_prim_t *par_in, charm_prim_t *par_out, int8_t face,
                                    charm_real_t *param, charm_real_t n[3]);
/* mass flow, param = (flux, static p, total T, direction cosines, outflow
 * p, composition): as the reference, a supersonic outflow stores its
 * pressure in param[6] for later calls (the device kernels keep param[6]
 * fixed, since that update depends on the face order); it reads the
 * material of the boundary charm_bnd_cond is evaluating, so a direct call
 * passes the state through and records an error */
void charm_bnd_cond_fn_mass_flow(charm_prim_t *par_in, charm_prim_t *par_out, int8_t face,
                                 charm_real_t *param, charm_real_t n[3]);
/* charm_bnd_cond (src/charm_bnd_cond.c:21-33): the BC function of cube side
 * `face`, then EOS flag 3 and e_tot */
void charm_bnd_cond(charm_forest_t *p4est, int treeid, int8_t face, charm_prim_t *par_in,
                    charm_prim_t *par_out, charm_real_t n[3]);
/* src/charm_globals.c:183-240 */
void charm_param_cons_to_prim(charm_forest_t *p4est, charm_prim_t *p, charm_cons_t *c);
void charm_param_prim_to_cons(charm_forest_t *p4est, charm_cons_t *c, charm_prim_t *p);
void charm_prim_cpy(charm_prim_t *dest, charm_prim_t *src);
/* VTK files of the current state, <prefix>_<timestep %08d>{_0000.vtu,.pvtu}
 * (src/charm_output.c:56-126; the time-step entry points call it when
 * FILE_OUTPUT_STEP divides the step number, src/models/euler/charm_model.c:507-510) */
void charm_write_solution(charm_forest_t *p4est);
void charm_calc_flux_lf(charm_forest_t *p4est, charm_prim_t prim[2], charm_real_t *qu,
                        charm_real_t *qv, charm_real_t *qw, charm_real_t *qe,
                        charm_real_t qc[], charm_real_t n[3]);
void charm_calc_flux_hllc(charm_forest_t *p4est, charm_prim_t prim[2], charm_real_t *qu,
                          charm_real_t *qv, charm_real_t *qw, charm_real_t *qe,
                          charm_real_t qc[], charm_real_t n[3]);
void charm_calc_flux_godunov(charm_forest_t *p4est, charm_prim_t prim[2], charm_real_t *qu,
                             charm_real_t *qv, charm_real_t *qw, charm_real_t *qe,
                             charm_real_t qc[], charm_real_t n[3]);
void charm_calc_flux_cd(charm_forest_t *p4est, charm_prim_t prim[2], charm_real_t *qu,
                        charm_real_t *qv, charm_real_t *qw, charm_real_t *qe,
                        charm_real_t qc[], charm_real_t n[3]);

#ifdef __cplusplus
}
#endif
#endif
