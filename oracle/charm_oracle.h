/*
 * charm_oracle.h -- FP64 CPU restatement of the CHARM_3D (zhrv/charm_3d)
 * discontinuous-Galerkin residual and time-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: only tests/, the
 * smoke check in __graft_entry__.py and the cpu_baseline leg of bench.py
 * load it.  The product library (charm_3d_amd/) never links, loads or calls
 * anything in this directory.
 *
 * The oracle works on its own data structures: an array-of-structs record per
 * leaf (the shape of the reference's charm_data_t, include/charm_globals.h:
 * 143-234) and its own face traversal (a hash-map neighbour search that
 * emits each face once with the side/hanging semantics of p4est_iterate as
 * consumed by src/models/euler/charm_model.c:140-385).  It shares no code
 * with the product's forest, flattening or kernels.
 *
 * Parity status of each part is stated in oracle/README.md.
 */
#ifndef CHARM_ORACLE_H
#define CHARM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_BF       4      /* CHARM_BASE_FN_COUNT (include/charm_def.h:177) */
#define OR_FGP      6      /* CHARM_FACE_GP_COUNT (include/charm_def.h:178) */
#define OR_QGP      8      /* CHARM_QUAD_GP_COUNT (include/charm_def.h:179) */
#define OR_NC_MAX   8      /* species bound of this restatement (reference: 128) */
#define OR_NV_MAX   (4 + OR_NC_MAX)
#define OR_EPS      1.e-12 /* CHARM_EPS (include/charm_def.h:158) */
#define OR_RIM_EPS  1.e-5  /* CHARM_RIM_EPS (include/charm_def.h:156) */
#define OR_RIM_NEWTON_STEPS 5000 /* include/charm_def.h:155 */
#define OR_MAXLEVEL 19     /* p8est P8EST_MAXLEVEL: root length 2^19 */
#define OR_ROOT_LEN (1 << OR_MAXLEVEL)

/* flux ids, in the order of the reference's FLUX_TYPE keys
 * (src/charm_init_yaml.cpp:352-379) */
enum { OR_FLUX_LF = 0, OR_FLUX_HLLC = 1, OR_FLUX_GODUNOV = 2, OR_FLUX_CD = 3 };
/* boundary types, include/charm_globals.h:432-439 */
enum { OR_BND_INLET = 0, OR_BND_OUTLET, OR_BND_WALL_SLIP, OR_BND_WALL_NO_SLIP,
       OR_BND_MASS_FLOW, OR_BND_PERIODIC = 100 };
/* charm_comp_cp_type_t / _kp_ / _ml_ (include/charm_globals.h:20-39) */
enum { OR_CP_CONST = 0, OR_CP_POLYNOM = 1 };
enum { OR_ML_CONST = 0, OR_ML_SATHERLAND = 1, OR_ML_NV = 2 };
enum { OR_EOS_IDEAL = 0, OR_EOS_MIX = 1 };

typedef struct {
    double r, u, v, w, e, e_tot, p, t, cz, gam, cp, cv, m;
    int    mat_id;
    double c[OR_NC_MAX];
} or_prim_t;

typedef struct {
    double ru, rv, rw, re;
    double rc[OR_NC_MAX];
    int    mat_id;
} or_cons_t;

typedef struct {
    double m, ml0, kp0, t0, ts, sig, ek, h0;
    int    cp_type, kp_type, ml_type;
    int    n_cp;
    double cp[8];
} or_comp_t;

typedef struct {
    int    left[3], right[3];
    double a, e, n;
} or_reaction_t;

typedef struct {
    int    n_comp;
    or_comp_t comp[OR_NC_MAX];
    int    eos;                /* OR_EOS_IDEAL / OR_EOS_MIX */
    int    flux;               /* OR_FLUX_* */
    int    limiter;            /* 0 none, 1 Barth-Jespersen */
    int    model;              /* 0 Euler forward step, 1 NS SSP-RK3 */
    int    bnd_type[6];        /* per domain face: OR_BND_* */
    double bnd_param[6][16];
    double grav[3];
    int    use_visc, use_diff;
    double t_ref;
    int    n_react;
    or_reaction_t react[4];
    double cfl;
    int    chem_parity;        /* 1: the reference's chem_w[species] index (see or_ns.c) */
    int    chem_energy_molar;  /* 1: energy source sum w_i M_i h_i instead of the
                                  reference's sum w_i h_i (see or_ns.c) */
    double visc_br2_eta;       /* 0: the reference's BR1 viscous flux; > 0: BR2
                                  lifting with this penalty factor (see or_ns.c) */
} or_config_t;

typedef struct or_sim or_sim_t;

/* lifecycle: leaves given as int32 [n][4] = (x, y, z, level) in root units */
or_sim_t *or_create(const or_config_t *cfg, int64_t n_leaves, const int32_t *leaves);
/* multi-tree forest refined uniformly to `level`: tree_vertices [n_trees][8][3]
 * (z-order corners), tree_bnd [n_trees][6] = the boundary slot of a tree face
 * or -1 where it meets another tree (conforming meshes) */
or_sim_t *or_create_trees(const or_config_t *cfg, int64_t n_trees, const double *tree_vertices,
                          const int32_t *tree_bnd, int level);
void      or_destroy(or_sim_t *s);
int       or_nvar(const or_sim_t *s);
int64_t   or_n_faces(const or_sim_t *s);
const char *or_error(const or_sim_t *s);

/* state: double [n][V][4] */
void or_set_state(or_sim_t *s, const double *u);
void or_get_state(const or_sim_t *s, double *u);

/* the residual integrals int_* after zero + convective pass: [n][V][4] */
void or_conv_residual(or_sim_t *s, double *out);
void or_limiter(or_sim_t *s);
void or_euler_step(or_sim_t *s, double dt);
void or_ns_step(or_sim_t *s, double dt);
void or_run(or_sim_t *s, double dt, int nsteps);

/* Navier-Stokes pieces for tests: full stage residual (zero + viscous +
 * convective), auxiliary tau[6][4] q[3][4] per leaf ([n][9][4]), diffusion
 * coefficients [n][n_comp], chemistry applied to the state (chem_rhs [n]) */
void or_ns_residual(or_sim_t *s, double *out);
void or_ns_aux(or_sim_t *s, double *out);
void or_ns_aux_own(or_sim_t *s, double *out);
void or_ns_visc(or_sim_t *s, double *out);
void or_ns_diff_coeffs(or_sim_t *s, double *out);
void or_ns_chem_apply(or_sim_t *s, double dt, double *chem_rhs);

/* face listing for flattening checks: per face
 * [nsides, face0, face1, hang0, hang1, q0[4], q1[4], bnd, wrap_axis] = 16 int32 */
void or_faces(const or_sim_t *s, int32_t *out);
/* geometry of one leaf: n[6][3] fc[6][3] face_gp[6][6][3] face_gw[6][6]
 * face_gj[6][6] quad_gp[8][3] quad_gw[8] quad_gj[8] area[6] volume c[3] dh[3]
 * a[4][4] a_inv[4][4] packed in that order (580 doubles) */
void or_geom(const or_sim_t *s, int64_t leaf, double *out);

/* diagnostics: FP64 conservation sums sum_e V_e U[e][v][0] (V values) and the
 * CFL dt of the reference's disabled branch (mode 0) or h-based (mode 1) */
void   or_conservation_sums(const or_sim_t *s, double *out);
double or_cfl_dt(const or_sim_t *s, int mode);

/* adaptation: estimator (err2, centre density per leaf), refine (+1) /
 * coarsen (-1) flags, and the L2 projection onto a new leaf set given the
 * new->old map (>= 0 source leaf, -1 - first child of a coarsened family) */
void or_amr_indicator(or_sim_t *s, double *err2, double *rho_c);
void or_amr_flags(const or_sim_t *s, const double *err2, const double *rho_c, double max_err,
                  int min_level, int max_level, int8_t *flags);
void or_amr_project(int nv, int64_t n_old, const int32_t *old_leaves, const double *u_old,
                    int64_t n_new, const int32_t *new_leaves, const int64_t *map, double *u_new);

/* pointwise numerics (batched for tests) */
void or_cons_to_prim(const or_config_t *cfg, const or_cons_t *c, or_prim_t *p);
void or_eos(const or_config_t *cfg, or_prim_t *p, int flag);
void or_flux(const or_config_t *cfg, int flux, const or_prim_t prim[2],
             const double n[3], double *q /* 4 + n_comp */);
void or_bnd(const or_config_t *cfg, int face, const or_prim_t *in,
            or_prim_t *out, const double n[3]);
void or_matr_inv(const double a[4][4], double am[4][4]);
void or_rim(double *RI, double *EI, double *PI, double *UI, double *VI, double *WI,
            double RB, double PB, double UB, double VB, double WB,
            double RE, double PE, double UE, double VE, double WE,
            double gam_b, double gam_e);
double or_gas_r(void);
size_t or_prim_size(void);
size_t or_cons_size(void);
size_t or_config_size(void);

#ifdef __cplusplus
}
#endif
#endif
