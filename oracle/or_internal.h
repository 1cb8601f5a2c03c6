/*
 * or_internal.h -- private types of the oracle (TEST INFRASTRUCTURE ONLY).
 */
#ifndef OR_INTERNAL_H
#define OR_INTERNAL_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "charm_oracle.h"

#define OR_MAX(X, Y) ((X) > (Y) ? (X) : (Y))
#define OR_MIN(X, Y) ((X) < (Y) ? (X) : (Y))

/* Species bound of the per-leaf record.  OR_FAITHFUL (the bench's
 * reference-faithful CPU timing build, never the checker) sizes the record
 * for the reference's CHARM_MAX_COMPONETS_COUNT = 128 species
 * (include/charm_def.h:181; about the reference's 24 KB per leaf) and runs
 * the field evaluation and the RK combine over all of them, as
 * src/charm_base_func.c:198-211 and src/models/ns/charm_model.c:129-178 do.
 * The public structs keep OR_NC_MAX. */
#ifdef OR_FAITHFUL
#define OR_REC_NC 128
#define OR_FIELD_NV(s) (4 + OR_REC_NC)
#else
#define OR_REC_NC OR_NC_MAX
#define OR_FIELD_NV(s) ((s)->nv)
#endif
#define OR_REC_NV (4 + OR_REC_NC)

extern const double or_gR;

/* per-leaf geometry, the `geom` member of charm_param_t
 * (include/charm_globals.h:361-377) */
typedef struct {
    double n[6][3];
    double face_gp[6][OR_FGP][3];
    double face_gw[6][OR_FGP];
    double face_gj[6][OR_FGP];
    double quad_gp[OR_QGP][3];
    double quad_gw[OR_QGP];
    double quad_gj[OR_QGP];
    double area[6];
    double volume;
    double c[3];
    double fc[6][3];
    double dh[3];
    double a[4][4];
    double a_inv[4][4];
} or_geom_t;

/* per-leaf record in the shape of charm_data_t (include/charm_globals.h:
 * 143-234); coefficients stored [v][k] with v = ru, rv, rw, re, rc[...] */
typedef struct {
    double c[OR_REC_NV][OR_BF];
    double c_old[OR_REC_NV][OR_BF];
    or_geom_t g;
    int    mat_id;
    double grav[3];
    struct {
        int    count;
        double u[OR_REC_NV][7];
    } l;
    double grad_u[3];
    double integ[OR_REC_NV][OR_BF];          /* int_ru .. int_rc */
    /* Navier-Stokes auxiliaries (charm_param_t.model.ns) */
    double tau[6][OR_BF];                    /* xx yy zz xy xz yz */
    double q[3][OR_BF];
    double d[OR_NC_MAX];
    double chem_rhs;
    double int_tau[6][OR_BF];
    double int_q[3][OR_BF];
    /* BR2 option: the auxiliary field with every interior face term taken
     * from the leaf's own trace (tau_h less the interior face liftings) */
    double tau_own[6][OR_BF];
    double q_own[3][OR_BF];
    double int_tau_own[6][OR_BF];
    double int_q_own[3][OR_BF];
    int32_t quad[4];                         /* x y z level */
    int32_t tree;                            /* multi-tree forests: the leaf's tree */
} or_data_t;

/* one face visit in the sense of p4est_iter_face_info_t */
typedef struct {
    int nsides;
    int face[2];
    int is_hanging[2];
    int quad[2][4];
    int bnd;          /* domain face for a one-sided face */
    int wrap_axis;    /* periodic wrap axis, -1 if none */
} or_face_t;

struct or_sim {
    or_config_t cfg;
    int64_t     n;
    or_data_t  *d;
    int64_t     nf;
    or_face_t  *f;
    int         nv;
    double      dt;          /* ctx->dt: the chemistry's step */
    char        err[256];
    /* OpenMP builds only (the bench's CPU timing): each leaf's face
     * incidences in face order, and per face its flux values at the face
     * points, so that the convective pass runs its faces in parallel and
     * still adds every leaf's contributions in the serial order */
    int64_t    *inc_ptr;     /* [n + 1] */
    int64_t    *inc;         /* face << 4 | side << 3 | subface (7: all four) */
    int64_t    *fq_off;      /* [nf + 1], in doubles */
    double     *fq;
    double     *lq;          /* [nf][2][OR_NV_MAX]: the limiter's per-face neighbour values */
    double     *tv;          /* multi-tree forests: [n_trees][8][3] tree vertices (NULL: the unit cube) */
};

/* the 8 vertices of leaf d (its tree's trilinear map) */
void or_leaf_vertices(const or_sim_t *s, const or_data_t *d, double v[8][3]);

double or_scalar_prod(const double a[3], const double b[3]);
double or_vect_length(const double v[3]);
void   or_vect_prod(const double a[3], const double b[3], double r[3]);
double or_matr3_det(double a[3][3]);
void   or_matr_vect_mult(double a[4][4], const double b[4], double r[4]);
double or_comp_calc_cp(const or_comp_t *c, double t);
double or_comp_calc_cp_dt(const or_comp_t *c, double t);
double or_comp_calc_ml(const or_comp_t *c, double t);
double or_comp_calc_kp(const or_comp_t *c, double t);
double or_comp_calc_enthalpy(const or_comp_t *c, double t, double t_ref);
void   or_prim_to_cons(const or_config_t *cfg, or_cons_t *c, const or_prim_t *p);
int    or_flux_err(const or_config_t *cfg, int flux, const or_prim_t prim[2],
                   const double n[3], double *q);
void   or_quad_vertices(const int32_t q[4], double v[8][3]);
extern const double or_unit_tree[8][3];
void   or_quad_vertices_tv(const double tv[8][3], const int32_t q[4], double v[8][3]);
void   or_geom_calc_tv(const double tv[8][3], const int32_t q[4], or_geom_t *g);
double or_bf(const or_geom_t *g, const double x[3], int k);
double or_bf_dx(const or_geom_t *g, int k);
double or_bf_dy(const or_geom_t *g, int k);
double or_bf_dz(const or_geom_t *g, int k);
void   or_geom_calc(const int32_t q[4], or_geom_t *g);
void   or_get_fields(const or_sim_t *s, const or_data_t *d, const double x[3], or_cons_t *c);
void   or_get_fields_avg(const or_sim_t *s, const or_data_t *d, or_cons_t *c);
void   or_face_normal(const or_data_t *d, int face, double n[3]);
void   or_shift(const or_face_t *f, int side, int owner, const double x[3], double xs[3]);

#endif
