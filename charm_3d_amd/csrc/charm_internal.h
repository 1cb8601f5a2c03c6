/*
 * charm_internal.h -- private host types of libcharm3d (C99 side).
 */
#ifndef CHARM_INTERNAL_H
#define CHARM_INTERNAL_H

#include <stdint.h>
#include "charm_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CHARM_MAXLEVEL 19                 /* root length 2^19 as p8est */
#define CHARM_ROOT_LEN (1 << CHARM_MAXLEVEL)

/* face info word: dir0 (3 bits) | kind << 3 (2 bits) | subface << 5 (2 bits) */
enum { CHARM_FACE_CONF = 0, CHARM_FACE_HANG0 = 1, CHARM_FACE_HANG1 = 2, CHARM_FACE_BND = 3 };
#define CHARM_FACE_INFO(dir, kind, sub) ((dir) | ((kind) << 3) | ((sub) << 5))
#define CHARM_NBR_AUX_BASE 8              /* nbr <= -8: hanging group -(8 + g) */

struct charm_forest {
    int64_t   n;
    int32_t  *q;            /* [n][4] x y z level, Morton order */
    uint64_t *key;          /* Morton key of each anchor */
    int       periodic[3];
    int       balanced;     /* known 2:1 face balanced (uniform, or balanced since) */
    void     *user;
};

struct charm_mesh {
    int64_t   n;
    uint8_t  *level;        /* [n] */
    int32_t  *q;            /* [n][4] */
    int       periodic[3];
    /* all (sub)faces, each once, side 0 = lower-coordinate element */
    int64_t   nf;
    int32_t  *f4;           /* [nf][4] e0 e1 info 0 */
    /* limiter neighbour table */
    int32_t  *nbr;          /* [n][6] */
    int64_t   naux;
    int32_t  *aux;          /* [naux][4] */
    /* workgroup chunks */
    int       chunk_elems, chunk_faces;
    int64_t   nch;
    int32_t  *ch_ebeg;      /* [nch+1] */
    int32_t  *ch_fbeg;      /* [nch+1] */
    int64_t   ncf;
    int32_t  *cf4;          /* [ncf][4] chunk-local copies of face records */
    int32_t  *slot_ptr;     /* [n+1] */
    int64_t   nslot;
    uint16_t *slots;        /* local face index << 1 | side */
};

/* the forest's user pointer (the reference's charm_ctx_t) */
struct charm_ctx {
    charm_config_t   cfg;
    charm_forest_t  *forest;
    charm_gpu_ctx_t *gpu;
    int              chunk_elems, chunk_faces;   /* flattening of every (re)upload */
    double           t;          /* ctx->t */
    int              timestep;   /* ctx->timestep */
    int              err;
};

charm_ctx_t *charm_ctx_of(charm_forest_t *f);
/* host pointwise EOS pieces (charm_pointwise.c) */
double charm_comp_cp(const charm_comp_desc_t *q, double t);
double charm_comp_cp_dt(const charm_comp_desc_t *q, double t);
double charm_calc_t(const charm_config_t *cfg, const double *c, double e);
void   charm_eos_ideal_cfg(const charm_config_t *cfg, charm_prim_t *p, int flag);
void   charm_eos_mix_cfg(const charm_config_t *cfg, charm_prim_t *p, int flag);
uint64_t charm_morton(int32_t x, int32_t y, int32_t z);
int64_t  charm_forest_find(const charm_forest_t *f, int32_t x, int32_t y, int32_t z);
void     charm_set_error(const char *fmt, ...);
/* FP64 Gauss-Jordan inverse with partial pivoting (charm_geom.c); 0 on a zero pivot */
int      charm_gj_inverse4(const double src[4][4], double inv[4][4]);

/* Multi-tree meshes (charm_gmesh.c): the device arrays, structure of arrays
 * [q][n] / [q][nf].  Per leaf (CHARM_GM_LEAFQ): 1/dh (0-2), basis values at the
 * 8 volume points (3-26), their weights wJ (27-34), A^-1 row-major (35-50),
 * basis values at the 8 vertices (51-74), the leaf means of the 3 linear
 * basis functions (75-77).  Per face (CHARM_GM_FACEQ): basis values of side 0
 * at the 6 points (0-17); side 1's as the affine image of side 0's, b1_a =
 * b0_a r_a + o_a with r_a = dh0_a / dh1_a (18-20) and o_a = (c0_a - c1_a) /
 * dh1_a (21-23), zero on a boundary face; the points' weights wJ (24-29). */
#define CHARM_GM_LEAFQ 78
#define CHARM_GM_FACEQ 30
typedef struct {
    int64_t n, nf;
    const double *lg, *fg, *fn;   /* fn: [3][nf] side-0 outward normals */
    const int32_t *fe;            /* [2][nf] side-0 leaf, side-1 leaf or -1 - boundary slot */
    const int32_t *ff;            /* [2][nf] each side's local face 0-5 (side 1: -1 on a boundary) */
    const int32_t *lnb;           /* [6][n] neighbour leaf or -1 - boundary face index */
} charm_gmesh_arrays_t;
int charm_gmesh_arrays(const charm_gmesh_t *m, charm_gmesh_arrays_t *a);
/* a task's region index per tree (charm_task_gmesh), or NULL */
int charm_gmesh_set_regions(charm_gmesh_t *m, const int32_t *tree_reg);
const int32_t *charm_gmesh_regions(const charm_gmesh_t *m);

#ifdef __cplusplus
}
#endif
#endif
