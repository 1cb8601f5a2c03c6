// charm_kernels.h -- launch interface between the context code and the
// kernels (C++ linkage, internal to libcharm3d).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "charm_device.hpp"

namespace charm {

struct StageArgs {
    const void *U;            // state the residual is evaluated on
    const void *U0;           // stage-start state for the SSP-RK combine
    void *out;                // residual (mode 0) or updated state (mode 1)
    const uint8_t *level;
    const int4 *cf4;          // chunk face records (e0, e1, info, 0)
    const int *ch_ebeg, *ch_fbeg, *slot_ptr;
    const uint16_t *slots;
    int *err;
    int64_t n_elem;
    int n_chunks, max_faces, chunk_elems;
    int n_xcd;                // XCDs the workgroups are spread over
    int mode;                 // 0: write int_*; 1: c - dt A^-1 int
    int combine;              // 0: none, 1: w/4 + 3 u0/4, 2: 2w/3 + u0/3
    double dt;
    const void *visc;         // [16][n] viscous integrals added before A^-1 (or null)
    // the point EOS closures of U (NsArgs::tpt) for the PH = 1 variant, or
    // null (each point evaluates its own)
    const void *tpt;
    // the context's constants in device memory: the Navier-Stokes / real-gas
    // variant (PH = 1) reads them here, so that its runtime indices (power
    // series up to n_cp, the BC record of a lane's face) are loads rather
    // than a private copy of the whole by-value block
    const DevParams *prm;
};

struct LimArgs {
    const void *W, *U0;
    void *out;
    const int *nbr;
    const int4 *aux;
    int64_t n;
    int limit;                // apply Barth-Jespersen
    int n_xcd;
    int combine;              // as StageArgs::combine
};

struct RedArgs {
    const void *U;
    const uint8_t *level;
    double *partial, *result;
    int64_t n;
    int mode;                 // dt length: 0 volume, 1 edge
};

struct FluxArgs {
    const double *prim_l, *prim_r, *normals;
    double *out;
    int64_t n;
    // vector-Jacobian product: qbar [n][V] in, d(qbar . q)/d(prim rows) out
    const double *qbar;
    double *bar_l, *bar_r;
};

// density-gradient error estimator of the adaptation (one thread per leaf)
struct AmrArgs {
    const void *U;
    const uint8_t *level;
    const int *nbr;
    const int4 *aux;
    double *err2, *rho_c;     // [n] each
    int64_t n;
};

// projection of the state onto a new leaf set: map[i] >= 0 names the old
// leaf that contains new leaf i, map[i] = -1 - f names the coarsened family
// whose first member is old leaf f (new leaf i is that parent or inside it)
struct ProjArgs {
    const void *Uold;
    void *Unew;
    const int4 *old_q, *new_q;   // (x, y, z, level) in root units
    const int64_t *map;
    int64_t n_old, n_new;
    int nvk;                     // 4 V coefficients per leaf
};

// Multi-tree (general hexahedral) meshes, charm_gmesh.hip: the stage as a
// face pass (FQ[2][4V][nf]) and an element pass; the limiter.  Geometry
// arrays in the state precision, layouts of charm_internal.h (CHARM_GM_*).
struct GArgs {
    const void *U;            // stage: the state the residual is evaluated on; limiter: W
    const void *U0;           // SSP-RK combine base
    void *out;                // mode 0: the residual integrals; else the new state
    void *fq;                 // [6][4V][n] each leaf's face sums by local face (side 1 negated)
    void *mb;                 // [V][n] cell means of the stage output for the limiter (or null)
    const void *lg, *fg;      // [CHARM_GM_LEAFQ][n], [CHARM_GM_FACEQ][nf]
    const double *fn;         // [3][nf] side-0 outward normals (FP64: the reference's face frame)
    const int *fe, *ff, *lnb;
    int *err;
    int64_t n, nf;
    int mode, combine;
    int n_xcd;                // XCDs: the face pass and limiter give each a contiguous range
    double dt;
};
hipError_t launch_gstage(int prec, int nc, int flux, const GArgs &A, const DevParams &P, hipStream_t st);
hipError_t launch_glimiter(int prec, int nc, const GArgs &A, const DevParams &P, hipStream_t st);

// Navier-Stokes kernels (one thread per leaf).  BR2 liftings: per leaf 24
// (sub)face slots (direction d * 4 + sub-face j) of kLiftFields coefficients
// (tau xx yy zz xy xz yz and q along the face axis, 4 each)
constexpr int kLiftSlots = 24;
constexpr int kLiftFields = 28;

struct NsArgs {
    const void *U;            // state
    void *out;                // chemistry: state with the reacted species means
    const uint8_t *level;
    const int *nbr;
    const int4 *aux;
    double *chem_rhs;         // [n] energy source of the chemistry (or null)
    double *dcoef;            // [n_comp][n] diffusion coefficients (or null)
    const void *aux_in;       // [36][n] tau, q coefficients (viscous integrals)
    void *aux_out;            // [36][n] (gradient pass)
    void *visc_out;           // [16][n] viscous integrals of ru rv rw re
    void *aux_own;            // [36][n] BR2: own-trace tau, q (written by code 4, read by code 5)
    void *lift;               // [24][kLiftFields][n] BR2: each (sub)face slot's lifting (codes 4, 5)
    // point EOS closures and transport of the state U, in its precision
    // (code 6 writes them; the viscous kernels and the stage read them where
    // non-null): planes m, cv, gamma, mu, k of kMemoRows rows, row
    // (4 d + j) * 6 + k at point k of (sub)face slot 4 d + j of each leaf's
    // own face points, row kMemoVolRow + g at volume point g:
    // [kMemoPlanes][kMemoRows][n]
    void *tpt;
    int64_t n;
    double dt;
    const DevParams *prm;     // the context's constants in device memory (read by every NS kernel)
};

// adjoint of one SSP-RK stage (one thread per leaf):
// OUT = c_out (sb + (d int / d U)^T W) + c_acc ACC + c_acc2 ACC2,  sb = c_in X,
// W = -dt A^-1 sb
struct AdjArgs {
    const void *U, *X, *ACC, *ACC2;
    void *OUT;
    double c_in, c_out, c_acc, c_acc2, dt;
    const uint8_t *level;
    const int *nbr;
    const int4 *aux;
    double *theta_bar;        // [n][3] += d/d(inlet u, T, p) (or null)
    int theta_face;           // the inlet face of the parameters (-1: none)
    int64_t n;
    const int *bnd_elems;     // leaves with at least one boundary face
    const int *bnd_ptr;       // [n_bnd + 1]: their ranges in bnd_faces
    const int *bnd_faces;     // boundary faces, leaf * 8 + direction, by leaf
    int n_bnd, n_bface;
    void *bnd_part;           // [n_bface][6][4 V] point contributions (state type)
    double *bnd_theta;        // [n_bface][6][3] parameter contributions
    // chunked face-centric variant (chunks as for the stage kernel; null
    // cf4: the element-centric kernel)
    const int4 *cf4;
    const int *ch_ebeg, *ch_fbeg, *slot_ptr;
    const uint16_t *slots;
    int n_chunks, max_faces, chunk_elems, n_xcd;
    // chunked variant: per leaf its index in bnd_elems (or -1); the chunk
    // kernel then adds the boundary parts itself (k_adj_bnd_pts runs first)
    const int *bnd_of;
};

// objective over a box of leaves: mode 0 per-block partial sums of
// sum_gp wJ p, mode 1 lambda += coef dQ/dU
struct ObjArgs {
    const void *U;
    void *lambda;
    const int4 *q;
    double lo[3], hi[3];
    double coef;
    double *partial;
    int mode;
    int64_t n;
};

hipError_t launch_adj_stage(int prec, int nc, int flux, const AdjArgs &A, const DevParams &P, hipStream_t st);
// LDS bytes per workgroup of the chunked adjoint stage kernel
size_t adj_chunk_smem_bytes(int prec, int nc, int max_faces);
hipError_t launch_adj_objective(int prec, int nc, const ObjArgs &A, const DevParams &P, hipStream_t st);

hipError_t launch_stage_f32_ph0(int nc, int flux, const StageArgs &A, const DevParams &P, hipStream_t st);
hipError_t launch_stage_f32_ph1(int nc, int flux, const StageArgs &A, const DevParams &P, hipStream_t st);
hipError_t launch_stage_f64_ph0(int nc, int flux, const StageArgs &A, const DevParams &P, hipStream_t st);
hipError_t launch_stage_f64_ph1(int nc, int flux, const StageArgs &A, const DevParams &P, hipStream_t st);
hipError_t launch_stage_mix_ph0(int nc, int flux, const StageArgs &A, const DevParams &P, hipStream_t st);
const void *stage_fn_f32_ph0(int nc, int flux, int bs);
const void *stage_fn_f32_ph1(int nc, int flux, int bs);
const void *stage_fn_f64_ph0(int nc, int flux, int bs);
const void *stage_fn_f64_ph1(int nc, int flux, int bs);
const void *stage_fn_mix_ph0(int nc, int flux, int bs);
// prec: 0 FP64, 1 FP32, 2 mixed (FP64 state, FP32 residual)
hipError_t launch_stage(int prec, int nc, int flux, int ph, const StageArgs &A, const DevParams &P, hipStream_t st);
size_t stage_smem_bytes(int prec, int nc, int max_faces);
// host address of the stage kernel a launch with these parameters uses
// (bs = the workgroup size: 128 or 256)
const void *stage_kernel(int prec, int nc, int flux, int ph, int bs);
hipError_t launch_limiter(int prec, int nc, int ph, const LimArgs &A, const DevParams &P, hipStream_t st);
hipError_t launch_reduce(int prec, int nc, int ph, const RedArgs &A, const DevParams &P, hipStream_t st);
// as launch_reduce, writing the first nout of the V sums and the CFL minimum
hipError_t launch_reduce_n(int prec, int nc, int ph, const RedArgs &A, const DevParams &P, hipStream_t st,
                           int nout);
// device address and size of the static block-partial array of the
// allocation-free conservation sums, for the current device (no allocation)
hipError_t noalloc_scratch(void **ptr, size_t *bytes);
hipError_t launch_flux_points(int prec, int nc, int flux, const FluxArgs &A, const DevParams &P, hipStream_t st);
// FP64 pointwise flux VJP (LF, HLLC, CD; 1-2 species) on dual numbers
hipError_t launch_flux_points_vjp(int nc, int flux, const FluxArgs &A, const DevParams &P, hipStream_t st);
hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st);
// what: 0 chemistry, 1 diffusion coefficients, 2 / 4 tau / q (BR1 / BR2),
// 3 / 5 viscous integrals (BR1 / BR2), 6 point EOS closures
hipError_t launch_ns(int prec, int what, int nc, const NsArgs &A, const DevParams &P, hipStream_t st);
hipError_t launch_amr_indicator(int prec, int nc, int ph, const AmrArgs &A, const DevParams &P, hipStream_t st);
hipError_t launch_amr_project(int prec, const ProjArgs &A, hipStream_t st);
hipError_t launch_convert(int prec, int to_f64, const void *src, void *dst, int64_t n, int nvk, hipStream_t st);
// HLLC branch comparison of two FP64 states on one chunked octree mesh
constexpr int kHllcBits = 9;
struct FlipArgs {
    const double *Ua, *Ub;    // the two states, [(4 v + k) n + e]
    const int4 *cf4;
    const int *ch_ebeg, *ch_fbeg;
    int64_t n;
    int n_chunks;
    unsigned long long *out;  // [2 + kHllcBits]: interior face points, points whose branch
                              // differs, points where each branch bit differs
};
hipError_t launch_hllc_flips(int nc, const FlipArgs &A, const DevParams &P, hipStream_t st);

} // namespace charm
