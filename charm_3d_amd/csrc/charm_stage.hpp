// charm_stage.hpp -- the stage kernel of the DG residual / time-step path
// (included by the translation units that instantiate it per precision and
// physics variant, so that they compile in parallel).
//
// State layout (structure of arrays): U[(4 v + k) n + e] for element e of n,
// variable v = ru rv rw re rc[..], Taylor coefficient k (1, (x-xc)/h,
// (y-yc)/h, (z-zc)/h): a wave's per-element accesses are unit-stride.
//
// One RK stage = k_stage over workgroup chunks (at most one element per
// thread of the 128- or 256-thread workgroup, Morton-consecutive, and a cap
// on the chunk's faces that sizes its LDS):
//   phase 1  one thread per (sub)face touching the chunk: both traces at the
//            six face Gauss points, the Riemann flux, and the three flux
//            moments per variable (sum q wJ, sum q phi_t1 wJ, sum q phi_t2 wJ in
//            the small side's face frame) written to LDS;
//   phase 2  one thread per element: volume integral at the 8 Gauss points,
//            gather of its face moments from LDS (no atomics: a fixed slot
//            list per element), A^-1 (diagonal on the cube octree) and the
//            SSP-RK update / combine, one write of the new state.
// Faces shared by two chunks are evaluated by both (each keeps its side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include "charm_device.hpp"
#include "charm_dev.hpp"
#include "charm_kernels.h"

namespace charm {

// SSP-RK combines rk_1 (c = c/4 + 3 c_old/4) and rk_2 (c = 2c/3 + c_old/3)
// in the reference's operation order, src/models/ns/charm_model.c:136-184
template <typename R>
__device__ __forceinline__ R rk_combine(int mode, R w, R u0)
{
    if (mode == 1) { R c = w * R(0.25); return c + R(0.75) * u0; }
    R c = w * R(2);
    c /= R(3);
    return c + u0 / R(3);
}

// per-(v, k) array bases are wave-uniform (scalar registers); the element
// index is the 32-bit per-lane offset
template <typename R, int NV>
__device__ __forceinline__ void load_elem(const R *__restrict__ U, int64_t n, int e, R *u)
{
    #pragma unroll
    for (int i = 0; i < 4 * NV; i++) u[i] = ldp<R>(U, n, i, e);
}

// load a state stored in S into the arithmetic type R
template <typename S, typename R, int NV>
__device__ __forceinline__ void load_conv(const S *__restrict__ U, int64_t n, int e, R *u)
{
    #pragma unroll
    for (int i = 0; i < 4 * NV; i++) u[i] = (R) ldp<S>(U, n, i, e);
}

template <typename R, int NV>
__device__ __forceinline__ void trace(const R *u, R px, R py, R pz, R *t)
{
    #pragma unroll
    for (int v = 0; v < NV; v++)
        t[v] = u[4*v] + px * u[4*v+1] + py * u[4*v+2] + pz * u[4*v+3];
}

// spread (normal, t1, t2) components onto (x, y, z)
template <typename R>
__device__ __forceinline__ void to_xyz(int a, R n, R t1, R t2, R &x, R &y, R &z)
{
    x = (a == 0) ? n : t1;
    y = (a == 1) ? n : ((a == 0) ? t1 : t2);
    z = (a == 2) ? n : t2;
}

// Moments of one face normal to axis AX from the two element states u0, u1
// (u1 unused on a boundary face):
// out[3][NV] = wJ * (sum q, sum q phi_t1, sum q phi_t2) over the six points,
// phi_t the tangential Taylor coordinates of the small (or side-0) face.
// DEV: u0, u1 hold deviations from the chunk reference *rf (coefficient 0
// minus the reference) and q is F - F(U_r) (charm_dev.hpp).
// tl / tr (PH = 1, may be null): the EOS closure of side 0 / 1 at point k
// from the point pass starts at tl[k * ts] / tr[k * ts], its planes
// kMemoRows * ts apart (global memory: the point loop is not unrolled, so a
// local array would sit in scratch)
template <typename R, int NC, int FLUX, int AX, bool BND, int PH, bool DEV = false>
__device__ __forceinline__ bool face_core(const DevParams &P, const R *u0, const R *u1, int info,
                                          int lown, R *out, int bnd_axis, const DevRef &rf,
                                          const R *tl = nullptr, const R *tr = nullptr,
                                          int64_t ts = 0)
{
    constexpr int NV = 4 + NC;
    constexpr int T1 = (AX == 0) ? 1 : 0, T2 = (AX == 2) ? 1 : 2;
    const int dir0 = info & 7, kind = (info >> 3) & 3, sub = (info >> 5) & 3;
    const R s = (dir0 & 1) ? R(1) : R(-1);              // outward normal of side 0
    const R h = ldexp(R(1), -lown);
    const R wj = h * h * R(1.0 / 6.0);
    // normal-basis values on the face: side 0 at +-1/2 by its face parity
    const R pa0 = R(0.5) * s, pa1 = -pa0;
    // the big side of a hanging face sees the small face's points at
    // delta + offset / 2 in its own units
    const R d1 = (sub & 1) ? R(0.25) : R(-0.25);
    const R d2 = (sub & 2) ? R(0.25) : R(-0.25);
    const bool big0 = (kind == FACE_HANG1), big1 = (kind == FACE_HANG0);
    const R m0 = big0 ? R(0.5) : R(1), m1 = big1 ? R(0.5) : R(1);
    const R o01 = big0 ? d1 : R(0), o02 = big0 ? d2 : R(0);
    const R o11 = big1 ? d1 : R(0), o12 = big1 ? d2 : R(0);
    R S0[NV], S1[NV], S2[NV];
    #pragma unroll
    for (int v = 0; v < NV; v++) { S0[v] = R(0); S1[v] = R(0); S2[v] = R(0); }
    // the normal-direction part of each trace is the same at all six points
    R c0[NV], c1[NV];
    #pragma unroll
    for (int v = 0; v < NV; v++) {
        c0[v] = u0[4*v] + pa0 * u0[4*v+1+AX];
        c1[v] = BND ? R(0) : u1[4*v] + pa1 * u1[4*v+1+AX];
    }
    bool ok = true;
    // a rolled loop over the points (k is wave-uniform: the point's constants
    // are scalar selects): a fifth of the unrolled code and fewer registers
    #pragma unroll 1
    for (int k = 0; k < 6; k++) {
        const R g1 = gp_t1<R>(k), g2 = gp_t2<R>(k);
        R t[NV];
        #pragma unroll
        for (int v = 0; v < NV; v++)
            t[v] = c0[v] + (o01 + m0 * g1) * u0[4*v+1+T1] + (o02 + m0 * g2) * u0[4*v+1+T2];
        R q[NV], qg[NV];
        if constexpr (DEV) {
            DPrim dl, dr;
            dev_prim(rf, t, dl);
            if (BND) {
                if (P.bnd_type[dir0] == BND_OUTLET) {
                    dr = dl;                   // transmissive: the exterior state is the interior one
                } else {
                    Prim<float, 1> pl, pr;
                    prim_of_dev(P, rf, dl, pl);
                    bnd_state<float, 1, 0>(P, dir0, AX, s, pl, pr, bnd_axis);
                    dev_from_prim(rf, pr, dr);
                }
            } else {
                #pragma unroll
                for (int v = 0; v < NV; v++)
                    t[v] = c1[v] + (o11 + m1 * g1) * u1[4*v+1+T1] + (o12 + m1 * g2) * u1[4*v+1+T2];
                dev_prim(rf, t, dr);
            }
            ok &= flux_dev<FLUX>(dev_frame<AX>(dl, s), dev_frame<AX>(dr, s), dev_ref_frame<AX>(rf, s), q);
        } else {
        Prim<R, NC> pl, pr;
        cons_to_prim<R, NC, PH>(P, t, pl, tl ? tl + k * ts : nullptr, kMemoRows * ts);
        if (BND) {
            bnd_state<R, NC, PH>(P, dir0, AX, s, pl, pr, bnd_axis);
        } else {
            #pragma unroll
            for (int v = 0; v < NV; v++)
                t[v] = c1[v] + (o11 + m1 * g1) * u1[4*v+1+T1] + (o12 + m1 * g2) * u1[4*v+1+T2];
            cons_to_prim<R, NC, PH>(P, t, pr, tr ? tr + k * ts : nullptr, kMemoRows * ts);
        }
        // face frame (s e_AX, e_T1, e_T2)
        Frame<R, NC> fl, fr;
        {
            const R vl[3] = {pl.u, pl.v, pl.w}, vr[3] = {pr.u, pr.v, pr.w};
            fl.r = pl.r; fl.e_tot = pl.e_tot; fl.p = pl.p; fl.cz = pl.cz; fl.gam = pl.gam;
            fr.r = pr.r; fr.e_tot = pr.e_tot; fr.p = pr.p; fr.cz = pr.cz; fr.gam = pr.gam;
            fl.un = s * vl[AX]; fr.un = s * vr[AX];
            fl.ut1 = vl[T1]; fr.ut1 = vr[T1];
            fl.ut2 = vl[T2]; fr.ut2 = vr[T2];
            #pragma unroll
            for (int i = 0; i < NC; i++) { fl.c[i] = pl.c[i]; fr.c[i] = pr.c[i]; }
        }
        ok &= flux_x<R, NC, FLUX>(fl, fr, q);
        }
        qg[AX] = s * q[0]; qg[T1] = q[1]; qg[T2] = q[2];
        #pragma unroll
        for (int v = 3; v < NV; v++) qg[v] = q[v];
        #pragma unroll
        for (int v = 0; v < NV; v++) { S0[v] += qg[v]; S1[v] += qg[v] * g1; S2[v] += qg[v] * g2; }
    }
    #pragma unroll
    for (int v = 0; v < NV; v++) {
        out[v] = S0[v] * wj;
        out[NV + v] = S1[v] * wj;
        out[2 * NV + v] = S2[v] * wj;
    }
    return ok;
}

// interior face from the two preloaded element states; the small side's
// level sets the face size
template <typename R, int NC, int FLUX, int AX, int PH, bool DEV = false>
__device__ __forceinline__ bool face_moments(const DevParams &P, const R *u0, const R *u1, int info,
                                             int lev0, R *out, const DevRef &rf, const R *tl = nullptr,
                                             const R *tr = nullptr, int64_t ts = 0)
{
    const int kind = (info >> 3) & 3;
    return face_core<R, NC, FLUX, AX, false, PH, DEV>(P, u0, u1, info, kind == FACE_HANG1 ? lev0 + 1 : lev0, out,
                                                      AX, rf, tl, tr, ts);
}

// Boundary faces: the element state is rotated into the frame (a, t1, t2)
// of its face (Taylor slopes and momentum components permuted), evaluated
// with the x-normal code, and the momentum moments are rotated back; one
// code path serves the three axes.
template <typename R, int NC, int FLUX, int PH, bool DEV = false>
__device__ __forceinline__ bool face_moments_bnd(const DevParams &P, const R *u, int info, int lev, R *out,
                                                 const DevRef &rf, const R *tl = nullptr, int64_t ts = 0)
{
    constexpr int NV = 4 + NC;
    const int ax = (info & 7) >> 1;
    R w[4 * NV];
    #pragma unroll
    for (int v = 0; v < NV; v++) {
        // source variable: momentum components follow the axis permutation
        R c[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            if (v < 3) {
                const R x0 = u[4 * 0 + k], x1 = u[4 * 1 + k], x2 = u[4 * 2 + k];
                // frame var v -> global var: v=0 -> ax, v=1 -> t1, v=2 -> t2
                const int gv = (v == 0) ? ax : ((v == 1) ? (ax == 0 ? 1 : 0) : (ax == 2 ? 1 : 2));
                c[k] = (gv == 0) ? x0 : ((gv == 1) ? x1 : x2);
            } else {
                c[k] = u[4 * v + k];
            }
        }
        const R ca = (ax == 0) ? c[1] : ((ax == 1) ? c[2] : c[3]);
        const R ct1 = (ax == 0) ? c[2] : c[1];
        const R ct2 = (ax == 2) ? c[2] : c[3];
        w[4 * v] = c[0]; w[4 * v + 1] = ca; w[4 * v + 2] = ct1; w[4 * v + 3] = ct2;
    }
    R o[3 * NV];
    bool ok;
    if constexpr (DEV) {
        // the chunk reference in the face's frame, like the state
        ok = face_core<R, NC, FLUX, 0, true, PH, true>(P, w, w, info, lev, o, ax, permute_ref(rf, ax));
    } else {
        ok = face_core<R, NC, FLUX, 0, true, PH>(P, w, w, info, lev, o, ax, rf, tl, nullptr, ts);
    }
    #pragma unroll
    for (int m = 0; m < 3; m++) {
        const R f0 = o[m * NV + 0], f1 = o[m * NV + 1], f2 = o[m * NV + 2];
        // global var g takes frame var: g == ax -> 0, g == t1 -> 1, g == t2 -> 2
        out[m * NV + 0] = (ax == 0) ? f0 : f1;
        out[m * NV + 1] = (ax == 1) ? f0 : ((ax == 0) ? f1 : f2);
        out[m * NV + 2] = (ax == 2) ? f0 : f2;
        #pragma unroll
        for (int v = 3; v < NV; v++) out[m * NV + v] = o[m * NV + v];
    }
    return ok;
}

// volume integral, src/models/euler/charm_model.c:66-132: with the Taylor
// basis on a cube, grad phi_k = e_k / h and phi_k(gp) = +-1/(2 sqrt 3), so
// int[v][1+d] -= (wJ/h) sum_gp F_d,v and the gravity term adds
// -g_v sum_gp rho phi_k wJ.
// tv (PH = 1, may be null): the EOS closure at volume point g starts at
// tv[g * ts], its planes kMemoRows * ts apart
template <typename R, int NC, int PH>
__device__ __forceinline__ void volume_integral(const DevParams &P, const R *u, R h, R *res,
                                                const R *tv = nullptr, int64_t ts = 0)
{
    constexpr int NV = 4 + NC;
    const R ph = R(0.28867513459481288225);           // 1/(2 sqrt 3)
    const R hinv = rcp(h);
    const R wj = h * h * h * R(0.125);
    R SF[3][NV];
    R G[4] = {R(0), R(0), R(0), R(0)};
    #pragma unroll
    for (int d = 0; d < 3; d++)
        #pragma unroll
        for (int v = 0; v < NV; v++) SF[d][v] = R(0);
    #pragma unroll
    for (int g = 0; g < 8; g++) {
        const R px = (g & 1) ? ph : -ph, py = (g & 2) ? ph : -ph, pz = (g & 4) ? ph : -ph;
        R t[NV];
        trace<R, NV>(u, px, py, pz, t);
        Prim<R, NC> p;
        cons_to_prim<R, NC, PH>(P, t, p, tv ? tv + g * ts : nullptr, kMemoRows * ts);
        const R m[3] = {t[0], t[1], t[2]};
        const R vel[3] = {p.u, p.v, p.w};
        #pragma unroll
        for (int d = 0; d < 3; d++) {
            SF[d][0] += m[d] * p.u + (d == 0 ? p.p : R(0));
            SF[d][1] += m[d] * p.v + (d == 1 ? p.p : R(0));
            SF[d][2] += m[d] * p.w + (d == 2 ? p.p : R(0));
            SF[d][3] += m[d] * p.e_tot + p.p * vel[d];
            #pragma unroll
            for (int i = 0; i < NC; i++) SF[d][4 + i] += m[d] * p.c[i];
        }
        G[0] += p.r; G[1] += p.r * px; G[2] += p.r * py; G[3] += p.r * pz;
    }
    const R f = wj * hinv;
    #pragma unroll
    for (int v = 0; v < NV; v++) {
        res[4*v+0] = R(0);
        res[4*v+1] = -SF[0][v] * f;
        res[4*v+2] = -SF[1][v] * f;
        res[4*v+3] = -SF[2][v] * f;
    }
    #pragma unroll
    for (int v = 0; v < 3; v++) {
        const R gv = cst<R>(P.grav[v], P.grav_f[v]);
        #pragma unroll
        for (int k = 0; k < 4; k++) res[4*v+k] -= gv * G[k] * wj;
    }
}

// the volume integral of F - F(U_r) from deviation coefficients u (one
// species): the momentum, energy and mass fluxes along each axis d as in
// dev_phys; the gravity term takes the full density
__device__ __forceinline__ void volume_integral_dev(const DevParams &P, const DevRef &rf, const float *u, float h,
                                                    float *res)
{
    const float ph = 0.28867513459481288225f;          // 1/(2 sqrt 3)
    const float hinv = rcp(h);
    const float wj = h * h * h * 0.125f;
    float SF[3][5];
    float G[4] = {0.f, 0.f, 0.f, 0.f};
    #pragma unroll
    for (int d = 0; d < 3; d++)
        #pragma unroll
        for (int v = 0; v < 5; v++) SF[d][v] = 0.f;
    #pragma unroll
    for (int g = 0; g < 8; g++) {
        const float px = (g & 1) ? ph : -ph, py = (g & 2) ? ph : -ph, pz = (g & 4) ? ph : -ph;
        float t[5];
        trace<float, 5>(u, px, py, pz, t);
        DPrim q;
        dev_prim(rf, t, q);
        #pragma unroll
        for (int d = 0; d < 3; d++) {
            #pragma unroll
            for (int j = 0; j < 3; j++) SF[d][j] += q.dm[d] * q.u[j] + rf.m[d] * q.du[j] + (d == j ? q.dp : 0.f);
            SF[d][3] += (q.dE + q.dp) * q.u[d] + rf.Epp * q.du[d];
            SF[d][4] += q.dm[d];
        }
        G[0] += q.r; G[1] += q.r * px; G[2] += q.r * py; G[3] += q.r * pz;
    }
    const float f = wj * hinv;
    #pragma unroll
    for (int v = 0; v < 5; v++) {
        res[4*v+0] = 0.f;
        res[4*v+1] = -SF[0][v] * f;
        res[4*v+2] = -SF[1][v] * f;
        res[4*v+3] = -SF[2][v] * f;
    }
    #pragma unroll
    for (int v = 0; v < 3; v++) {
        const float gv = P.grav_f[v];
        #pragma unroll
        for (int k = 0; k < 4; k++) res[4*v+k] -= gv * G[k] * wj;
    }
}

// S: storage type of the state; R: arithmetic type of the residual (S = R,
// or the mixed mode S = double, R = float: FP32 fluxes and integrals, FP64
// state, update and combine).  PH = 0: constant-cp Euler terms; PH = 1: any
// cp (FP64 Newton T) and the viscous integrals of the Navier-Stokes model
// added before A^-1.
// PH = 1 reads the constants from device memory (runtime-indexed power
// series and BC records), PH = 0 from the by-value kernel argument
template <int PH>
__device__ __forceinline__ const DevParams &stage_params(const StageArgs &A, const DevParams &Pv)
{
    if constexpr (PH == 1) return *A.prm;
    else return Pv;
}

// The FP32 deviation-form instances are held to at most 168 VGPRs (three
// waves per SIMD) so that a change of the code cannot drop them below the
// occupancy their LDS allows; with the rolled point loop they need 117, and
// four 40 KB workgroups share a CU
template <typename R, int NC, int FLUX, int PH>
constexpr int stage_min_waves()
{
    return uses_dev<R, NC, FLUX, PH>() ? 3 : 1;
}

template <typename S, typename R, int NC, int FLUX, int BS, int PH>
__global__ __launch_bounds__(BS, (stage_min_waves<R, NC, FLUX, PH>()))
void k_stage(StageArgs A, DevParams Pv)
{
    const DevParams &P = stage_params<PH>(A, Pv);
    constexpr int NV = 4 + NC;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    R *mom = reinterpret_cast<R *>(smem_raw);                       // [maxf][3][NV]
    int *finfo = reinterpret_cast<int *>(smem_raw + (size_t) A.max_faces * 3 * NV * sizeof(R));
    const S *__restrict__ U = reinterpret_cast<const S *>(A.U);
    const int c = xcd_block(A.n_xcd);
    if (c >= A.n_chunks) return;                 // whole workgroup exits together
    const int fb = A.ch_fbeg[c], fe = A.ch_fbeg[c + 1];
    const int eb = A.ch_ebeg[c], ee = A.ch_ebeg[c + 1];
    const int64_t n = A.n_elem;
    // FP32, one species: deviation form against the chunk's reference (the
    // cell mean of its first leaf), the same for every face and leaf of the
    // chunk
    constexpr bool DEV = uses_dev<R, NC, FLUX, PH>();
    DevRef rf{};
    if constexpr (DEV)
        rf = make_ref(P, (double) U[eb], (double) (U + 4 * n)[eb], (double) (U + 8 * n)[eb],
                      (double) (U + 12 * n)[eb], (double) (U + 16 * n)[eb]);
    bool ok = true;
    for (int f = fb + (int) threadIdx.x; f < fe; f += BS) {
        const int4 rec = A.cf4[f];
        R *dst = mom + (size_t) (f - fb) * 3 * NV;
        R ua[4 * NV];
        if constexpr (DEV) load_dev<S>(U, n, rec.x, rf, ua);
        else load_conv<S, R, NV>(U, n, rec.x, ua);
        const int lev0 = A.level[rec.x];
        // faces of a chunk are sorted by (axis, kind): waves are mostly
        // uniform in the branch taken here
        const int ax = (rec.z & 7) >> 1;
        // PH = 1: each side's point EOS closures from the point pass, at
        // the side's own (sub)face slot 4 d + j (point k at k * n)
        const R *pl = nullptr, *pr = nullptr;
        if (PH == 1 && A.tpt) {
            const R *TP = reinterpret_cast<const R *>(A.tpt);
            const int dir0 = rec.z & 7, kind = (rec.z >> 3) & 3, sub = (rec.z >> 5) & 3;
            const int s0 = 4 * dir0 + (kind == FACE_HANG1 ? sub : 0);
            const int s1 = 4 * (dir0 ^ 1) + (kind == FACE_HANG0 ? sub : 0);
            pl = TP + (int64_t) 6 * s0 * n + rec.x;
            if (kind != FACE_BND) pr = TP + (int64_t) 6 * s1 * n + rec.y;
        }
        if (((rec.z >> 3) & 3) == FACE_BND) {
            ok &= face_moments_bnd<R, NC, FLUX, PH, DEV>(P, ua, rec.z, lev0, dst, rf, pl, n);
        } else {
            R ub[4 * NV];
            if constexpr (DEV) load_dev<S>(U, n, rec.y, rf, ub);
            else load_conv<S, R, NV>(U, n, rec.y, ub);
            if (ax == 0)
                ok &= face_moments<R, NC, FLUX, 0, PH, DEV>(P, ua, ub, rec.z, lev0, dst, rf, pl, pr, n);
            else if (ax == 1)
                ok &= face_moments<R, NC, FLUX, 1, PH, DEV>(P, ua, ub, rec.z, lev0, dst, rf, pl, pr, n);
            else
                ok &= face_moments<R, NC, FLUX, 2, PH, DEV>(P, ua, ub, rec.z, lev0, dst, rf, pl, pr, n);
        }
        finfo[f - fb] = rec.z;
    }
    __syncthreads();
    const int e = eb + (int) threadIdx.x;
    if (e < ee) {
        R u[4 * NV], res[4 * NV];
        const int lev = A.level[e];
        const R h = ldexp(R(1), -lev);
        if constexpr (DEV) {
            load_dev<S>(U, n, e, rf, u);
            volume_integral_dev(P, rf, u, h, res);
        } else {
            load_conv<S, R, NV>(U, n, e, u);
            volume_integral<R, NC, PH>(P, u, h, res,
                                       (PH == 1 && A.tpt) ? reinterpret_cast<const R *>(A.tpt) + (int64_t) kMemoVolRow * n + e
                                                          : nullptr, n);
        }
        if constexpr (PH == 1) {
            if (A.visc) {
                const S *__restrict__ VI = reinterpret_cast<const S *>(A.visc);
                #pragma unroll
                for (int i = 0; i < 16; i++) res[i] += (R) (VI + i * A.n_elem)[e];
            }
        }
        const int s_end = A.slot_ptr[e + 1];
        for (int si = A.slot_ptr[e]; si < s_end; si++) {
            const int sl = A.slots[si];
            const int fl = sl >> 1, side = sl & 1;
            const int info = finfo[fl];
            const int dir0 = info & 7, kind = (info >> 3) & 3, sub = (info >> 5) & 3;
            const int a = dir0 >> 1;
            const R sg = side ? R(-1) : R(1);
            const R pa = ((dir0 & 1) ^ side) ? R(0.5) : R(-0.5);
            const bool big = (kind == FACE_HANG0 && side == 1) || (kind == FACE_HANG1 && side == 0);
            const R mu = big ? R(0.5) : R(1);
            const R d1 = big ? ((sub & 1) ? R(0.25) : R(-0.25)) : R(0);
            const R d2 = big ? ((sub & 2) ? R(0.25) : R(-0.25)) : R(0);
            const R *m = mom + (size_t) fl * 3 * NV;
            #pragma unroll
            for (int v = 0; v < NV; v++) {
                const R M0 = sg * m[v], M1 = sg * m[NV + v], M2 = sg * m[2 * NV + v];
                R cx, cy, cz;
                to_xyz<R>(a, pa * M0, mu * M1 + d1 * M0, mu * M2 + d2 * M0, cx, cy, cz);
                res[4*v+0] += M0;
                res[4*v+1] += cx;
                res[4*v+2] += cy;
                res[4*v+3] += cz;
            }
        }
        S *out = reinterpret_cast<S *>(A.out);
        if (A.mode == 0) {
            #pragma unroll
            for (int i = 0; i < 4 * NV; i++) stp<S>(out, n, i, e, (S) res[i]);
        } else {
            // c -= dt * A^-1 int (src/models/euler/charm_model.c:402-433); A^-1
            // of a level-l cube is the FP64 level-0 inverse times 8^l (exact
            // power-of-two scaling), diagonal on the cube octree; then the
            // SSP-RK combine when no limiter sits in between
            // (mixed mode: the stored FP64 coefficient is re-read, the FP32
            // integral widened, and the update and combine run in FP64)
            const S sc = ldexp(S(1), 3 * lev);
            const S dt = S(A.dt);
            // not __restrict__: the last RK stage writes its output over u^n
            const S *U0 = reinterpret_cast<const S *>(A.U0);
            #pragma unroll
            for (int v = 0; v < NV; v++)
                #pragma unroll
                for (int k = 0; k < 4; k++) {
                    const S ai = cst<S>(P.ainv0[k], P.ainv0_f[k]) * sc;
                    // the stored coefficient (the registers hold deviations in DEV)
                    const S uk = (std::is_same<S, R>::value && !DEV) ? (S) u[4*v+k] : ldp<S>(U, n, 4*v+k, e);
                    const S w = uk - dt * (ai * (S) res[4*v+k]);
                    stp<S>(out, n, 4*v+k, e, A.combine ? rk_combine<S>(A.combine, w, ldp<S>(U0, n, 4*v+k, e)) : w);
                }
        }
    }
    if (!ok) A.err[0] = 1;
}

template <typename S, typename R, int NC, int BS, int PH>
static hipError_t launch_stage_bs(int flux, const StageArgs &A, const DevParams &P, hipStream_t st)
{
    constexpr int NV = 4 + NC;
    const size_t sh = (size_t) A.max_faces * (3 * NV * sizeof(R) + sizeof(int));
    const dim3 grid((unsigned) (((A.n_chunks + A.n_xcd - 1) / A.n_xcd) * A.n_xcd)), block(BS);
    switch (flux) {
    case FLUX_LF: hipLaunchKernelGGL((k_stage<S, R, NC, FLUX_LF, BS, PH>), grid, block, sh, st, A, P); break;
    case FLUX_HLLC: hipLaunchKernelGGL((k_stage<S, R, NC, FLUX_HLLC, BS, PH>), grid, block, sh, st, A, P); break;
    case FLUX_GODUNOV: hipLaunchKernelGGL((k_stage<S, R, NC, FLUX_GODUNOV, BS, PH>), grid, block, sh, st, A, P); break;
    case FLUX_CD: hipLaunchKernelGGL((k_stage<S, R, NC, FLUX_CD, BS, PH>), grid, block, sh, st, A, P); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// host address of one stage kernel instance (for its attributes)
template <typename S, typename R, int NC, int BS, int PH>
static const void *stage_fn_bs(int flux)
{
    switch (flux) {
    case FLUX_LF: return reinterpret_cast<const void *>(&k_stage<S, R, NC, FLUX_LF, BS, PH>);
    case FLUX_HLLC: return reinterpret_cast<const void *>(&k_stage<S, R, NC, FLUX_HLLC, BS, PH>);
    case FLUX_GODUNOV: return reinterpret_cast<const void *>(&k_stage<S, R, NC, FLUX_GODUNOV, BS, PH>);
    case FLUX_CD: return reinterpret_cast<const void *>(&k_stage<S, R, NC, FLUX_CD, BS, PH>);
    }
    return nullptr;
}

template <typename S, typename R, int PH>
const void *stage_fn_t(int nc, int flux, int bs)
{
    const bool wide = PH == 1 || bs > 128;
    switch (nc) {
    case 1: return wide ? stage_fn_bs<S, R, 1, 256, PH>(flux) : stage_fn_bs<S, R, 1, 128, PH>(flux);
    case 2: return wide ? stage_fn_bs<S, R, 2, 256, PH>(flux) : stage_fn_bs<S, R, 2, 128, PH>(flux);
    case 3: return wide ? stage_fn_bs<S, R, 3, 256, PH>(flux) : stage_fn_bs<S, R, 3, 128, PH>(flux);
    case 4: return wide ? stage_fn_bs<S, R, 4, 256, PH>(flux) : stage_fn_bs<S, R, 4, 128, PH>(flux);
    }
    return nullptr;
}

// one thread per chunk element: 128- or 256-element chunks (the
// Navier-Stokes variant always runs 256-thread workgroups)
template <typename S, typename R, int PH>
hipError_t launch_stage_t(int nc, int flux, const StageArgs &A, const DevParams &P, hipStream_t st)
{
    if (A.chunk_elems > 256) return hipErrorInvalidValue;
    if constexpr (PH == 1) {
        switch (nc) {
        case 1: return launch_stage_bs<S, R, 1, 256, PH>(flux, A, P, st);
        case 2: return launch_stage_bs<S, R, 2, 256, PH>(flux, A, P, st);
        case 3: return launch_stage_bs<S, R, 3, 256, PH>(flux, A, P, st);
        case 4: return launch_stage_bs<S, R, 4, 256, PH>(flux, A, P, st);
        }
        return hipErrorInvalidValue;
    } else {
        const bool wide = A.chunk_elems > 128;
        switch (nc) {
        case 1: return wide ? launch_stage_bs<S, R, 1, 256, PH>(flux, A, P, st) : launch_stage_bs<S, R, 1, 128, PH>(flux, A, P, st);
        case 2: return wide ? launch_stage_bs<S, R, 2, 256, PH>(flux, A, P, st) : launch_stage_bs<S, R, 2, 128, PH>(flux, A, P, st);
        case 3: return wide ? launch_stage_bs<S, R, 3, 256, PH>(flux, A, P, st) : launch_stage_bs<S, R, 3, 128, PH>(flux, A, P, st);
        case 4: return wide ? launch_stage_bs<S, R, 4, 256, PH>(flux, A, P, st) : launch_stage_bs<S, R, 4, 128, PH>(flux, A, P, st);
        }
        return hipErrorInvalidValue;
    }
}

} // namespace charm
