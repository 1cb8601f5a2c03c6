// charm_dev.hpp -- the FP32 residual in deviation form (one species,
// constant gamma; LF, HLLC and central fluxes).
//
// Every chunk of the stage kernel picks a reference state U_r (the cell mean
// of its first leaf, taken as FP32 primitive values rho_r, u_r, p_r; its
// conserved values follow exactly in FP64).  All fluxes of the chunk's faces
// and volume points are evaluated as F(U) - F(U_r), with every term written
// so that no two O(|F|) quantities are subtracted: velocity deviations
// du = (dm - u_r drho) / rho, kinetic-energy deviation dK = (dm . u + rho_r
// u_r . du) / 2, dp = (gamma - 1)(dE - dK), and for HLLC the contact-speed
// offsets S* - u_K = (dp_R - dp_L - a_{R|L} (du_R - du_L)) / (a_L - a_R) with
// a_K = rho_K (S_K - u_K).  The dropped constant flux F(U_r) has an exactly
// zero DG residual (divergence theorem with the exact quadrature of the
// linear basis: sum over faces of F_r . n phi_k minus the volume integral of
// F_r . grad phi_k vanishes for every k, hanging faces included), so the
// residual of F - F(U_r) is the residual of F.  The FP32 rounding is then
// relative to the deviations (about 1% of |F| on the benchmark's perturbed
// Riemann box) instead of to |F| itself.
//
// The algebra is the reference's (src/charm_fluxes.c:255-315 LF, 319-437 HLLC
// x-split core, 513-559 central; src/models/euler/charm_model.c:66-132 volume
// fluxes); the FP64 path keeps the reference's arithmetic.
#pragma once
#include <hip/hip_runtime.h>
#include "charm_device.hpp"

namespace charm {

// chunk reference state: FP32 definitions and the FP64 conserved values they
// imply (m_r = rho_r u_r, E_r = p_r / (gamma - 1) + rho_r |u_r|^2 / 2)
struct DevRef {
    float r, u[3], p;
    float m[3], E, Epp;          // rounded rho_r u_r, E_r, E_r + p_r (multiply small terms only)
    double c0[5];                // exact (m_r[3], E_r, rho_r): subtracted from coefficient 0
    float c0h[5], c0l[5];        // c0 as the sum of two floats, c0h = c0 rounded
    float gam, gm1, gq;          // gamma, gamma - 1, (gamma + 1) / (2 gamma)
};

// reference from a cell mean (ru, rv, rw, re, rho), any finite positive
// density; a non-physical mean gives another (still exact) definition
__device__ __forceinline__ DevRef make_ref(const DevParams &P, double m0, double m1, double m2, double e0, double r0)
{
    DevRef R;
    R.gam = P.gam_f;
    R.gm1 = R.gam - 1.f;
    R.gq = (R.gam + 1.f) / (2.f * R.gam);
    if (!(r0 > 0.0) || !(r0 < 1e300)) { r0 = 1.0; m0 = m1 = m2 = 0.0; e0 = 1.0; }
    R.r = (float) r0;
    const double ri = 1.0 / (double) R.r;
    R.u[0] = (float) (m0 * ri); R.u[1] = (float) (m1 * ri); R.u[2] = (float) (m2 * ri);
    double kr = 0.0;
    #pragma unroll
    for (int d = 0; d < 3; d++) {
        R.c0[d] = (double) R.r * (double) R.u[d];          // exact product of two floats
        kr += R.c0[d] * (double) R.u[d];
    }
    kr *= 0.5;
    R.p = (float) ((double) R.gm1 * (e0 - kr));
    R.c0[3] = (double) R.p / (double) R.gm1 + kr;
    R.c0[4] = (double) R.r;
    #pragma unroll
    for (int d = 0; d < 3; d++) R.m[d] = (float) R.c0[d];
    #pragma unroll
    for (int i = 0; i < 5; i++) {
        R.c0h[i] = (float) R.c0[i];
        R.c0l[i] = (float) (R.c0[i] - (double) R.c0h[i]);
    }
    R.E = (float) R.c0[3];
    R.Epp = (float) (R.c0[3] + (double) R.p);
    return R;
}

// v_a of (v0, v1, v2) for a runtime axis a.  The arguments are values: a
// conditional over the lvalues R.u[0], R.u[1] would select their address,
// which keeps the whole reference record in private memory.
template <typename T>
__device__ __forceinline__ T vpick3(int a, T v0, T v1, T v2)
{
    return a == 0 ? v0 : (a == 1 ? v1 : v2);
}

// the reference in the frame (a, t1, t2) of a boundary face whose element
// state is permuted that way (momentum components and slopes): frame
// component 0 = axis a, 1 = the lower other axis, 2 = the upper one
__device__ __forceinline__ DevRef permute_ref(const DevRef &R, int a)
{
    DevRef Q;
    Q.r = R.r; Q.p = R.p; Q.E = R.E; Q.Epp = R.Epp;
    Q.gam = R.gam; Q.gm1 = R.gm1; Q.gq = R.gq;
    Q.c0[3] = R.c0[3]; Q.c0[4] = R.c0[4];
    #pragma unroll
    for (int i = 3; i < 5; i++) { Q.c0h[i] = R.c0h[i]; Q.c0l[i] = R.c0l[i]; }
    Q.c0h[0] = vpick3(a, R.c0h[0], R.c0h[1], R.c0h[2]);
    Q.c0h[1] = vpick3(a, R.c0h[1], R.c0h[0], R.c0h[0]);
    Q.c0h[2] = vpick3(a, R.c0h[2], R.c0h[2], R.c0h[1]);
    Q.c0l[0] = vpick3(a, R.c0l[0], R.c0l[1], R.c0l[2]);
    Q.c0l[1] = vpick3(a, R.c0l[1], R.c0l[0], R.c0l[0]);
    Q.c0l[2] = vpick3(a, R.c0l[2], R.c0l[2], R.c0l[1]);
    Q.u[0] = vpick3(a, R.u[0], R.u[1], R.u[2]);
    Q.u[1] = vpick3(a, R.u[1], R.u[0], R.u[0]);
    Q.u[2] = vpick3(a, R.u[2], R.u[2], R.u[1]);
    Q.m[0] = vpick3(a, R.m[0], R.m[1], R.m[2]);
    Q.m[1] = vpick3(a, R.m[1], R.m[0], R.m[0]);
    Q.m[2] = vpick3(a, R.m[2], R.m[2], R.m[1]);
    Q.c0[0] = vpick3(a, R.c0[0], R.c0[1], R.c0[2]);
    Q.c0[1] = vpick3(a, R.c0[1], R.c0[0], R.c0[0]);
    Q.c0[2] = vpick3(a, R.c0[2], R.c0[2], R.c0[1]);
    return Q;
}

// element coefficients as deviations: coefficient 0 minus the reference,
// slopes as they are.  An FP64 state subtracts in FP64 (one rounding of the
// difference); an FP32 state subtracts the two-float reference c0h + c0l in
// FP32, (x - c0h) - c0l, two roundings, each relative to the deviation.
template <typename S>
__device__ __forceinline__ void load_dev(const S *__restrict__ U, int64_t n, int e, const DevRef &R, float *u)
{
    #pragma unroll
    for (int i = 0; i < 20; i++) {
        const S x = ldp<S>(U, n, i, e);
        if constexpr (sizeof(S) == 4)
            u[i] = (i & 3) ? x : (x - R.c0h[i >> 2]) - R.c0l[i >> 2];
        else
            u[i] = (i & 3) ? (float) x : (float) ((double) x - R.c0[i >> 2]);
    }
}

// state at a point from its deviation trace t = (dm[3], dE, drho)
struct DPrim {
    float r, ri, u[3], p, E, cz;
    float dr, dm[3], du[3], dE, dp;
};

__device__ __forceinline__ void dev_prim(const DevRef &R, const float *t, DPrim &q)
{
    q.dm[0] = t[0]; q.dm[1] = t[1]; q.dm[2] = t[2]; q.dE = t[3]; q.dr = t[4];
    q.r = R.r + q.dr;
    q.ri = rcp(q.r);
    #pragma unroll
    for (int d = 0; d < 3; d++) {
        q.du[d] = (q.dm[d] - R.u[d] * q.dr) * q.ri;
        q.u[d] = R.u[d] + q.du[d];
    }
    const float dk = 0.5f * (q.dm[0] * q.u[0] + q.dm[1] * q.u[1] + q.dm[2] * q.u[2]
                             + R.m[0] * q.du[0] + R.m[1] * q.du[1] + R.m[2] * q.du[2]);
    q.dp = R.gm1 * (q.dE - dk);
    q.p = R.p + q.dp;
    q.E = R.E + q.dE;
    q.cz = dsqrt(R.gam * q.p * q.ri);
}

// deviation state of a full primitive state (exterior states of boundary
// faces other than the transmissive one): plain FP32 accuracy
__device__ __forceinline__ void dev_from_prim(const DevRef &R, const Prim<float, 1> &p, DPrim &q)
{
    q.r = p.r; q.ri = rcp(p.r);
    const float v[3] = {p.u, p.v, p.w};
    #pragma unroll
    for (int d = 0; d < 3; d++) {
        q.u[d] = v[d];
        q.du[d] = v[d] - R.u[d];
        q.dm[d] = p.r * v[d] - R.m[d];
    }
    q.p = p.p; q.dp = p.p - R.p;
    q.E = p.r * p.e_tot; q.dE = q.E - R.E;
    q.dr = p.r - R.r;
    q.cz = p.cz;
}

// full primitive state of a deviation state (input of the BC functions)
__device__ __forceinline__ void prim_of_dev(const DevParams &P, const DevRef &R, const DPrim &q, Prim<float, 1> &p)
{
    p.r = q.r; p.u = q.u[0]; p.v = q.u[1]; p.w = q.u[2];
    p.p = q.p; p.cz = q.cz; p.gam = R.gam;
    p.e_tot = q.E * q.ri;
    p.e = q.p * q.ri * rcp(R.gm1);
    p.cv = P.cv_f;
    p.m = P.m_f;
    p.c[0] = 1.f;
}

// face-frame view (normal axis AX with sign s, tangents T1 < T2)
struct DFrame {
    float r, ri, un, ut1, ut2, p, E, cz;
    float dr, dmn, dmt1, dmt2, dun, dE, dp;
};

template <int AX>
__device__ __forceinline__ DFrame dev_frame(const DPrim &q, float s)
{
    constexpr int T1 = (AX == 0) ? 1 : 0, T2 = (AX == 2) ? 1 : 2;
    DFrame f;
    f.r = q.r; f.ri = q.ri; f.p = q.p; f.E = q.E; f.cz = q.cz;
    f.un = s * q.u[AX]; f.ut1 = q.u[T1]; f.ut2 = q.u[T2];
    f.dr = q.dr; f.dE = q.dE; f.dp = q.dp;
    f.dmn = s * q.dm[AX]; f.dmt1 = q.dm[T1]; f.dmt2 = q.dm[T2];
    f.dun = s * q.du[AX];
    return f;
}

// reference values in the same frame
struct DRefFrame {
    float mn, mt1, mt2, Epp, gq, gam;
};

template <int AX>
__device__ __forceinline__ DRefFrame dev_ref_frame(const DevRef &R, float s)
{
    constexpr int T1 = (AX == 0) ? 1 : 0, T2 = (AX == 2) ? 1 : 2;
    DRefFrame f;
    f.mn = s * R.m[AX]; f.mt1 = R.m[T1]; f.mt2 = R.m[T2]; f.Epp = R.Epp; f.gq = R.gq; f.gam = R.gam;
    return f;
}

__device__ __forceinline__ DFrame sel_dframe(bool c, const DFrame &a, const DFrame &b)
{
    DFrame o;
    o.r = sel(c, a.r, b.r); o.ri = sel(c, a.ri, b.ri); o.un = sel(c, a.un, b.un); o.ut1 = sel(c, a.ut1, b.ut1); o.ut2 = sel(c, a.ut2, b.ut2);
    o.p = sel(c, a.p, b.p); o.E = sel(c, a.E, b.E); o.cz = sel(c, a.cz, b.cz);
    o.dr = sel(c, a.dr, b.dr); o.dmn = sel(c, a.dmn, b.dmn); o.dmt1 = sel(c, a.dmt1, b.dmt1);
    o.dmt2 = sel(c, a.dmt2, b.dmt2); o.dun = sel(c, a.dun, b.dun); o.dE = sel(c, a.dE, b.dE);
    o.dp = sel(c, a.dp, b.dp);
    return o;
}

// physical flux deviation F_n(U) - F_n(U_r) in the face frame
// (q = (normal momentum, t1, t2, energy, mass))
__device__ __forceinline__ void dev_phys(const DFrame &K, const DRefFrame &R, float *q)
{
    q[0] = K.dmn * K.un + R.mn * K.dun + K.dp;
    q[1] = K.dmt1 * K.un + R.mt1 * K.dun;
    q[2] = K.dmt2 * K.un + R.mt2 * K.dun;
    q[3] = (K.dE + K.dp) * K.un + R.Epp * K.dun;
    q[4] = K.dmn;
}

// HLLC (src/charm_fluxes.c:319-437) minus F(U_r): the upwind branches are
// the physical flux deviation, the star branch adds S_K (U*_K - U_K) =
// S_K g (rho_K S_K, rho_K ut1, rho_K ut2, E_K + a_K S* + p_K, rho_K),
// g = (S* - u_K) / (S_K - S*)
__device__ __forceinline__ bool hllc_dev(const DFrame &L, const DFrame &Rt, const DRefFrame &rf, float *q)
{
    const float p_pvrs = 0.5f * (L.p + Rt.p) - 0.5f * (Rt.un - L.un) * 0.25f * (L.r + Rt.r) * (L.cz + Rt.cz);
    const float p_star = p_pvrs > 0.f ? p_pvrs : 0.f;
    // c_K q_K with q_K = sqrt(1 + gq (p* / p_K - 1)) for p* > p_K, else 1, as
    // one root: c_K q_K = sqrt(gamma / rho_K (p_K + gq max(p* - p_K, 0)))
    float sl = L.un - dsqrt(rf.gam * L.ri * (L.p + rf.gq * fmaxf(p_star - L.p, 0.f)));
    float sr = Rt.un + dsqrt(rf.gam * Rt.ri * (Rt.p + rf.gq * fmaxf(p_star - Rt.p, 0.f)));
    if (sl > sr) { const float t = sl; sl = sr; sr = t; }
    const float al = L.r * (sl - L.un), ar = Rt.r * (sr - Rt.un);
    const float dpn = Rt.dp - L.dp, dun = Rt.dun - L.dun;
    const float den = rcp(al - ar);
    float dl = (dpn - ar * dun) * den;              // S* - u_L
    float dr = (dpn - al * dun) * den;              // S* - u_R
    float s_star = L.un + dl;
    if (s_star < sl) { s_star = sl; dl = sl - L.un; dr = sl - Rt.un; }
    if (s_star > sr) { s_star = sr; dl = sr - L.un; dr = sr - Rt.un; }
    const bool ok = (sl <= s_star) && (s_star <= sr);
    // one branch-free body: the upwind side of a supersonic face (sl >= 0:
    // left, sr <= 0: right) with no star term, else the star region of the
    // side s* falls on
    const bool sup = (sl >= 0.f) || (sr <= 0.f);
    const bool left = (sl >= 0.f) || (!(sr <= 0.f) && s_star >= 0.f);
    const DFrame K = sel_dframe(left, L, Rt);
    const float sk = left ? sl : sr, dk = left ? dl : dr;
    dev_phys(K, rf, q);
    const float sg = sup ? 0.f : sk * dk * rcp(sk - s_star);          // S_K g
    const float ak = K.r * (sk - K.un);
    q[0] += sg * (K.r * sk);
    q[1] += sg * (K.r * K.ut1);
    q[2] += sg * (K.r * K.ut2);
    q[3] += sg * (K.E + ak * s_star + K.p);
    q[4] += sg * K.r;
    return ok;
}

// Lax-Friedrichs / central (src/charm_fluxes.c:255-315, 513-559) minus F(U_r)
__device__ __forceinline__ void lf_dev(const DFrame &L, const DFrame &Rt, const DRefFrame &rf, float *q,
                                       bool central)
{
    float fl[5], fr[5];
    dev_phys(L, rf, fl);
    dev_phys(Rt, rf, fr);
    const float alpha = central ? 0.f : fmaxf(fabsf(L.un) + L.cz, fabsf(Rt.un) + Rt.cz);
    q[0] = 0.5f * (fr[0] + fl[0] - alpha * (Rt.dmn - L.dmn));
    q[1] = 0.5f * (fr[1] + fl[1] - alpha * (Rt.dmt1 - L.dmt1));
    q[2] = 0.5f * (fr[2] + fl[2] - alpha * (Rt.dmt2 - L.dmt2));
    q[3] = 0.5f * (fr[3] + fl[3] - alpha * (Rt.dE - L.dE));
    q[4] = 0.5f * (fr[4] + fl[4] - alpha * (Rt.dr - L.dr));
}

template <int FLUX>
__device__ __forceinline__ bool flux_dev(const DFrame &L, const DFrame &Rt, const DRefFrame &rf, float *q)
{
    if constexpr (FLUX == FLUX_HLLC) {
        return hllc_dev(L, Rt, rf, q);
    } else {
        lf_dev(L, Rt, rf, q, FLUX == FLUX_CD);
        return true;
    }
}

// whether a stage kernel instance evaluates its FP32 residual in deviation form
template <typename R, int NC, int FLUX, int PH>
constexpr bool uses_dev()
{
    return sizeof(R) == 4 && NC == 1 && PH == 0 && FLUX != FLUX_GODUNOV;
}

} // namespace charm
