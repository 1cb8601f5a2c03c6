// charm_gmesh.hip -- gfx950 kernels of the multi-tree (general hexahedral)
// meshes: the residual of a stage as a face pass and an element pass, and
// the Barth-Jespersen limiter, on the per-point basis values the host
// computed from each leaf's trilinear geometry (charm_gmesh.c).
//
// Face pass, one thread per face: at its 6 points both traces (side 1 from
// the boundary condition on a boundary face), the flux along the side-0
// normal in the reference's face frame, and the moment sums of both sides,
// each written to its leaf's slot for that local face, FQ[6][4V][n] (side 1
// negated; on a conforming forest every local face of a leaf has exactly one
// face).  Element pass, one thread per leaf: the volume integral at 8 points
// (flux tensor against grad phi = e / dh, gravity on the momentum), its six
// face slots (coalesced, no atomics), the full 4x4 A^-1, the update and the
// SSP-RK combine; ahead of the limiter also the cell means of the new state.  The arithmetic is the
// reference's Euler model (src/models/euler/charm_model.c:66-212, 215-372,
// 402-433) with the geometry of src/charm_geom.c.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "charm_device.hpp"
#include "charm_kernels.h"
#include "charm_stage.hpp"

namespace charm {

namespace {

// exterior state of a boundary face for an arbitrary unit outward normal n
// (src/charm_bnd_cond.c:21-146; the axis-aligned bnd_state with n = s e_a)
template <typename R, int NC>
__device__ __forceinline__ void bnd_state_n(const DevParams &P, int slot, const R n[3],
                                            const Prim<R, NC> &in, Prim<R, NC> &out)
{
    out = in;
    R t = dv(in.e, in.cv);
    R pr = in.p;
    R vel[3] = {in.u, in.v, in.w};
    const int type = P.bnd_type[slot];
    if (type == BND_INLET) {
        #pragma unroll
        for (int i = 0; i < 3; i++) vel[i] = cst<R>(P.bnd_param[slot][i], P.bnd_param_f[slot][i]);
        t = cst<R>(P.bnd_param[slot][3], P.bnd_param_f[slot][3]);
        pr = cst<R>(P.bnd_param[slot][4], P.bnd_param_f[slot][4]);
    } else if (type == BND_WALL_SLIP || type == BND_WALL_NO_SLIP) {
        const R vn = vel[0] * n[0] + vel[1] * n[1] + vel[2] * n[2];
        #pragma unroll
        for (int i = 0; i < 3; i++) {
            vel[i] -= n[i] * vn;
            if (type == BND_WALL_SLIP) vel[i] -= n[i] * vn;
        }
        if (type == BND_WALL_NO_SLIP) t = cst<R>(P.bnd_param[slot][0], P.bnd_param_f[slot][0]);
    } else if (type == BND_MASS_FLOW) {
        R cs[3];
        #pragma unroll
        for (int i = 0; i < 3; i++) cs[i] = cst<R>(P.bnd_param[slot][3 + i], P.bnd_param_f[slot][3 + i]);
        const R fd = n[0] * cs[0] + n[1] * cs[1] + n[2] * cs[2];
        const R mflow = cst<R>(P.bnd_param[slot][0], P.bnd_param_f[slot][0]);
        if (fd > R(0)) {
            const R vn = vel[0] * n[0] + vel[1] * n[1] + vel[2] * n[2];
            if (vn < in.cz) pr = cst<R>(P.bnd_param[slot][6], P.bnd_param_f[slot][6]);
        } else {
            Prim<R, NC> q = in;
            #pragma unroll
            for (int i = 0; i < NC; i++) q.c[i] = cst<R>(P.bnd_param[slot][7 + i], P.bnd_param_f[slot][7 + i]);
            eos_coeffs<R, NC>(P, q);
            const R r = in.r < R(kEps) ? R(kEps) : in.r;
            const R p0 = r * in.e * (q.gam - R(1));
            const R v2 = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
            pr = v2 > dv(q.gam * p0, r) ? cst<R>(P.bnd_param[slot][1], P.bnd_param_f[slot][1]) : p0;
            const R cv = q.cv, cp = q.gam * cv;
            const R x = dv(mflow * (cp - cv), pr * fd);
            const R k = dv(x * x, R(2) * cp);
            const R ttot = cst<R>(P.bnd_param[slot][2], P.bnd_param_f[slot][2]);
            t = dv(R(-1) + dsqrt(R(1) + R(4) * k * ttot), R(2) * k);
            const R vm = fabs(dv(mflow, r * fd));
            vel[0] = cs[0] * vm; vel[1] = cs[1] * vm; vel[2] = cs[2] * vm;
            #pragma unroll
            for (int i = 0; i < NC; i++) out.c[i] = q.c[i];
        }
    }
    out.u = vel[0]; out.v = vel[1]; out.w = vel[2];
    eos_coeffs<R, NC>(P, out);
    if (pr < R(kEps)) pr = R(kEps);
    out.p = pr;
    out.r = dv(pr * out.m, t * R(kGasR));
    out.cz = dsqrt(dv(out.gam * pr, out.r));
    out.e = dv(pr, out.r * (out.gam - R(1)));
    out.e_tot = out.e + R(0.5) * (out.u * out.u + out.v * out.v + out.w * out.w);
}

template <typename R, int NV>
__device__ __forceinline__ void trace_at(const R *u, const R b[3], R *t)
{
    #pragma unroll
    for (int v = 0; v < NV; v++) t[v] = u[4 * v] + b[0] * u[4 * v + 1] + b[1] * u[4 * v + 2] + b[2] * u[4 * v + 3];
}

template <typename R, int NC>
__device__ __forceinline__ void prim_to_cons(const Prim<R, NC> &p, R *c)
{
    c[0] = p.r * p.u; c[1] = p.r * p.v; c[2] = p.r * p.w;
    c[3] = p.r * (p.e + R(0.5) * (p.u * p.u + p.v * p.v + p.w * p.w));
    #pragma unroll
    for (int i = 0; i < NC; i++) c[4 + i] = p.r * p.c[i];
}

// blocks rounded up to a multiple of the XCD count
static inline int64_t xcd_grid(int64_t blocks, int n_xcd)
{
    return (blocks + n_xcd - 1) / n_xcd * n_xcd;
}

} // namespace

template <typename R, int NC, int FLUX>
__global__ __launch_bounds__(256)
void k_gface(GArgs A, DevParams P)
{
    constexpr int NV = 4 + NC;
    const int64_t nf = A.nf, n = A.n;
    const int64_t f = (int64_t) xcd_block(A.n_xcd) * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const R *__restrict__ U = reinterpret_cast<const R *>(A.U);
    const R *__restrict__ FG = reinterpret_cast<const R *>(A.fg) + f;
    const int e0 = A.fe[f], e1 = A.fe[nf + f];
    R u0[4 * NV], u1[4 * NV];
    #pragma unroll
    for (int i = 0; i < 4 * NV; i++) u0[i] = (U + i * n)[e0];
    #pragma unroll
    for (int i = 0; i < 4 * NV; i++) u1[i] = e1 >= 0 ? (U + i * n)[e1] : R(0);
    double nn[3] = {A.fn[f], A.fn[nf + f], A.fn[2 * nf + f]}, nt[3][3];
    ref_face_frame(nn, nt);
    const R n3[3] = {R(nn[0]), R(nn[1]), R(nn[2])};
    R s0[NV][4], s1[NV][4];
    #pragma unroll
    for (int v = 0; v < NV; v++)
        #pragma unroll
        for (int k = 0; k < 4; k++) { s0[v][k] = R(0); s1[v][k] = R(0); }
    bool ok = true;
    // side 1's basis is the affine image of side 0's: b1 = b0 r + o per axis
    const R r1[3] = {FG[18 * nf], FG[19 * nf], FG[20 * nf]}, o1[3] = {FG[21 * nf], FG[22 * nf], FG[23 * nf]};
    for (int p = 0; p < 6; p++) {
        const R b0[3] = {FG[(3 * p) * nf], FG[(3 * p + 1) * nf], FG[(3 * p + 2) * nf]};
        const R b1[3] = {b0[0] * r1[0] + o1[0], b0[1] * r1[1] + o1[1], b0[2] * r1[2] + o1[2]};
        const R wj = FG[(24 + p) * nf];
        R t[NV];
        Prim<R, NC> pa, pb;
        trace_at<R, NV>(u0, b0, t);
        cons_to_prim<R, NC, 0>(P, t, pa);
        if (e1 >= 0) {
            trace_at<R, NV>(u1, b1, t);
            cons_to_prim<R, NC, 0>(P, t, pb);
        } else {
            bnd_state_n<R, NC>(P, -1 - e1, n3, pa, pb);
        }
        const Frame<R, NC> fa = to_frame<R, NC>(pa, R(nt[0][0]), R(nt[0][1]), R(nt[0][2]), R(nt[1][0]), R(nt[1][1]),
                                                R(nt[1][2]), R(nt[2][0]), R(nt[2][1]), R(nt[2][2]));
        const Frame<R, NC> fb = to_frame<R, NC>(pb, R(nt[0][0]), R(nt[0][1]), R(nt[0][2]), R(nt[1][0]), R(nt[1][1]),
                                                R(nt[1][2]), R(nt[2][0]), R(nt[2][1]), R(nt[2][2]));
        R q[NV], g[NV];
        ok &= flux_x<R, NC, FLUX>(fa, fb, q);
        #pragma unroll
        for (int d = 0; d < 3; d++) g[d] = q[0] * R(nt[0][d]) + q[1] * R(nt[1][d]) + q[2] * R(nt[2][d]);
        #pragma unroll
        for (int v = 3; v < NV; v++) g[v] = q[v];
        const R ph0[4] = {wj, b0[0] * wj, b0[1] * wj, b0[2] * wj};
        const R ph1[4] = {wj, b1[0] * wj, b1[1] * wj, b1[2] * wj};
        #pragma unroll
        for (int v = 0; v < NV; v++)
            #pragma unroll
            for (int k = 0; k < 4; k++) { s0[v][k] += g[v] * ph0[k]; s1[v][k] += g[v] * ph1[k]; }
    }
    R *__restrict__ FQ = reinterpret_cast<R *>(A.fq);
    const int l0 = A.ff[f], l1 = A.ff[nf + f];
    #pragma unroll
    for (int v = 0; v < NV; v++)
        #pragma unroll
        for (int k = 0; k < 4; k++) (FQ + (int64_t) (4 * NV * l0 + 4 * v + k) * n)[e0] = s0[v][k];
    if (e1 >= 0) {
        #pragma unroll
        for (int v = 0; v < NV; v++)
            #pragma unroll
            for (int k = 0; k < 4; k++) (FQ + (int64_t) (4 * NV * l1 + 4 * v + k) * n)[e1] = -s1[v][k];
    }
    if (!ok) A.err[0] = 1;
}

template <typename R, int NC>
__global__ __launch_bounds__(256)
void k_gelem(GArgs A, DevParams P)
{
    constexpr int NV = 4 + NC;
    const int64_t n = A.n;
    const int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const R *__restrict__ U = reinterpret_cast<const R *>(A.U);
    const R *__restrict__ LG = reinterpret_cast<const R *>(A.lg) + e;
    R u[4 * NV];
    #pragma unroll
    for (int i = 0; i < 4 * NV; i++) u[i] = (U + i * n)[e];
    R in[NV][4];
    #pragma unroll
    for (int v = 0; v < NV; v++)
        #pragma unroll
        for (int k = 0; k < 4; k++) in[v][k] = R(0);
    const R dhi[3] = {LG[0], LG[n], LG[2 * n]};
    const R grav[3] = {cst<R>(P.grav[0], P.grav_f[0]), cst<R>(P.grav[1], P.grav_f[1]), cst<R>(P.grav[2], P.grav_f[2])};
    for (int g = 0; g < 8; g++) {
        const R b[3] = {LG[(3 + 3 * g) * n], LG[(4 + 3 * g) * n], LG[(5 + 3 * g) * n]};
        const R wj = LG[(27 + g) * n];
        R t[NV];
        trace_at<R, NV>(u, b, t);
        Prim<R, NC> p;
        cons_to_prim<R, NC, 0>(P, t, p);
        const R vel[3] = {p.u, p.v, p.w};
        // flux along axis d: (m_d u + p e_d, m_d e_tot + p u_d, m_d c), m = (ru, rv, rw)
        #pragma unroll
        for (int d = 0; d < 3; d++) {
            const R md = t[d], gd = dhi[d] * wj;
            in[0][1 + d] -= (md * p.u + (d == 0 ? p.p : R(0))) * gd;
            in[1][1 + d] -= (md * p.v + (d == 1 ? p.p : R(0))) * gd;
            in[2][1 + d] -= (md * p.w + (d == 2 ? p.p : R(0))) * gd;
            in[3][1 + d] -= (md * p.e_tot + p.p * vel[d]) * gd;
            #pragma unroll
            for (int i = 0; i < NC; i++) in[4 + i][1 + d] -= md * p.c[i] * gd;
        }
        const R phi[4] = {wj, b[0] * wj, b[1] * wj, b[2] * wj};
        #pragma unroll
        for (int d = 0; d < 3; d++)
            #pragma unroll
            for (int k = 0; k < 4; k++) in[d][k] -= p.r * grav[d] * phi[k];
    }
    const R *__restrict__ FQ = reinterpret_cast<const R *>(A.fq);
    #pragma unroll 1
    for (int lf = 0; lf < 6; lf++)
        #pragma unroll
        for (int v = 0; v < NV; v++)
            #pragma unroll
            for (int k = 0; k < 4; k++) in[v][k] += (FQ + (int64_t) (4 * NV * lf + 4 * v + k) * n)[e];
    R *__restrict__ out = reinterpret_cast<R *>(A.out);
    if (A.mode == 0) {
        #pragma unroll
        for (int v = 0; v < NV; v++)
            #pragma unroll
            for (int k = 0; k < 4; k++) (out + (4 * v + k) * n)[e] = in[v][k];
        return;
    }
    // c -= dt A^-1 int, then the SSP-RK combine (U0 is read before out is written)
    const R dt = R(A.dt);
    const R *U0 = reinterpret_cast<const R *>(A.U0);
    R *MB = reinterpret_cast<R *>(A.mb);
    const R mb[3] = {LG[75 * n], LG[76 * n], LG[77 * n]};
    #pragma unroll
    for (int v = 0; v < NV; v++) {
        R w[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            R du = R(0);
            #pragma unroll
            for (int j = 0; j < 4; j++) du += LG[(35 + 4 * k + j) * n] * in[v][j];
            w[k] = u[4 * v + k] - dt * du;
            (out + (4 * v + k) * n)[e] = A.combine ? rk_combine<R>(A.combine, w[k], (U0 + (4 * v + k) * n)[e]) : w[k];
        }
        // the limiter's cell mean of this variable (the basis means of the leaf)
        if (MB) (MB + v * n)[e] = w[0] + w[1] * mb[0] + w[2] * mb[1] + w[3] * mb[2];
    }
}

// Barth-Jespersen on the general mesh (src/limiters/charm_limiter_bj.c:
// 11-332): the leaf's mean and its six neighbours' (a boundary face gives
// the mean pushed through the boundary condition), psi from the 8 vertex
// values, the slopes scaled, then the SSP-RK combine.  A leaf's mean is
// U0 + sum_d U_d mb_d (mb: the leaf means of the linear basis functions),
// read from the element pass's means when it wrote them for this state.
template <typename R, int NC>
__global__ __launch_bounds__(256)
void k_glimiter(GArgs A, DevParams P)
{
    constexpr int NV = 4 + NC;
    const int64_t n = A.n, nf = A.nf;
    const int64_t e = (int64_t) xcd_block(A.n_xcd) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const R *__restrict__ W = reinterpret_cast<const R *>(A.U);
    const R *__restrict__ LG = reinterpret_cast<const R *>(A.lg);
    R u[4 * NV];
    #pragma unroll
    for (int i = 0; i < 4 * NV; i++) u[i] = (W + i * n)[e];
    const R *__restrict__ MB = reinterpret_cast<const R *>(A.mb);
    auto mean = [&](int64_t j, int v) {
        if (MB) return (MB + v * n)[j];
        return (W + 4 * v * n)[j] + (W + (4 * v + 1) * n)[j] * (LG + 75 * n)[j] +
               (W + (4 * v + 2) * n)[j] * (LG + 76 * n)[j] + (W + (4 * v + 3) * n)[j] * (LG + 77 * n)[j];
    };
    R u0[NV], umin[NV], umax[NV];
    #pragma unroll
    for (int v = 0; v < NV; v++) { u0[v] = mean(e, v); umin[v] = u0[v]; umax[v] = u0[v]; }
    for (int d = 0; d < 6; d++) {
        const int nb = A.lnb[d * n + e];
        R m[NV];
        if (nb >= 0) {
            #pragma unroll
            for (int v = 0; v < NV; v++) m[v] = mean(nb, v);
        } else {
            const int fi = -1 - nb;
            const R n3[3] = {R(A.fn[fi]), R(A.fn[nf + fi]), R(A.fn[2 * nf + fi])};
            Prim<R, NC> pi, po;
            cons_to_prim<R, NC, 0>(P, u0, pi);
            bnd_state_n<R, NC>(P, -1 - A.fe[nf + fi], n3, pi, po);
            prim_to_cons<R, NC>(po, m);
        }
        #pragma unroll
        for (int v = 0; v < NV; v++) {
            umin[v] = m[v] < umin[v] ? m[v] : umin[v];
            umax[v] = m[v] > umax[v] ? m[v] : umax[v];
        }
    }
    R bv[8][3];
    #pragma unroll
    for (int vt = 0; vt < 8; vt++)
        #pragma unroll
        for (int a = 0; a < 3; a++) bv[vt][a] = (LG + (51 + 3 * vt + a) * n)[e];
    // psi = min over the vertices of min(1, (umax - u0) / df) for df > eps and
    // min(1, (umin - u0) / df) for df < -eps, df = u(vertex) - u0: with the
    // numerators' fixed signs, the quotient by the largest |df| on each side
    #pragma unroll
    for (int v = 0; v < NV; v++) {
        R dpos = R(0), dneg = R(0);
        #pragma unroll
        for (int vt = 0; vt < 8; vt++) {
            const R f = u[4 * v] + bv[vt][0] * u[4 * v + 1] + bv[vt][1] * u[4 * v + 2] + bv[vt][2] * u[4 * v + 3];
            const R df = f - u0[v];
            dpos = (df > R(kEps) && df > dpos) ? df : dpos;
            dneg = (-df > R(kEps) && -df > dneg) ? -df : dneg;
        }
        R psi = R(1);
        if (dpos > R(0)) psi = fmin(psi, dv(umax[v] - u0[v], dpos));
        if (dneg > R(0)) psi = fmin(psi, dv(u0[v] - umin[v], dneg));
        u[4 * v + 1] *= psi; u[4 * v + 2] *= psi; u[4 * v + 3] *= psi;
    }
    R *out = reinterpret_cast<R *>(A.out);
    const R *U0 = reinterpret_cast<const R *>(A.U0);
    #pragma unroll
    for (int i = 0; i < 4 * NV; i++)
        (out + i * n)[e] = A.combine ? rk_combine<R>(A.combine, u[i], (U0 + i * n)[e]) : u[i];
}

template <typename R, int NC>
static hipError_t launch_gstage_r(int flux, const GArgs &A, const DevParams &P, hipStream_t st)
{
    // the face grid is a multiple of the XCD count (xcd_block's bijection);
    // the padding blocks return
    const dim3 block(256), gf((unsigned) xcd_grid((A.nf + 255) / 256, A.n_xcd)), ge((unsigned) ((A.n + 255) / 256));
    switch (flux) {
    case FLUX_LF: hipLaunchKernelGGL((k_gface<R, NC, FLUX_LF>), gf, block, 0, st, A, P); break;
    case FLUX_HLLC: hipLaunchKernelGGL((k_gface<R, NC, FLUX_HLLC>), gf, block, 0, st, A, P); break;
    case FLUX_GODUNOV: hipLaunchKernelGGL((k_gface<R, NC, FLUX_GODUNOV>), gf, block, 0, st, A, P); break;
    case FLUX_CD: hipLaunchKernelGGL((k_gface<R, NC, FLUX_CD>), gf, block, 0, st, A, P); break;
    default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_gelem<R, NC>), ge, block, 0, st, A, P);
    return hipGetLastError();
}

hipError_t launch_gstage(int prec, int nc, int flux, const GArgs &A, const DevParams &P, hipStream_t st)
{
    if (A.n <= 0) return hipSuccess;
#define CHARM_G_CASE(NCV) \
    case NCV: return prec ? launch_gstage_r<float, NCV>(flux, A, P, st) : launch_gstage_r<double, NCV>(flux, A, P, st);
    switch (nc) {
    CHARM_G_CASE(1)
    CHARM_G_CASE(2)
    CHARM_G_CASE(3)
    CHARM_G_CASE(4)
    default: return hipErrorInvalidValue;
    }
#undef CHARM_G_CASE
}

hipError_t launch_glimiter(int prec, int nc, const GArgs &A, const DevParams &P, hipStream_t st)
{
    if (A.n <= 0) return hipSuccess;
    const dim3 block(256), ge((unsigned) xcd_grid((A.n + 255) / 256, A.n_xcd));
#define CHARM_G_CASE(NCV)                                                                              \
    case NCV:                                                                                          \
        if (prec) hipLaunchKernelGGL((k_glimiter<float, NCV>), ge, block, 0, st, A, P);               \
        else hipLaunchKernelGGL((k_glimiter<double, NCV>), ge, block, 0, st, A, P);                   \
        break;
    switch (nc) {
    CHARM_G_CASE(1)
    CHARM_G_CASE(2)
    CHARM_G_CASE(3)
    CHARM_G_CASE(4)
    default: return hipErrorInvalidValue;
    }
#undef CHARM_G_CASE
    return hipGetLastError();
}

} // namespace charm
