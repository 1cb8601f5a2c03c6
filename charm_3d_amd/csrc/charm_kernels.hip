// charm_kernels.hip -- gfx950 kernels of the DG residual / time-step path
// (limiter, reductions, pointwise fluxes, conversions; the stage kernel is
// in charm_stage.hpp, instantiated by charm_stage_*.hip).
//
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <mutex>
#include "charm_device.hpp"
#include "charm_kernels.h"
#include "charm_stage.hpp"

namespace charm {

// Barth-Jespersen limiter, src/limiters/charm_limiter_bj.c:11-332, fused
// with the SSP-RK combine that follows it (src/models/ns/charm_model.c:
// 136-184).  Neighbour means are coefficient 0 (the Taylor basis has zero
// mean for k > 0); a big element's hanging face contributes the volume
// average of its four small neighbours, a boundary face the mean pushed
// through the boundary condition.  Only coefficient 0 of neighbours is read,
// so the slopes of W may be limited while other threads read means.
template <typename R, int NC, int PH>
__global__ __launch_bounds__(256)
void k_limiter(LimArgs A, DevParams P)
{
    constexpr int NV = 4 + NC;
    const int e = xcd_block(A.n_xcd) * blockDim.x + threadIdx.x;
    if (e >= A.n) return;
    const R *__restrict__ W = reinterpret_cast<const R *>(A.W);
    const int64_t n = A.n;
    R u[4 * NV], u_old[4 * NV];
    load_elem<R, NV>(W, n, e, u);
    // the RK combine's old state is fetched up front with the rest
    if (A.combine) load_elem<R, NV>(reinterpret_cast<const R *>(A.U0), n, e, u_old);
    R umin[NV], umax[NV];
    #pragma unroll
    for (int v = 0; v < NV; v++) { umin[v] = u[4*v]; umax[v] = u[4*v]; }
    if (A.limit) {
        // all six neighbour indices, then all six neighbour means, are put
        // in flight together; the rare boundary / coarse-side neighbours
        // (negative codes) are replaced afterwards
        int nb[6];
        {
            // 24 bytes per element: three 8-byte loads (the array is
            // 256-byte aligned)
            const int2 *q = reinterpret_cast<const int2 *>(A.nbr + 6 * (int64_t) e);
            const int2 a = q[0], b = q[1], c = q[2];
            nb[0] = a.x; nb[1] = a.y; nb[2] = b.x; nb[3] = b.y; nb[4] = c.x; nb[5] = c.y;
        }
        R m[6][NV];
        #pragma unroll
        for (int d = 0; d < 6; d++) {
            const int j = nb[d] >= 0 ? nb[d] : e;
            #pragma unroll
            for (int v = 0; v < NV; v++) m[d][v] = (W + 4 * v * n)[j];
        }
        #pragma unroll
        for (int d = 0; d < 6; d++) {
            if (nb[d] >= 0) continue;
            if (nb[d] > -8) {
                // boundary: mean -> prim -> BC -> cons (charm_limiter_bj.c:79-91)
                R t[NV];
                #pragma unroll
                for (int v = 0; v < NV; v++) t[v] = u[4*v];
                Prim<R, NC> pi, po;
                cons_to_prim<R, NC, PH>(P, t, pi);
                const int face = -1 - nb[d], a = face >> 1;
                bnd_state<R, NC, PH>(P, face, a, (face & 1) ? R(1) : R(-1), pi, po, -1);
                m[d][0] = po.r * po.u; m[d][1] = po.r * po.v; m[d][2] = po.r * po.w;
                m[d][3] = po.r * (po.e + R(0.5) * (po.u * po.u + po.v * po.v + po.w * po.w));
                #pragma unroll
                for (int i = 0; i < NC; i++) m[d][4 + i] = po.r * po.c[i];
            } else {
                const int4 g = A.aux[-8 - nb[d]];
                const int q[4] = {g.x, g.y, g.z, g.w};
                #pragma unroll
                for (int v = 0; v < NV; v++) m[d][v] = R(0);
                for (int j = 0; j < 4; j++)
                    #pragma unroll
                    for (int v = 0; v < NV; v++) m[d][v] += (W + 4 * v * n)[q[j]];
                #pragma unroll
                for (int v = 0; v < NV; v++) m[d][v] *= R(0.25);
            }
        }
        #pragma unroll
        for (int d = 0; d < 6; d++)
            #pragma unroll
            for (int v = 0; v < NV; v++) {
                umin[v] = m[d][v] < umin[v] ? m[d][v] : umin[v];
                umax[v] = m[d][v] > umax[v] ? m[d][v] : umax[v];
            }
        // psi = min over the 8 vertices of min(1, (umax - u0) / df) where
        // df > eps and min(1, (umin - u0) / df) where df < -eps.  The
        // numerators have fixed signs (umin <= u0 <= umax), so the minimum
        // over vertices is the quotient by the largest |df| on each side:
        // two divisions per variable instead of eight, the same value
        // (division is monotone in the divisor).
        #pragma unroll
        for (int v = 0; v < NV; v++) {
            const R u0 = u[4*v];
            R dpos = R(0), dneg = R(0);
            #pragma unroll
            for (int vt = 0; vt < 8; vt++) {
                const R f = u0 + ((vt & 1) ? R(0.5) : R(-0.5)) * u[4*v+1]
                               + ((vt & 2) ? R(0.5) : R(-0.5)) * u[4*v+2]
                               + ((vt & 4) ? R(0.5) : R(-0.5)) * u[4*v+3];
                const R df = f - u0;
                dpos = (df > R(kEps) && df > dpos) ? df : dpos;
                dneg = (-df > R(kEps) && -df > dneg) ? -df : dneg;
            }
            R psi = R(1);
            if (dpos > R(0)) psi = fmin(psi, dv(umax[v] - u0, dpos));
            if (dneg > R(0)) psi = fmin(psi, dv(u0 - umin[v], dneg));
            u[4*v+1] *= psi; u[4*v+2] *= psi; u[4*v+3] *= psi;
        }
    }
    R *out = reinterpret_cast<R *>(A.out);
    if (A.combine) {
        #pragma unroll
        for (int i = 0; i < 4 * NV; i++) (out + i * n)[e] = rk_combine<R>(A.combine, u[i], u_old[i]);
    } else {
        #pragma unroll
        for (int i = 0; i < 4 * NV; i++) (out + i * n)[e] = u[i];
    }
}

// per-element CFL step and FP64 conservation sums, reduced per block
template <typename R, int NC, int PH>
__global__ __launch_bounds__(256)
void k_reduce(RedArgs A, DevParams P)
{
    constexpr int NV = 4 + NC;
    __shared__ double sh[256][NV + 1];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    double loc[NV + 1];
    #pragma unroll
    for (int v = 0; v <= NV; v++) loc[v] = 0.0;
    loc[NV] = 1.e300;
    if (e < A.n) {
        const R *u = reinterpret_cast<const R *>(A.U) + e;
        const double h = ldexp(1.0, -(int) A.level[e]);
        const double vol = h * h * h;
        #pragma unroll
        for (int v = 0; v < NV; v++) loc[v] = vol * (double) u[4 * v * A.n];
        R t[NV];
        #pragma unroll
        for (int v = 0; v < NV; v++) t[v] = u[4 * v * A.n];
        Prim<R, NC> p;
        cons_to_prim<R, NC, PH>(P, t, p);
        const double len = A.mode == 0 ? vol : h;
        const double spd = sqrt((double) (p.u * p.u + p.v * p.v + p.w * p.w)) + (double) p.cz;
        loc[NV] = P.cfl * len / spd;
    }
    #pragma unroll
    for (int v = 0; v <= NV; v++) sh[threadIdx.x][v] = loc[v];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int) threadIdx.x < s) {
            #pragma unroll
            for (int v = 0; v < NV; v++) sh[threadIdx.x][v] += sh[threadIdx.x + s][v];
            sh[threadIdx.x][NV] = fmin(sh[threadIdx.x][NV], sh[threadIdx.x + s][NV]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        #pragma unroll
        for (int v = 0; v <= NV; v++) A.partial[(size_t) blockIdx.x * (NV + 1) + v] = sh[0][v];
    }
}

// final pass of k_reduce over the block partials (one block); writes the
// first nout of the nv sums and the CFL minimum
__global__ __launch_bounds__(256)
void k_reduce_final(const double *partial, int nblk, int nv, double *out, int nout)
{
    __shared__ double sh[256];
    for (int v = 0; v <= nv; v++) {
        const bool isdt = (v == nv);
        double acc = isdt ? 1.e300 : 0.0;
        for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
            const double x = partial[(size_t) b * (nv + 1) + v];
            acc = isdt ? fmin(acc, x) : acc + x;
        }
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int) threadIdx.x < s)
                sh[threadIdx.x] = isdt ? fmin(sh[threadIdx.x], sh[threadIdx.x + s])
                                       : sh[threadIdx.x] + sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0 && v < nout) out[v] = sh[0];
        __syncthreads();
    }
}

// pointwise flux with an arbitrary unit normal: the reference's tangent
// frame (src/charm_fluxes.c:458-488), the frame flux, rotation back (:507-509)
template <typename R, int NC, int FLUX>
__global__ void k_flux_points(FluxArgs A, DevParams P)
{
    constexpr int NV = 4 + NC;
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    Prim<R, NC> pp[2];
    for (int s = 0; s < 2; s++) {
        const double *row = (s ? A.prim_r : A.prim_l) + i * (5 + NC);
        Prim<R, NC> &p = pp[s];
        p.r = R(row[0]); p.u = R(row[1]); p.v = R(row[2]); p.w = R(row[3]); p.p = R(row[4]);
        #pragma unroll
        for (int c = 0; c < NC; c++) p.c[c] = R(row[5 + c]);
        eos_coeffs<R, NC>(P, p);
        p.e = p.p / (p.r * (p.gam - R(1)));
        p.e_tot = p.e + R(0.5) * (p.u * p.u + p.v * p.v + p.w * p.w);
        p.cz = dsqrt(p.gam * p.p / p.r);
    }
    double nt[3][3];
    ref_face_frame(A.normals + 3 * i, nt);
    Frame<R, NC> f[2];
    for (int s = 0; s < 2; s++)
        f[s] = to_frame<R, NC>(pp[s], R(nt[0][0]), R(nt[0][1]), R(nt[0][2]), R(nt[1][0]), R(nt[1][1]),
                               R(nt[1][2]), R(nt[2][0]), R(nt[2][1]), R(nt[2][2]));
    R q[NV];
    flux_x<R, NC, FLUX>(f[0], f[1], q);
    double *o = A.out + i * NV;
    for (int d = 0; d < 3; d++)
        o[d] = (double) (q[0] * R(nt[0][d]) + q[1] * R(nt[1][d]) + q[2] * R(nt[2][d]));
    for (int v = 3; v < NV; v++) o[v] = (double) q[v];
}

// precision and layout conversions between the FP64 host layout [e][v][k]
// and the device state [(4 v + k) n + e]
template <typename R>
__global__ void k_from_f64(const double *src, R *dst, int64_t n, int nvk)
{
    const int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    for (int i = 0; i < nvk; i++) dst[i * n + e] = (R) src[e * nvk + i];
}
template <typename R>
__global__ void k_to_f64(const R *src, double *dst, int64_t n, int nvk)
{
    const int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    for (int i = 0; i < nvk; i++) dst[e * nvk + i] = (double) src[i * n + e];
}

// plain device copy in 16-byte words (state checkpoints; also the
// known-size reference for the HBM byte counters)
__global__ __launch_bounds__(256)
void k_copy16(const int4 *__restrict__ src, int4 *__restrict__ dst, int64_t n16)
{
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (int64_t) gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st)
{
    if (bytes % 16) return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st);
    const int64_t n16 = (int64_t) (bytes / 16);
    const unsigned blocks = (unsigned) ((n16 + 255) / 256 < 8192 ? (n16 + 255) / 256 : 8192);
    if (n16 == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, st, (const int4 *) src, (int4 *) dst, n16);
    return hipGetLastError();
}

// ---------------------------------------------------------------- launchers
hipError_t launch_stage(int prec, int nc, int flux, int ph, const StageArgs &A, const DevParams &P, hipStream_t st)
{
    if (prec == 2) return ph ? hipErrorInvalidValue : launch_stage_mix_ph0(nc, flux, A, P, st);
    if (prec) return ph ? launch_stage_f32_ph1(nc, flux, A, P, st) : launch_stage_f32_ph0(nc, flux, A, P, st);
    return ph ? launch_stage_f64_ph1(nc, flux, A, P, st) : launch_stage_f64_ph0(nc, flux, A, P, st);
}

const void *stage_kernel(int prec, int nc, int flux, int ph, int bs)
{
    if (prec == 2) return ph ? nullptr : stage_fn_mix_ph0(nc, flux, bs);
    if (prec) return ph ? stage_fn_f32_ph1(nc, flux, bs) : stage_fn_f32_ph0(nc, flux, bs);
    return ph ? stage_fn_f64_ph1(nc, flux, bs) : stage_fn_f64_ph0(nc, flux, bs);
}

size_t stage_smem_bytes(int prec, int nc, int max_faces)
{
    return (size_t) max_faces * (3 * (4 + nc) * (prec ? 4 : 8) + sizeof(int));
}

template <typename R, int PH>
static hipError_t launch_limiter_r(int nc, const LimArgs &A, const DevParams &P, hipStream_t st)
{
    const int64_t nb = (A.n + 255) / 256;
    const dim3 grid((unsigned) (((nb + A.n_xcd - 1) / A.n_xcd) * A.n_xcd)), block(256);
    switch (nc) {
    case 1: hipLaunchKernelGGL((k_limiter<R, 1, PH>), grid, block, 0, st, A, P); break;
    case 2: hipLaunchKernelGGL((k_limiter<R, 2, PH>), grid, block, 0, st, A, P); break;
    case 3: hipLaunchKernelGGL((k_limiter<R, 3, PH>), grid, block, 0, st, A, P); break;
    case 4: hipLaunchKernelGGL((k_limiter<R, 4, PH>), grid, block, 0, st, A, P); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_limiter(int prec, int nc, int ph, const LimArgs &A, const DevParams &P, hipStream_t st)
{
    if (ph) return prec ? launch_limiter_r<float, 1>(nc, A, P, st) : launch_limiter_r<double, 1>(nc, A, P, st);
    return prec ? launch_limiter_r<float, 0>(nc, A, P, st) : launch_limiter_r<double, 0>(nc, A, P, st);
}

template <typename R, int PH>
static hipError_t launch_reduce_r(int nc, const RedArgs &A, const DevParams &P, hipStream_t st, int nout)
{
    const dim3 grid((unsigned) ((A.n + 255) / 256)), block(256);
    switch (nc) {
    case 1: hipLaunchKernelGGL((k_reduce<R, 1, PH>), grid, block, 0, st, A, P); break;
    case 2: hipLaunchKernelGGL((k_reduce<R, 2, PH>), grid, block, 0, st, A, P); break;
    case 3: hipLaunchKernelGGL((k_reduce<R, 3, PH>), grid, block, 0, st, A, P); break;
    case 4: hipLaunchKernelGGL((k_reduce<R, 4, PH>), grid, block, 0, st, A, P); break;
    default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, A.partial, (int) grid.x, 4 + nc, A.result, nout);
    return hipGetLastError();
}

hipError_t launch_reduce_n(int prec, int nc, int ph, const RedArgs &A, const DevParams &P, hipStream_t st,
                           int nout)
{
    if (ph) return prec ? launch_reduce_r<float, 1>(nc, A, P, st, nout) : launch_reduce_r<double, 1>(nc, A, P, st, nout);
    return prec ? launch_reduce_r<float, 0>(nc, A, P, st, nout) : launch_reduce_r<double, 0>(nc, A, P, st, nout);
}

hipError_t launch_reduce(int prec, int nc, int ph, const RedArgs &A, const DevParams &P, hipStream_t st)
{
    return launch_reduce_n(prec, nc, ph, A, P, st, 4 + nc + 1);
}

// Block partials of charm_gpu_conservation_sums_noalloc: a static device
// array, one per GPU the library runs on, held until the process ends
// (kNoAllocBlocks blocks of 256 leaves, V + 1 <= 9 doubles each: 1.2 MB, up
// to 4,194,304 leaves).  Host code never passes the array's name to a
// kernel: that is a host-side stand-in, not a device address.  The device
// address comes from hipGetSymbolAddress, once per device, with that device
// current, and is kept per device.
constexpr int kNoAllocBlocks = 16384;
static __device__ __align__(256) double scratch[(size_t) kNoAllocBlocks * (4 + 4 + 1)];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];

hipError_t noalloc_scratch(void **ptr, size_t *bytes)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    *bytes = sizeof(scratch);
    return hipSuccess;
}

template <typename R, int NC>
static hipError_t launch_flux_nc(int flux, const FluxArgs &A, const DevParams &P, hipStream_t st)
{
    const dim3 grid((unsigned) ((A.n + 127) / 128)), block(128);
    switch (flux) {
    case FLUX_LF: hipLaunchKernelGGL((k_flux_points<R, NC, FLUX_LF>), grid, block, 0, st, A, P); break;
    case FLUX_HLLC: hipLaunchKernelGGL((k_flux_points<R, NC, FLUX_HLLC>), grid, block, 0, st, A, P); break;
    case FLUX_GODUNOV: hipLaunchKernelGGL((k_flux_points<R, NC, FLUX_GODUNOV>), grid, block, 0, st, A, P); break;
    case FLUX_CD: hipLaunchKernelGGL((k_flux_points<R, NC, FLUX_CD>), grid, block, 0, st, A, P); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R>
static hipError_t launch_flux_r(int nc, int flux, const FluxArgs &A, const DevParams &P, hipStream_t st)
{
    switch (nc) {
    case 1: return launch_flux_nc<R, 1>(flux, A, P, st);
    case 2: return launch_flux_nc<R, 2>(flux, A, P, st);
    case 3: return launch_flux_nc<R, 3>(flux, A, P, st);
    case 4: return launch_flux_nc<R, 4>(flux, A, P, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_flux_points(int prec, int nc, int flux, const FluxArgs &A, const DevParams &P, hipStream_t st)
{
    return prec ? launch_flux_r<float>(nc, flux, A, P, st) : launch_flux_r<double>(nc, flux, A, P, st);
}

// The branch hllc_x takes at one face point, as kHllcBits bits: 0 p* > p_L (shock
// factor on the left), 1 p* > p_R, 2 the wave speeds swapped, 3 S_L >= 0,
// 4 S_R <= 0, 5 S* >= 0, 6 S* clamped to S_L, 7 S* clamped to S_R, 8 the
// PVRS pressure clamped to zero.  The conditions restate hllc_x's
// (src/charm_fluxes.c:319-437) in the same arithmetic.
template <int NC>
__device__ int hllc_branch(const Frame<double, NC> &L, const Frame<double, NC> &Rt)
{
    const double p_pvrs = 0.5 * (L.p + Rt.p) - 0.5 * (Rt.un - L.un) * 0.25 * (L.r + Rt.r) * (L.cz + Rt.cz);
    const double p_star = p_pvrs > 0.0 ? p_pvrs : 0.0;
    const double gl = (L.gam + 1.0) / (2.0 * L.gam), gr = (Rt.gam + 1.0) / (2.0 * Rt.gam);
    const double ql = (p_star <= L.p) ? 1.0 : sqrt(1.0 + gl * (p_star / L.p - 1.0));
    const double qr = (p_star <= Rt.p) ? 1.0 : sqrt(1.0 + gr * (p_star / Rt.p - 1.0));
    double sl = L.un - L.cz * ql, sr = Rt.un + Rt.cz * qr;
    const bool swap = sl > sr;
    if (swap) { const double t = sl; sl = sr; sr = t; }
    double s_star = Rt.p - L.p;
    s_star += L.r * L.un * (sl - L.un);
    s_star -= Rt.r * Rt.un * (sr - Rt.un);
    s_star = s_star / (L.r * (sl - L.un) - Rt.r * (sr - Rt.un));
    const bool lo = s_star < sl;
    if (lo) s_star = sl;
    const bool hi = s_star > sr;
    if (hi) s_star = sr;
    return (p_star > L.p) | (p_star > Rt.p) << 1 | swap << 2 | (sl >= 0.0) << 3 | (sr <= 0.0) << 4 |
           (s_star >= 0.0) << 5 | lo << 6 | hi << 7 | !(p_pvrs > 0.0) << 8;
}

// the HLLC branch at the six points of an interior face for one state
// (the traces and face frame of face_core)
template <int NC, int AX>
__device__ void face_branches(const DevParams &P, const double *u0, const double *u1, int info, int *code)
{
    constexpr int NV = 4 + NC;
    constexpr int T1 = (AX == 0) ? 1 : 0, T2 = (AX == 2) ? 1 : 2;
    const int dir0 = info & 7, kind = (info >> 3) & 3, sub = (info >> 5) & 3;
    const double s = (dir0 & 1) ? 1.0 : -1.0;
    const double pa0 = 0.5 * s, pa1 = -pa0;
    const double d1 = (sub & 1) ? 0.25 : -0.25, d2 = (sub & 2) ? 0.25 : -0.25;
    const bool big0 = (kind == FACE_HANG1), big1 = (kind == FACE_HANG0);
    const double m0 = big0 ? 0.5 : 1.0, m1 = big1 ? 0.5 : 1.0;
    const double o01 = big0 ? d1 : 0.0, o02 = big0 ? d2 : 0.0, o11 = big1 ? d1 : 0.0, o12 = big1 ? d2 : 0.0;
    for (int k = 0; k < 6; k++) {
        const double g1 = gp_t1<double>(k), g2 = gp_t2<double>(k);
        double ta[NV], tb[NV];
        #pragma unroll
        for (int v = 0; v < NV; v++) {
            ta[v] = u0[4*v] + pa0 * u0[4*v+1+AX] + (o01 + m0 * g1) * u0[4*v+1+T1] + (o02 + m0 * g2) * u0[4*v+1+T2];
            tb[v] = u1[4*v] + pa1 * u1[4*v+1+AX] + (o11 + m1 * g1) * u1[4*v+1+T1] + (o12 + m1 * g2) * u1[4*v+1+T2];
        }
        Prim<double, NC> pl, pr;
        cons_to_prim<double, NC, 0>(P, ta, pl);
        cons_to_prim<double, NC, 0>(P, tb, pr);
        Frame<double, NC> fl, fr;
        const double vl[3] = {pl.u, pl.v, pl.w}, vr[3] = {pr.u, pr.v, pr.w};
        fl.r = pl.r; fl.e_tot = pl.e_tot; fl.p = pl.p; fl.cz = pl.cz; fl.gam = pl.gam;
        fr.r = pr.r; fr.e_tot = pr.e_tot; fr.p = pr.p; fr.cz = pr.cz; fr.gam = pr.gam;
        fl.un = s * vl[AX]; fr.un = s * vr[AX];
        fl.ut1 = vl[T1]; fr.ut1 = vr[T1];
        fl.ut2 = vl[T2]; fr.ut2 = vr[T2];
        #pragma unroll
        for (int i = 0; i < NC; i++) { fl.c[i] = pl.c[i]; fr.c[i] = pr.c[i]; }
        code[k] = hllc_branch<NC>(fl, fr);
    }
}

// one workgroup per chunk; each interior face counted in the chunk that
// holds its side-0 element (once)
template <int NC>
__global__ __launch_bounds__(256) void k_hllc_flips(FlipArgs A, DevParams P)
{
    constexpr int NV = 4 + NC;
    constexpr int NQ = 2 + kHllcBits;
    __shared__ unsigned long long red[NQ][256];
    const int c = blockIdx.x;
    const int fb = A.ch_fbeg[c], fe = A.ch_fbeg[c + 1], eb = A.ch_ebeg[c], ee = A.ch_ebeg[c + 1];
    unsigned long long q[NQ];
    #pragma unroll
    for (int i = 0; i < NQ; i++) q[i] = 0;
    for (int f = fb + (int) threadIdx.x; f < fe; f += 256) {
        const int4 rec = A.cf4[f];
        if (((rec.z >> 3) & 3) == FACE_BND || rec.x < eb || rec.x >= ee) continue;
        double a0[4 * NV], a1[4 * NV], b0[4 * NV], b1[4 * NV];
        load_elem<double, NV>(A.Ua, A.n, rec.x, a0);
        load_elem<double, NV>(A.Ua, A.n, rec.y, a1);
        load_elem<double, NV>(A.Ub, A.n, rec.x, b0);
        load_elem<double, NV>(A.Ub, A.n, rec.y, b1);
        int ca[6], cb[6];
        const int ax = (rec.z & 7) >> 1;
        if (ax == 0) { face_branches<NC, 0>(P, a0, a1, rec.z, ca); face_branches<NC, 0>(P, b0, b1, rec.z, cb); }
        else if (ax == 1) { face_branches<NC, 1>(P, a0, a1, rec.z, ca); face_branches<NC, 1>(P, b0, b1, rec.z, cb); }
        else { face_branches<NC, 2>(P, a0, a1, rec.z, ca); face_branches<NC, 2>(P, b0, b1, rec.z, cb); }
        for (int k = 0; k < 6; k++) {
            const int x = ca[k] ^ cb[k];
            q[0]++;
            q[1] += x != 0;
            #pragma unroll
            for (int i = 0; i < kHllcBits; i++) q[2 + i] += (x >> i) & 1;
        }
    }
    #pragma unroll
    for (int i = 0; i < NQ; i++) red[i][threadIdx.x] = q[i];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int) threadIdx.x < w)
            for (int i = 0; i < NQ; i++) red[i][threadIdx.x] += red[i][threadIdx.x + w];
        __syncthreads();
    }
    if ((int) threadIdx.x < NQ) atomicAdd(A.out + threadIdx.x, red[threadIdx.x][0]);
}

hipError_t launch_hllc_flips(int nc, const FlipArgs &A, const DevParams &P, hipStream_t st)
{
    if (A.n_chunks <= 0) return hipSuccess;
    const dim3 grid((unsigned) A.n_chunks), block(256);
    switch (nc) {
    case 1: hipLaunchKernelGGL(k_hllc_flips<1>, grid, block, 0, st, A, P); break;
    case 2: hipLaunchKernelGGL(k_hllc_flips<2>, grid, block, 0, st, A, P); break;
    case 3: hipLaunchKernelGGL(k_hllc_flips<3>, grid, block, 0, st, A, P); break;
    case 4: hipLaunchKernelGGL(k_hllc_flips<4>, grid, block, 0, st, A, P); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_convert(int prec, int to_f64, const void *src, void *dst, int64_t n, int nvk, hipStream_t st)
{
    const dim3 grid((unsigned) ((n + 255) / 256)), block(256);
    if (n <= 0) return hipSuccess;
    if (to_f64) {
        if (prec) hipLaunchKernelGGL(k_to_f64<float>, grid, block, 0, st, (const float *) src, (double *) dst, n, nvk);
        else hipLaunchKernelGGL(k_to_f64<double>, grid, block, 0, st, (const double *) src, (double *) dst, n, nvk);
    } else {
        if (prec) hipLaunchKernelGGL(k_from_f64<float>, grid, block, 0, st, (const double *) src, (float *) dst, n, nvk);
        else hipLaunchKernelGGL(k_from_f64<double>, grid, block, 0, st, (const double *) src, (double *) dst, n, nvk);
    }
    return hipGetLastError();
}

} // namespace charm
