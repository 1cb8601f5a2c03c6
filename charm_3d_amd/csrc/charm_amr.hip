// charm_amr.hip -- gfx950 kernels of the mesh adaptation: the density
// gradient error estimator and the projection of the state onto a new leaf
// set.  Both are per-leaf, HBM-bound gathers; arithmetic is FP64 whatever
// the storage precision of the state.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "charm_device.hpp"
#include "charm_kernels.h"

namespace charm {

namespace {

constexpr int kAuxBase = 8;            // nbr <= -8: group of four small neighbours
constexpr int kRootLen = 1 << 19;      // octree root edge in integer units

// density coefficients of leaf e: rho = sum_i rc_i is linear in the Taylor
// coefficients, so its coefficients are the species sums
template <typename S, int NC>
__device__ __forceinline__ void rho_coeffs(const S *__restrict__ U, int64_t n, int e, double r[4])
{
    #pragma unroll
    for (int k = 0; k < 4; k++) {
        double s = 0.0;
        #pragma unroll
        for (int i = 0; i < NC; i++) s += (double) (U + (4 * (4 + i) + k) * n)[e];
        r[k] = s;
    }
}

} // namespace

// Green-Gauss density gradient and the interpolation error estimate of
// src/charm_amr.c:350-599 and :12-30:
//   grad = (1/V) sum_faces int_f (rho_0 + rho_1)/2 n dA,
//   err2 = |grad|^2 / 12 * V^(5/3).
// The integrand is linear on each (sub)face, where the reference's six-point
// rule is exact, so each face contributes its area times the mean of the two
// traces at the (sub)face centre.  Boundary faces evaluate the boundary state
// at the six face Gauss points.  One thread per leaf
// gathers over its six face slots (no atomics); rho_c is the floored density
// of the cell mean (the value src/charm_amr.c:116-127 compares).
template <typename S, int NC, int PH>
__global__ __launch_bounds__(256)
void k_amr_indicator(AmrArgs A, DevParams P)
{
    constexpr int NV = 4 + NC;
    const int64_t n = A.n;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const S *__restrict__ U = reinterpret_cast<const S *>(A.U);
    const int lev = A.level[e];
    const double h = ldexp(1.0, -lev), area = h * h;
    double r[4];
    rho_coeffs<S, NC>(U, n, e, r);
    double g[3] = {0.0, 0.0, 0.0};
    #pragma unroll
    for (int d = 0; d < 6; d++) {
        const int a = d >> 1, t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
        const double sg = (d & 1) ? 1.0 : -1.0, half = 0.5 * sg;
        const double rs = r[0] + half * r[1 + a];
        const int nb = A.nbr[6 * e + d];
        double sum;
        if (nb >= 0) {
            double q[4];
            rho_coeffs<S, NC>(U, n, nb, q);
            double rn = q[0] - half * q[1 + a];
            if (A.level[nb] != lev) {
                // this leaf is sub-face j of the coarser neighbour's face:
                // j's bits select the tangential quarter offsets of its centre
                const int4 grp = A.aux[-kAuxBase - A.nbr[6 * nb + (d ^ 1)]];
                const int j = (grp.x == e) ? 0 : ((grp.y == e) ? 1 : ((grp.z == e) ? 2 : 3));
                rn += ((j & 1) ? 0.25 : -0.25) * q[1 + t1] + ((j & 2) ? 0.25 : -0.25) * q[1 + t2];
            }
            sum = area * (0.5 * (rs + rn));
        } else if (nb > -kAuxBase) {
            // the exterior density is not linear in general (a no-slip wall
            // imposes T: rho_b = p M / (R T_w)), so this face takes the
            // six-point rule, w J = h^2 / 6
            double c[NV][4];
            #pragma unroll
            for (int v = 0; v < NV; v++)
                #pragma unroll
                for (int k = 0; k < 4; k++) c[v][k] = (double) (U + (4 * v + k) * n)[e];
            sum = 0.0;
            for (int k = 0; k < 6; k++) {
                const double o1 = gp_t1<double>(k), o2 = gp_t2<double>(k);
                double u[NV];
                #pragma unroll
                for (int v = 0; v < NV; v++) u[v] = c[v][0] + half * c[v][1 + a] + o1 * c[v][1 + t1] + o2 * c[v][1 + t2];
                Prim<double, NC> pi, po;
                cons_to_prim<double, NC, PH>(P, u, pi);
                bnd_state<double, NC, PH>(P, -1 - nb, a, sg, pi, po, -1);
                sum += (area * (1.0 / 6.0)) * (0.5 * (pi.r + po.r));
            }
        } else {
            const int4 grp = A.aux[-kAuxBase - nb];
            const int q4[4] = {grp.x, grp.y, grp.z, grp.w};
            sum = 0.0;
            #pragma unroll
            for (int j = 0; j < 4; j++) {
                double q[4];
                rho_coeffs<S, NC>(U, n, q4[j], q);
                const double rsj = rs + ((j & 1) ? 0.25 : -0.25) * r[1 + t1] + ((j & 2) ? 0.25 : -0.25) * r[1 + t2];
                sum += (0.25 * area) * (0.5 * (rsj + (q[0] - half * q[1 + a])));
            }
        }
        g[a] += sg * sum;
    }
    const double vol = area * h;
    double diff2 = 0.0;
    #pragma unroll
    for (int i = 0; i < 3; i++) { g[i] /= vol; diff2 += g[i] * g[i]; }
    A.err2[e] = diff2 * (1.0 / 12.0) * ldexp(1.0, -5 * lev);     // V^(5/3) = 2^(-5 level)
    A.rho_c[e] = r[0] < kEps ? kEps : r[0];
}

// Restriction of a P1 Taylor field of leaf (s, ls) to leaf (q, lq) inside it:
// c0' = c0 + sum_d c_d delta_d, c_d' = c_d h_q / h_s, delta = (x_q - x_s) / h_s
// between the centres.  This is the L2 projection of charm_replace_quads'
// refinement branch (src/charm_amr.c:258-306), which is exact for P1.
__device__ __forceinline__ void restrict_p1(double c[4], int4 s, int4 q)
{
    const int hs = kRootLen >> s.w, hq = kRootLen >> q.w;
    const double inv = 1.0 / (double) (2 * hs);
    const double d0 = (double) ((2 * q.x + hq) - (2 * s.x + hs)) * inv;
    const double d1 = (double) ((2 * q.y + hq) - (2 * s.y + hs)) * inv;
    const double d2 = (double) ((2 * q.z + hq) - (2 * s.z + hs)) * inv;
    const double ratio = ldexp(1.0, s.w - q.w);
    c[0] = c[0] + c[1] * d0 + c[2] * d1 + c[3] * d2;
    c[1] *= ratio; c[2] *= ratio; c[3] *= ratio;
}

// One thread per new leaf.  Unchanged leaves copy their coefficients; a leaf
// inside an old leaf restricts it; a leaf from a coarsened family first forms
// the parent's L2 projection of the eight children (charm_replace_quads'
// coarsening branch, src/charm_amr.c:219-256):
//   c0_P = (1/8) sum_c c0_c,  c_d,P = sum_c [ (3/8) sigma_d,c c0_c + c_d,c / 16 ],
// sigma_d,c = +-1 the side of child c along d, and restricts that if the
// balance refined the parent again.
template <typename S>
__global__ __launch_bounds__(256)
void k_amr_project(ProjArgs A)
{
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n_new) return;
    const S *__restrict__ Uo = reinterpret_cast<const S *>(A.Uold);
    S *__restrict__ Un = reinterpret_cast<S *>(A.Unew);
    const int64_t no = A.n_old, nn = A.n_new;
    const int nvar = A.nvk / 4;
    const int4 q = A.new_q[i];
    const int64_t m = A.map[i];
    if (m >= 0) {
        const int4 s = A.old_q[m];
        if (s.w == q.w) {
            for (int c = 0; c < A.nvk; c++) (Un + c * nn)[i] = (Uo + c * no)[m];
            return;
        }
        for (int v = 0; v < nvar; v++) {
            double c[4];
            #pragma unroll
            for (int k = 0; k < 4; k++) c[k] = (double) (Uo + (4 * v + k) * no)[m];
            restrict_p1(c, s, q);
            #pragma unroll
            for (int k = 0; k < 4; k++) (Un + (4 * v + k) * nn)[i] = (S) c[k];
        }
        return;
    }
    const int64_t f = -1 - m;
    const int4 c0q = A.old_q[f];
    const int4 par = make_int4(c0q.x, c0q.y, c0q.z, c0q.w - 1);
    for (int v = 0; v < nvar; v++) {
        double c[4] = {0.0, 0.0, 0.0, 0.0};
        for (int ch = 0; ch < 8; ch++) {
            const int4 cq = A.old_q[f + ch];
            const double u0 = (double) (Uo + 4 * v * no)[f + ch];
            const double s0 = (cq.x != par.x) ? 0.375 : -0.375;
            const double s1 = (cq.y != par.y) ? 0.375 : -0.375;
            const double s2 = (cq.z != par.z) ? 0.375 : -0.375;
            c[0] += u0;
            c[1] += s0 * u0 + (double) (Uo + (4 * v + 1) * no)[f + ch] * 0.0625;
            c[2] += s1 * u0 + (double) (Uo + (4 * v + 2) * no)[f + ch] * 0.0625;
            c[3] += s2 * u0 + (double) (Uo + (4 * v + 3) * no)[f + ch] * 0.0625;
        }
        c[0] *= 0.125;
        if (q.w != par.w) restrict_p1(c, par, q);
        #pragma unroll
        for (int k = 0; k < 4; k++) (Un + (4 * v + k) * nn)[i] = (S) c[k];
    }
}

template <typename S, int PH>
static hipError_t launch_amr_indicator_r(int nc, const AmrArgs &A, const DevParams &P, hipStream_t st)
{
    const dim3 grid((unsigned) ((A.n + 255) / 256)), block(256);
    switch (nc) {
    case 1: hipLaunchKernelGGL((k_amr_indicator<S, 1, PH>), grid, block, 0, st, A, P); break;
    case 2: hipLaunchKernelGGL((k_amr_indicator<S, 2, PH>), grid, block, 0, st, A, P); break;
    case 3: hipLaunchKernelGGL((k_amr_indicator<S, 3, PH>), grid, block, 0, st, A, P); break;
    case 4: hipLaunchKernelGGL((k_amr_indicator<S, 4, PH>), grid, block, 0, st, A, P); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_amr_indicator(int prec, int nc, int ph, const AmrArgs &A, const DevParams &P, hipStream_t st)
{
    if (A.n <= 0) return hipSuccess;
    if (ph) return prec ? launch_amr_indicator_r<float, 1>(nc, A, P, st) : launch_amr_indicator_r<double, 1>(nc, A, P, st);
    return prec ? launch_amr_indicator_r<float, 0>(nc, A, P, st) : launch_amr_indicator_r<double, 0>(nc, A, P, st);
}

hipError_t launch_amr_project(int prec, const ProjArgs &A, hipStream_t st)
{
    if (A.n_new <= 0) return hipSuccess;
    const dim3 grid((unsigned) ((A.n_new + 255) / 256)), block(256);
    if (prec) hipLaunchKernelGGL(k_amr_project<float>, grid, block, 0, st, A);
    else hipLaunchKernelGGL(k_amr_project<double>, grid, block, 0, st, A);
    return hipGetLastError();
}

} // namespace charm
