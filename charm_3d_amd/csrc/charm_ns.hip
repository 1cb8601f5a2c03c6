// charm_ns.hip -- gfx950 kernels of the Navier-Stokes terms: chemistry,
// multicomponent diffusion coefficients, the BR1 auxiliary stress / heat
// flux and the viscous integrals.
//
// The viscous kernels are element-centric: one thread per leaf walks its six
// face slots (conforming, hanging in either direction, boundary) through the
// neighbour table and evaluates the face points itself, so each interior
// face point is evaluated from both sides (no atomics, no face records).
// Their output feeds the stage kernel: the auxiliary field AUX[(9*4) E]
// (tau xx yy zz xy xz yz, q x y z; Taylor coefficient-major) and the viscous
// integrals VI[(4*4) E] of ru rv rw re, which the stage adds before A^-1.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "charm_device.hpp"
#include "charm_kernels.h"
#include "charm_elemfaces.hpp"

namespace charm {

namespace {

// mole-fraction weighted viscosity and mass-fraction weighted conductivity
// at the state's temperature T = e / cv (charm_get_visc_mu / charm_get_heat_k,
// src/charm_globals.c:515-561; Sutherland's law, :652-676)
template <typename R, int NC>
__device__ __forceinline__ void transport(const DevParams &P, const Prim<R, NC> &p, R &mu, R &kt)
{
    const double t = (double) p.e / (double) p.cv;
    double m = 0.0, s = 0.0, k = 0.0;
    #pragma unroll
    for (int i = 0; i < NC; i++) {
        const double cm = (double) p.c[i] * P.comp_minv[i];
        double ml = P.ml0[i], kp = P.kp0[i];
        if (P.ml_type[i] == 1 || P.kp_type[i] == 1) {
            // one Sutherland factor for both (same T0, S per species)
            const double r = t * P.t0inv[i];
            const double f = sqrt(r * r * r) * (P.t0[i] + P.ts[i]) / (t + P.ts[i]);
            if (P.ml_type[i] == 1) ml = P.ml0[i] * f;
            if (P.kp_type[i] == 1) kp = P.kp0[i] * f;
        }
        s += cm;
        m += cm * ml;
        k += (double) p.c[i] * kp;
    }
    mu = (R) (m / s);
    kt = (R) k;
}

// the transport of a point state: the point pass's mu and k (planes 3 and
// 4 of its row tk, qs apart) when tk is set, else evaluated
template <typename R, int NC>
__device__ __forceinline__ void transport_at(const DevParams &P, const Prim<R, NC> &p, const R *tk, int64_t qs,
                                             R &mu, R &kt)
{
    if (tk) { mu = tk[3 * qs]; kt = tk[4 * qs]; }
    else transport<R, NC>(P, p, mu, kt);
}

// enthalpy h0 + sum_k cp_k (T^(k+1) - T_ref^(k+1)) / (k+1),
// charm_comp_calc_enthalpy (src/charm_globals.c:679-694), as the polynomial
// hH0 + T (hc0 + T (hc1 + ...)) with hc_k = cp_k / (k+1) and the T_ref terms
// folded into hH0 on the host (the coefficients beyond n_cp are zero)
__device__ __forceinline__ double enthalpy(const DevParams &P, int i, double t)
{
    double h = P.hc[i][7];
    #pragma unroll
    for (int k = 6; k >= 0; k--) h = fma(h, t, P.hc[i][k]);
    return fma(h, t, P.hH0[i]);
}

// tk: the point's EOS closure from the point pass (planes kMemoRows * n
// apart), or null
template <typename R, int NC>
__device__ __forceinline__ void state_at(const DevParams &P, const R *u, const R ph[3], Prim<R, NC> &p,
                                         const R *tk = nullptr, int64_t qs = 0)
{
    constexpr int NV = 4 + NC;
    R t[NV];
    #pragma unroll
    for (int v = 0; v < NV; v++) t[v] = at<R>(&u[4 * v], ph[0], ph[1], ph[2]);
    cons_to_prim<R, NC, 1>(P, t, p, tk, qs);
}

// the point pass's row at point k of (sub)face slot 4 d + j of leaf e, or
// at volume point g (null when the pass did not run on this state)
template <typename R>
__device__ __forceinline__ const R *tpt_at(const NsArgs &A, int slot, int k, int e)
{
    return A.tpt ? reinterpret_cast<const R *>(A.tpt) + ((int64_t) (6 * slot + k) * A.n + e) : nullptr;
}

template <typename R>
__device__ __forceinline__ const R *tvol_at(const NsArgs &A, int g, int e)
{
    return A.tpt ? reinterpret_cast<const R *>(A.tpt) + ((int64_t) (kMemoVolRow + g) * A.n + e) : nullptr;
}

// velocity of the P1 state at reference point ph, formed as cons_to_prim
// forms it (the momenta over rho = sum of the partial densities)
template <typename R, int NC>
__device__ __forceinline__ void velocity_at(const R *u, const R ph[3], R vel[3])
{
    R r = at<R>(&u[16], ph[0], ph[1], ph[2]);
    #pragma unroll
    for (int i = 1; i < NC; i++) r += at<R>(&u[4 * (4 + i)], ph[0], ph[1], ph[2]);
    const R ri = rcp(r);
    #pragma unroll
    for (int d = 0; d < 3; d++) vel[d] = at<R>(&u[4 * d], ph[0], ph[1], ph[2]) * ri;
}

// face terms of the gradient pass at one point
// (src/models/ns/charm_model_ns_diff_grad.c:279-298): q6 = the stress of the
// velocity (fu, fv, fw) along nn, qq = (k T + sum_i h_i(T) rho D_i Y_i) nn, at
// the coefficients lp, lm, mu, kt, dm
template <typename R, int NC>
__device__ __forceinline__ void grad_terms(const DevParams &P, R lp, R lm, R mu, R kt, R fu, R fv, R fw, R ft,
                                           R fr, const R *fc, const double *dm, const R nn[3], R q6[6], R qq[3])
{
    q6[0] = lp * fu * nn[0] + lm * fv * nn[1] + lm * fw * nn[2];
    q6[1] = lm * fu * nn[0] + lp * fv * nn[1] + lm * fw * nn[2];
    q6[2] = lm * fu * nn[0] + lm * fv * nn[1] + lp * fw * nn[2];
    q6[3] = mu * (fu * nn[1] + fv * nn[0]);
    q6[4] = mu * (fw * nn[0] + fu * nn[2]);
    q6[5] = mu * (fw * nn[1] + fv * nn[2]);
    double sp = 0.0;
    #pragma unroll
    for (int i = 0; i < NC; i++) sp += enthalpy(P, i, (double) ft) * (double) fr * dm[i] * (double) fc[i];
    #pragma unroll
    for (int d = 0; d < 3; d++) qq[d] = kt * ft * nn[d] + (R) sp * nn[d];
}

} // namespace

// Chemistry on the cell means (_charm_model_ns_chem_iter_fn,
// src/models/ns/charm_model_ns_chem.c:111-156), FP64: molar concentrations
// C_i = rho Y_i / M_i, two semi-implicit passes
// C^ = (C + dt psi (1 + dt phi / 2)) / (1 + dt phi + (dt phi)^2 / 2) with the
// rates at C_ = (C^ + C) / 2, the energy source sum_i w_i h_i(T) (:66-108),
// then rc_i <- (C^_i M_i, 0, 0, 0).  Other variables pass through.
template <typename S, int NC>
__global__ __launch_bounds__(256)
void k_ns_chem(NsArgs A, DevParams)
{
    const DevParams &P = *A.prm;         // reaction tables indexed at run time
    constexpr int NV = 4 + NC;
    const int64_t n = A.n;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const S *__restrict__ U = reinterpret_cast<const S *>(A.U);
    S *__restrict__ out = reinterpret_cast<S *>(A.out);
    double u0[NV];
    #pragma unroll
    for (int v = 0; v < NV; v++) u0[v] = (double) (U + 4 * v * n)[e];
    Prim<double, NC> p;
    cons_to_prim<double, NC, 1>(P, u0, p);
    const double t = p.e / p.cv, dt = A.dt;
    double c[NC], ch[NC], c_[NC], k[4], w[4];
    #pragma unroll
    for (int i = 0; i < NC; i++) { c[i] = u0[4 + i] / P.comp_m[i]; ch[i] = c[i]; }
    for (int r = 0; r < P.n_react; r++) k[r] = P.r_a[r] * pow(t, P.r_n[r]) * exp(-P.r_e[r] / t / kGasR);
    for (int it = 0; it < 3; it++) {
        #pragma unroll
        for (int i = 0; i < NC; i++) c_[i] = (it < 2) ? (ch[i] + c[i]) * 0.5 : ch[i];
        for (int r = 0; r < P.n_react; r++) {
            w[r] = k[r];
            #pragma unroll
            for (int m = 0; m < 3; m++) {
                const int i = P.r_left[r][m];
                #pragma unroll
                for (int q = 0; q < NC; q++) if (i == q) w[r] *= c_[q];
            }
        }
        if (it < 2) {
            #pragma unroll
            for (int i = 0; i < NC; i++) {
                double phi = 0.0, psi = 0.0;
                for (int r = 0; r < P.n_react; r++)
                    for (int m = 0; m < 3; m++) {
                        if (P.r_left[r][m] == i) {
                            double ww = k[r];
                            for (int j = 1; j < 3; j++) {
                                const int ic = P.r_left[r][(m + j) % 3];
                                #pragma unroll
                                for (int q = 0; q < NC; q++) if (ic == q) ww *= c_[q];
                            }
                            phi += ww;
                        }
                        if (P.r_right[r][m] == i) psi += w[r];
                    }
                const double x = dt * phi;
                ch[i] = (c[i] + dt * psi * (1.0 + x * 0.5)) / (1.0 + x + x * x * 0.5);
            }
        } else {
            double rhs = 0.0;
            #pragma unroll
            for (int i = 0; i < NC; i++) {
                double ws = 0.0;
                for (int r = 0; r < P.n_react; r++)
                    for (int m = 0; m < 3; m++) {
                        if (P.r_left[r][m] == i) {
                            // the reference subtracts chem_w[species] (index quirk); past the
                            // reaction array it reads memory it does not own: 0 here
                            if (P.chem_parity) ws -= (i < P.n_react) ? w[i < 4 ? i : 0] : 0.0;
                            else ws -= w[r];
                        }
                        if (P.r_right[r][m] == i) ws += w[r];
                    }
                // the reference's molar rate times mass enthalpy, or (corrected
                // units) times the molar enthalpy M_i h_i
                rhs += ws * (P.chem_energy_molar ? P.comp_m[i] : 1.0) * enthalpy(P, i, t);
            }
            A.chem_rhs[e] = rhs;
        }
    }
    for (int v = 0; v < 4; v++)
        #pragma unroll
        for (int kk = 0; kk < 4; kk++) (out + (4 * v + kk) * n)[e] = (U + (4 * v + kk) * n)[e];
    #pragma unroll
    for (int i = 0; i < NC; i++) {
        (out + (4 * (4 + i)) * n)[e] = (S) (ch[i] * P.comp_m[i]);
        #pragma unroll
        for (int kk = 1; kk < 4; kk++) (out + (4 * (4 + i) + kk) * n)[e] = S(0);
    }
}

// multicomponent diffusion coefficients at the leaf centre
// (src/models/ns/charm_model_ns_diff.c:18-71), FP64: Chapman-Enskog binary
// coefficients with the fitted collision integral, Wilke's mixture rule
template <typename S, int NC>
__global__ __launch_bounds__(256)
void k_ns_diffusion(NsArgs A, DevParams)
{
    const DevParams &P = *A.prm;
    constexpr int NV = 4 + NC;
    const int64_t n = A.n;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const S *__restrict__ U = reinterpret_cast<const S *>(A.U);
    double u0[NV];
    #pragma unroll
    for (int v = 0; v < NV; v++) u0[v] = (double) (U + 4 * v * n)[e];
    Prim<double, NC> p;
    cons_to_prim<double, NC, 1>(P, u0, p);
    const double t = p.e / p.cv, pabs = p.p / 101325.0;
    double x[NC], sm = 0.0;
    #pragma unroll
    for (int i = 0; i < NC; i++) { x[i] = p.c[i] / P.comp_m[i]; sm += x[i]; }
    #pragma unroll
    for (int i = 0; i < NC; i++) x[i] /= sm;
    #pragma unroll
    for (int i = 0; i < NC; i++) {
        double di = 0.0;
        if (!(fabs(p.c[i] - 1.0) <= kEps)) {
            double s_xd = 0.0;
            #pragma unroll
            for (int j = 0; j < NC; j++) {
                const double td = t / sqrt(P.ek[i] * P.ek[j]);
                const double wd = 1.06036 / pow(td, 0.1561) + 0.1930 / exp(0.47635 * td) +
                                  1.03587 / exp(1.52996 * td) + 1.76474 / exp(3.89411 * td);
                const double sg = 0.5 * (P.sig[i] + P.sig[j]);
                const double dij = 2.628e-7 * sqrt(0.5 * (1.0 / (1000.0 * P.comp_m[i]) + 1.0 / (1000.0 * P.comp_m[j]))) *
                                   sqrt(t * t * t) / (pabs * (sg * sg) * wd);
                s_xd += (i == j) ? 0.0 : x[j] / dij;
            }
            di = (1.0 - x[i]) / s_xd;
        }
        A.dcoef[i * n + e] = di;
    }
}

// BR1 auxiliary stress and heat flux (src/models/ns/charm_model_ns_diff_grad.c:
// 21-507): int = -int f(u) . grad phi + sum_faces int avg(f) . n phi, then
// A^-1 (diagonal on the cube octree).  tau = (lambda + 4 mu / 3, lambda - 2 mu
// / 3, mu) applied to the velocity (lambda = 0), q = -(k T + sum_i h_i rho D_i
// Y_i) with the sign folded as in the reference.
// BR2 (the option; no counterpart in the reference) also writes, for each
// interior (sub)face slot of the leaf (d * 4 + j: face direction d, sub-face
// j of a face with four small neighbours, else 0), the leaf's lifting
//   l = A^-1 sum_points sgn [Q(avg) - Q(own)] phi wJ
// in 7 fields (tau xx yy zz xy xz yz, q along the face axis: the lifting's
// other q components are zero), and the own-trace auxiliary field
// tau_own = tau_h - sum of the leaf's liftings.
template <typename S, int NC, bool BR2>
__global__ __launch_bounds__(128, (sizeof(S) == 4 ? 2 : 1))
void k_ns_grad(NsArgs A, DevParams)
{
    const DevParams &P = *A.prm;         // power series and BC records indexed at run time
    using R = S;
    constexpr int NV = 4 + NC;
    const int64_t n = A.n;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const R *__restrict__ U = reinterpret_cast<const R *>(A.U);
    const int lev = A.level[e];
    const R h = ldexp(R(1), -lev), hinv = ldexp(R(1), lev);
    R u[4 * NV];
    load_coeffs<R, NV>(U, n, e, u);
    double ds[NC];
    #pragma unroll
    for (int i = 0; i < NC; i++) ds[i] = A.dcoef ? A.dcoef[i * n + e] : 0.0;
    R acc[9][4];
    #pragma unroll
    for (int f = 0; f < 9; f++)
        #pragma unroll
        for (int k = 0; k < 4; k++) acc[f][k] = R(0);
    // volume: grad phi_{1+d} = e_d / h, wJ = h^3 / 8
    const R gph = R(0.28867513459481288225);
    const R vw = h * h * h * R(0.125) * hinv;
    for (int g = 0; g < 8; g++) {
        const R ph[3] = {(g & 1) ? gph : -gph, (g & 2) ? gph : -gph, (g & 4) ? gph : -gph};
        Prim<R, NC> p;
        const R *tg = tvol_at<R>(A, g, e);
        state_at<R, NC>(P, u, ph, p, tg, kMemoRows * n);
        R mu, kt;
        transport_at<R, NC>(P, p, tg, kMemoRows * n, mu, kt);
        const R lp = R(4.0 / 3.0) * mu, lm = R(-2.0 / 3.0) * mu;
        const R vel[3] = {p.u, p.v, p.w};
        const double t = (double) p.e / (double) p.cv;
        double psi = (double) kt * t;
        #pragma unroll
        for (int i = 0; i < NC; i++) psi += enthalpy(P, i, t) * (double) p.r * ds[i] * (double) p.c[i];
        #pragma unroll
        for (int d = 0; d < 3; d++) {
            // phi_x etc. of basis 1+d: only component d is nonzero
            acc[0][1 + d] -= (d == 0 ? lp * vel[0] : lm * vel[d]) * vw;
            acc[1][1 + d] -= (d == 1 ? lp * vel[1] : lm * vel[d]) * vw;
            acc[2][1 + d] -= (d == 2 ? lp * vel[2] : lm * vel[d]) * vw;
        }
        acc[3][1] -= mu * p.v * vw; acc[3][2] -= mu * p.u * vw;      // xy: u phi_y + v phi_x
        acc[4][1] -= mu * p.w * vw; acc[4][3] -= mu * p.u * vw;      // xz: w phi_x + u phi_z
        acc[5][2] -= mu * p.w * vw; acc[5][3] -= mu * p.v * vw;      // yz: w phi_y + v phi_z
        acc[6][1] += (R) psi * vw; acc[7][2] += (R) psi * vw; acc[8][3] += (R) psi * vw;
    }
    // face terms from the averaged state: q6 = tau(avg u) . n, qq = k T n + ...
    auto add = [&](const R q6[6], const R qq[3], const R ph[3], R wj) {
        const R phi[4] = {R(1), ph[0], ph[1], ph[2]};
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            const R b = phi[k] * wj;
            #pragma unroll
            for (int f = 0; f < 6; f++) acc[f][k] += q6[f] * b;
            #pragma unroll
            for (int f = 0; f < 3; f++) acc[6 + f][k] -= qq[f] * b;
        }
    };
    auto terms = [&](R lp, R lm, R mu, R kt, R fu, R fv, R fw, R ft, R fr, const R *fc, const double *dm,
                     const R nn[3], R q6[6], R qq[3]) {
        grad_terms<R, NC>(P, lp, lm, mu, kt, fu, fv, fw, ft, fr, fc, dm, nn, q6, qq);
    };
    // one interior face point: the BR1 term into acc; with BR2 also the
    // lifting integrand [Q(avg) - Q(own)] sgn phi wJ into L
    auto inner = [&](int d, const R *un, int nb, const R ps[3], const R pn[3], R wj, R sign, R (*L)[4], int k,
                     int js, int jn) {
        Prim<R, NC> pa, pb;
        const R *ta_ = tpt_at<R>(A, 4 * d + js, k, e), *tb_ = tpt_at<R>(A, 4 * (d ^ 1) + jn, k, nb);
        state_at<R, NC>(P, u, ps, pa, ta_, kMemoRows * n);
        state_at<R, NC>(P, un, pn, pb, tb_, kMemoRows * n);
        R mua, kta, mub, ktb;
        transport_at<R, NC>(P, pa, ta_, kMemoRows * n, mua, kta);
        transport_at<R, NC>(P, pb, tb_, kMemoRows * n, mub, ktb);
        const R fmu = R(0.5) * (mua + mub), kt = R(0.5) * (kta + ktb);
        const R lp = R(4.0 / 3.0) * fmu, lm = R(-2.0 / 3.0) * fmu;
        const R ta = dv(pa.e, pa.cv);
        const R ft = R(0.5) * (ta + dv(pb.e, pb.cv));
        R fc[NC];
        double dm[NC];
        #pragma unroll
        for (int i = 0; i < NC; i++) {
            fc[i] = R(0.5) * (pa.c[i] + pb.c[i]);
            dm[i] = 0.5 * (ds[i] + (A.dcoef ? A.dcoef[i * n + nb] : 0.0));
        }
        // the face normal is +e_a for both sides; e adds with its side's sign
        const int a = d >> 1;
        const R nn[3] = {a == 0 ? R(1) : R(0), a == 1 ? R(1) : R(0), a == 2 ? R(1) : R(0)};
        R q6[6], qq[3];
        terms(lp, lm, fmu, kt, R(0.5) * (pa.u + pb.u), R(0.5) * (pa.v + pb.v), R(0.5) * (pa.w + pb.w),
              ft, R(0.5) * (pa.r + pb.r), fc, dm, nn, q6, qq);
        add(q6, qq, ps, sign * wj);
        if constexpr (BR2) {
            R q6o[6], qqo[3];
            terms(lp, lm, fmu, kt, pa.u, pa.v, pa.w, ta, pa.r, pa.c, dm, nn, q6o, qqo);
            const R dq = pick3(a, qq[0], qq[1], qq[2]) - pick3(a, qqo[0], qqo[1], qqo[2]);
            const R phi[4] = {R(1), ps[0], ps[1], ps[2]};
            #pragma unroll
            for (int c = 0; c < 4; c++) {
                const R b = phi[c] * sign * wj;
                #pragma unroll
                for (int f = 0; f < 6; f++) L[f][c] += (q6[f] - q6o[f]) * b;
                L[6][c] -= dq * b;
            }
        }
    };
    auto bnd = [&](int d, int face, int k, const R ps[3], R wj) {
            const int a = d >> 1;
            const R sg = (d & 1) ? R(1) : R(-1);
            Prim<R, NC> pa, pb;
            const R *tk = tpt_at<R>(A, 4 * d, k, e);
            state_at<R, NC>(P, u, ps, pa, tk, kMemoRows * n);
            bnd_state<R, NC, 1>(P, face, a, sg, pa, pb, -1);
            R mu, kt;
            transport_at<R, NC>(P, pa, tk, kMemoRows * n, mu, kt);
            const R lp = R(4.0 / 3.0) * mu, lm = R(-2.0 / 3.0) * mu;
            R fu, fv, fw, ft;
            if (P.bnd_type[face] == BND_WALL_NO_SLIP) {
                fu = fv = fw = R(0);
                ft = cst<R>(P.bnd_param[face][0], P.bnd_param_f[face][0]);
            } else {
                fu = R(0.5) * (pa.u + pb.u); fv = R(0.5) * (pa.v + pb.v); fw = R(0.5) * (pa.w + pb.w);
                ft = R(0.5) * (dv(pa.e, pa.cv) + dv(pb.e, pb.cv));
            }
            R fc[NC];
            #pragma unroll
            for (int i = 0; i < NC; i++) fc[i] = R(0.5) * (pa.c[i] + pb.c[i]);
            const R nn[3] = {a == 0 ? sg : R(0), a == 1 ? sg : R(0), a == 2 ? sg : R(0)};
            R q6[6], qq[3];
            terms(lp, lm, mu, kt, fu, fv, fw, ft, R(0.5) * (pa.r + pb.r), fc, ds, nn, q6, qq);
            add(q6, qq, ps, wj);
    };
    R ai[4];
    #pragma unroll
    for (int k = 0; k < 4; k++) ai[k] = cst<R>(P.ainv0[k], P.ainv0_f[k]) * ldexp(R(1), 3 * lev);
    R *__restrict__ lift = reinterpret_cast<R *>(A.lift);
    // boundary points, then each interior (sub)face with the neighbour's
    // coefficients loaded once for its six points
    face_points<R>(A.nbr, A.aux, A.level, e, h, [&](int, int, const R *, const R *, R, R, int, int, int) {}, bnd);
    face_subfaces<R>(A.nbr, A.aux, A.level, e, h,
        [&](int d, int nb, R so1, R so2, R sm, R no1, R no2, R nm, R wj, R sign, int js, int jn) {
            R un[4 * NV];
            load_coeffs<R, NV>(U, n, nb, un);
            R L[7][4];
            #pragma unroll
            for (int f = 0; f < 7; f++)
                #pragma unroll
                for (int c = 0; c < 4; c++) L[f][c] = R(0);
            for (int k = 0; k < 6; k++) {
                R ps[3], pn[3];
                face_point<R>(d, k, so1, so2, sm, ps);
                face_point<R>(d ^ 1, k, no1, no2, nm, pn);
                inner(d, un, nb, ps, pn, wj, sign, L, k, js, jn);
            }
            if constexpr (BR2) {
                const int slot = 4 * d + js;
                #pragma unroll
                for (int f = 0; f < 7; f++)
                    #pragma unroll
                    for (int c = 0; c < 4; c++)
                        (lift + ((size_t) kLiftFields * slot + 4 * f + c) * n)[e] = L[f][c] * ai[c];
            }
        });
    R *__restrict__ aux = reinterpret_cast<R *>(A.aux_out);
    #pragma unroll
    for (int k = 0; k < 4; k++)
        #pragma unroll
        for (int f = 0; f < 9; f++) {
            acc[f][k] *= ai[k];
            (aux + (4 * f + k) * n)[e] = acc[f][k];
        }
    if constexpr (BR2) {
        // tau_own = tau_h less the liftings this thread has just written
        for (int d = 0; d < 6; d++) {
            const int nb = A.nbr[6 * e + d];
            const int ns = nb >= 0 ? 1 : (nb <= -kNbrAuxBase ? 4 : 0);
            const int a = d >> 1;
            for (int j = 0; j < ns; j++) {
                const R *l = lift + (size_t) kLiftFields * (4 * d + j) * n + e;
                #pragma unroll
                for (int c = 0; c < 4; c++) {
                    #pragma unroll
                    for (int f = 0; f < 6; f++) acc[f][c] -= l[(4 * f + c) * n];
                    const R lq = l[(24 + c) * n];
                    acc[6][c] -= a == 0 ? lq : R(0);
                    acc[7][c] -= a == 1 ? lq : R(0);
                    acc[8][c] -= a == 2 ? lq : R(0);
                }
            }
        }
        R *__restrict__ own = reinterpret_cast<R *>(A.aux_own);
        #pragma unroll
        for (int k = 0; k < 4; k++)
            #pragma unroll
            for (int f = 0; f < 9; f++) (own + (4 * f + k) * n)[e] = acc[f][k];
    }
}

// Viscous integrals of ru rv rw re (src/models/ns/charm_model_ns_diff_integrals.c:
// 21-408): + int tau . grad phi (+ (tau u) . grad phi + chem_rhs phi for the
// energy) - sum_faces int (avg(tau) . n, avg(tau u) . n - avg(q) . n) phi;
// boundary faces contribute only on no-slip walls, one point at the face
// centre with the cell-mean velocity: -(-mu v_t / |l|, -q . n).
// BR2 (no counterpart in the reference): on an interior (sub)face each
// side's stress and heat flux are its own-trace auxiliary field plus eta
// times the (sub)face's lifting
//   l_s = A_s^-1 sum_points sgn_s [Q(avg) - Q(own_s)] phi_s wJ,
// Q the gradient pass's face term at the face-averaged coefficients, from
// the averaged and from side s's own trace; the volume and boundary terms
// are BR1's.  Of q only the normal component enters the flux.
template <typename S, int NC, bool BR2>
__global__ __launch_bounds__(128)
void k_ns_visc(NsArgs A, DevParams)
{
    const DevParams &P = *A.prm;
    using R = S;
    constexpr int NV = 4 + NC;
    const int64_t n = A.n;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const R *__restrict__ U = reinterpret_cast<const R *>(A.U);
    const R *__restrict__ X = reinterpret_cast<const R *>(A.aux_in);
    const int lev = A.level[e];
    const R h = ldexp(R(1), -lev), hinv = ldexp(R(1), lev);
    R u[4 * NV], x[36];
    load_coeffs<R, NV>(U, n, e, u);
    #pragma unroll
    for (int i = 0; i < 36; i++) x[i] = (X + i * n)[e];
    const R chem = A.chem_rhs ? (R) A.chem_rhs[e] : R(0);
    R acc[4][4];
    #pragma unroll
    for (int v = 0; v < 4; v++)
        #pragma unroll
        for (int k = 0; k < 4; k++) acc[v][k] = R(0);
    const R gph = R(0.28867513459481288225);
    const R wjv = h * h * h * R(0.125), vw = wjv * hinv;
    for (int g = 0; g < 8; g++) {
        const R ph[3] = {(g & 1) ? gph : -gph, (g & 2) ? gph : -gph, (g & 4) ? gph : -gph};
        Prim<R, NC> p;
        state_at<R, NC>(P, u, ph, p, tvol_at<R>(A, g, e), kMemoRows * n);
        R t[6];
        #pragma unroll
        for (int f = 0; f < 6; f++) t[f] = at<R>(&x[4 * f], ph[0], ph[1], ph[2]);
        // rows of tau: x (xx xy xz), y (xy yy yz), z (xz yz zz)
        const R row[3][3] = {{t[0], t[3], t[4]}, {t[3], t[1], t[5]}, {t[4], t[5], t[2]}};
        #pragma unroll
        for (int d = 0; d < 3; d++) {
            acc[0][1 + d] += row[0][d] * vw;
            acc[1][1 + d] += row[1][d] * vw;
            acc[2][1 + d] += row[2][d] * vw;
            acc[3][1 + d] += (row[d][0] * p.u + row[d][1] * p.v + row[d][2] * p.w) * vw;
        }
        const R phi[4] = {R(1), ph[0], ph[1], ph[2]};
        #pragma unroll
        for (int k = 0; k < 4; k++) acc[3][k] += chem * phi[k] * wjv;
    }
    auto add = [&](const R q[4], const R ph[3], R wj) {
        const R phi[4] = {R(1), ph[0], ph[1], ph[2]};
        #pragma unroll
        for (int k = 0; k < 4; k++)
            #pragma unroll
            for (int v = 0; v < 4; v++) acc[v][k] -= q[v] * phi[k] * wj;
    };
    face_points<R>(A.nbr, A.aux, A.level, e, h,
        [&](int d, int nb, const R ps[3], const R pn[3], R wj, R sign, int k, int js, int jn) {
            if constexpr (BR2) return;               // the interior faces are done above
            R un[4 * NV], xn[36];
            load_coeffs<R, NV>(U, n, nb, un);
            #pragma unroll
            for (int i = 0; i < 36; i++) xn[i] = (X + i * n)[nb];
            const int a = d >> 1;
            R ftau[6], fu[3] = {R(0), R(0), R(0)}, ft = R(0);
            #pragma unroll
            for (int f = 0; f < 6; f++) ftau[f] = R(0);
            #pragma unroll
            for (int side = 0; side < 2; side++) {
                const R *uu = side ? un : u, *xx = side ? xn : x, *pp = side ? pn : ps;
                Prim<R, NC> p;
                state_at<R, NC>(P, uu, pp, p, side ? tpt_at<R>(A, 4 * (d ^ 1) + jn, k, nb) : tpt_at<R>(A, 4 * d + js, k, e),
                                 kMemoRows * n);
                R t[6];
                #pragma unroll
                for (int f = 0; f < 6; f++) { t[f] = at<R>(&xx[4 * f], pp[0], pp[1], pp[2]); ftau[f] += t[f]; }
                fu[0] += p.u * t[0] + p.v * t[3] + p.w * t[4];
                fu[1] += p.u * t[3] + p.v * t[1] + p.w * t[5];
                fu[2] += p.u * t[4] + p.v * t[5] + p.w * t[2];
                // q . n, n = +e_a: the three components evaluated and one
                // selected (a runtime index into the register array would
                // put it in private memory)
                ft += pick3(a, at<R>(&xx[24], pp[0], pp[1], pp[2]), at<R>(&xx[28], pp[0], pp[1], pp[2]),
                            at<R>(&xx[32], pp[0], pp[1], pp[2]));
            }
            // avg(tau) . n with n = +e_a: the row a of the averaged tensor
            R q[4];
            q[0] = R(0.5) * pick3(a, ftau[0], ftau[3], ftau[4]);
            q[1] = R(0.5) * pick3(a, ftau[3], ftau[1], ftau[5]);
            q[2] = R(0.5) * pick3(a, ftau[4], ftau[5], ftau[2]);
            q[3] = R(0.5) * pick3(a, fu[0], fu[1], fu[2]) - R(0.5) * ft;
            add(q, ps, sign * wj);
        },
        [&](int d, int face, int k, const R *, R) {
            // no-slip walls: one point at the face centre, weight = area, taken
            // once per face; every other boundary type has no viscous term
            if (P.bnd_type[face] != BND_WALL_NO_SLIP || k != 0) return;
            const int a = d >> 1;
            const R sg = (d & 1) ? R(1) : R(-1);
            R pc[3] = {R(0), R(0), R(0)};
            pc[a] = R(0.5) * sg;
            Prim<R, NC> pf, pm;
            state_at<R, NC>(P, u, pc, pf);
            R mu, kt;
            transport<R, NC>(P, pf, mu, kt);
            const R zero[3] = {R(0), R(0), R(0)};
            state_at<R, NC>(P, u, zero, pm);                // cell mean = centre value
            const R vv[3] = {pm.u, pm.v, pm.w};
            const R vt[3] = {a == 0 ? R(0) : vv[0], a == 1 ? R(0) : vv[1], a == 2 ? R(0) : vv[2]};
            const R ll = R(0.5) * h;
            const R qn = pick3(a, at<R>(&x[24], pc[0], pc[1], pc[2]), at<R>(&x[28], pc[0], pc[1], pc[2]),
                               at<R>(&x[32], pc[0], pc[1], pc[2])) * sg;
            R q[4];
            q[0] = -mu * dv(vt[0], ll); q[1] = -mu * dv(vt[1], ll); q[2] = -mu * dv(vt[2], ll);
            q[3] = -qn;
            add(q, pc, h * h);
        });
    if constexpr (BR2) {        // after the walls: x (tau_h) is dead here
        const R *__restrict__ XO = reinterpret_cast<const R *>(A.aux_own);
        const R *__restrict__ LF = reinterpret_cast<const R *>(A.lift);
        const R eta = cst<R>(P.br2_eta, P.br2_eta_f);
        face_subfaces<R>(A.nbr, A.aux, A.level, e, h,
            [&](int d, int nb, R so1, R so2, R sm, R no1, R no2, R nm, R wj, R sign, int js, int jn) {
                const int a = d >> 1;
                R un[4 * NV];
                load_coeffs<R, NV>(U, n, nb, un);
                // [side][field][coefficient], side 0 = e, 1 = nb; fields tau xx yy
                // zz xy xz yz and q_a: tau_own + eta l, l the side's lifting of
                // this (sub)face from the gradient pass
                R cs[2][7][4];
                const int el[2] = {e, nb};
                const int slot[2] = {4 * d + js, 4 * (d ^ 1) + jn};
                #pragma unroll
                for (int s = 0; s < 2; s++) {
                    const R *l = LF + (size_t) kLiftFields * slot[s] * n + el[s];
                    #pragma unroll
                    for (int c = 0; c < 4; c++) {
                        #pragma unroll
                        for (int f = 0; f < 6; f++)
                            cs[s][f][c] = (XO + (4 * f + c) * n)[el[s]] + eta * l[(4 * f + c) * n];
                        cs[s][6][c] = (XO + (24 + 4 * a + c) * n)[el[s]] + eta * l[(24 + c) * n];
                    }
                }
                // the BR1 surface flux with these stresses
                for (int k = 0; k < 6; k++) {
                    R ps[3], pn[3];
                    face_point<R>(d, k, so1, so2, sm, ps);
                    face_point<R>(d ^ 1, k, no1, no2, nm, pn);
                    R ftau[6], fu[3] = {R(0), R(0), R(0)}, ft = R(0);
                    #pragma unroll
                    for (int f = 0; f < 6; f++) ftau[f] = R(0);
                    #pragma unroll
                    for (int s = 0; s < 2; s++) {
                        const R *pp = s ? pn : ps;
                        R vel[3];
                        velocity_at<R, NC>(s ? un : u, pp, vel);
                        R t[6];
                        #pragma unroll
                        for (int f = 0; f < 6; f++) { t[f] = at<R>(cs[s][f], pp[0], pp[1], pp[2]); ftau[f] += t[f]; }
                        fu[0] += vel[0] * t[0] + vel[1] * t[3] + vel[2] * t[4];
                        fu[1] += vel[0] * t[3] + vel[1] * t[1] + vel[2] * t[5];
                        fu[2] += vel[0] * t[4] + vel[1] * t[5] + vel[2] * t[2];
                        ft += at<R>(cs[s][6], pp[0], pp[1], pp[2]);
                    }
                    R q[4];
                    q[0] = R(0.5) * pick3(a, ftau[0], ftau[3], ftau[4]);
                    q[1] = R(0.5) * pick3(a, ftau[3], ftau[1], ftau[5]);
                    q[2] = R(0.5) * pick3(a, ftau[4], ftau[5], ftau[2]);
                    q[3] = R(0.5) * pick3(a, fu[0], fu[1], fu[2]) - R(0.5) * ft;
                    add(q, ps, sign * wj);
                }
            });
    }
    R *__restrict__ vi = reinterpret_cast<R *>(A.visc_out);
    #pragma unroll
    for (int v = 0; v < 4; v++)
        #pragma unroll
        for (int k = 0; k < 4; k++) (vi + (4 * v + k) * n)[e] = acc[v][k];
}

// The point EOS closures of the state U: at each leaf's 8 volume points and
// at its own face points (the points face_points and face_subfaces visit: a
// (sub)face slot 4 d + j, points k = 0..5), the FP64 Newton temperature of
// _calc_t (src/charm_eos.c:12-54) and the mixture's m, cv and gamma at it
// (src/charm_eos.c:98-148), and the viscosity and conductivity there
// (src/charm_globals.c:515-561), stored in the state's precision for the
// PH = 1 closures and transport of the viscous passes and the stage kernel
// on the same state.
// Each interior face point is evaluated once per side instead of once per
// kernel and side.  One thread per leaf and part (blockIdx.y: face d = 0..5,
// or 6 the volume points), so that the long dependent Newton chains of a
// leaf run in parallel.
template <typename S, int NC>
__global__ __launch_bounds__(256)
void k_ns_temps(NsArgs A, DevParams)
{
    const DevParams &P = *A.prm;
    using R = S;
    constexpr int NV = 4 + NC;
    const int64_t n = A.n;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int part = blockIdx.y;
    const R *__restrict__ U = reinterpret_cast<const R *>(A.U);
    R u[4 * NV];
    load_coeffs<R, NV>(U, n, e, u);
    // the EOS closure (m, cv, gamma) of the point state, as cons_to_prim
    // forms it (the FP64 Newton temperature, the mixture's cp at it), and
    // its transport coefficients (mu, k), into row `row` of the five planes
    R *__restrict__ W = reinterpret_cast<R *>(A.tpt);
    const int64_t qs = (int64_t) kMemoRows * n;
    auto closure = [&](const R ph[3], int row) {
        R t[NV];
        #pragma unroll
        for (int v = 0; v < NV; v++) t[v] = at<R>(&u[4 * v], ph[0], ph[1], ph[2]);
        Prim<R, NC> p;
        cons_to_prim<R, NC, 1>(P, t, p);
        R mu, kt;
        transport<R, NC>(P, p, mu, kt);
        R *w = W + (int64_t) row * n + e;
        w[0] = p.m; w[qs] = p.cv; w[2 * qs] = p.gam; w[3 * qs] = mu; w[4 * qs] = kt;
    };
    if (part == 6) {
        const R gph = R(0.28867513459481288225);
        for (int g = 0; g < 8; g++) {
            const R ph[3] = {(g & 1) ? gph : -gph, (g & 2) ? gph : -gph, (g & 4) ? gph : -gph};
            closure(ph, kMemoVolRow + g);
        }
        return;
    }
    const int d = part;
    const int nb = A.nbr[6 * e + d];
    // the big side of a hanging face: its four sub-faces' points
    const int nsub = nb <= -kNbrAuxBase ? 4 : 1;
    for (int j = 0; j < nsub; j++) {
        const R o1 = nsub == 4 ? ((j & 1) ? R(0.25) : R(-0.25)) : R(0);
        const R o2 = nsub == 4 ? ((j & 2) ? R(0.25) : R(-0.25)) : R(0);
        const R m = nsub == 4 ? R(0.5) : R(1);
        for (int k = 0; k < 6; k++) {
            R ph[3];
            face_point<R>(d, k, o1, o2, m, ph);
            closure(ph, 6 * (4 * d + j) + k);
        }
    }
}

template <typename S>
static hipError_t launch_ns_r(int what, int nc, const NsArgs &A, const DevParams &P, hipStream_t st)
{
    const unsigned bs = (what >= 2 && what != 6) ? 128u : 256u;
    // the point pass runs a thread per leaf and part (6 faces, volume)
    const dim3 grid((unsigned) ((A.n + bs - 1) / bs), what == 6 ? 7u : 1u), block(bs);
#define CHARM_NS_CASE(NCV)                                                                         \
    case NCV:                                                                                      \
        if (what == 0) hipLaunchKernelGGL((k_ns_chem<S, NCV>), grid, block, 0, st, A, P);          \
        else if (what == 1) hipLaunchKernelGGL((k_ns_diffusion<S, NCV>), grid, block, 0, st, A, P);\
        else if (what == 2) hipLaunchKernelGGL((k_ns_grad<S, NCV, false>), grid, block, 0, st, A, P); \
        else if (what == 4) hipLaunchKernelGGL((k_ns_grad<S, NCV, true>), grid, block, 0, st, A, P);  \
        else if (what == 3) hipLaunchKernelGGL((k_ns_visc<S, NCV, false>), grid, block, 0, st, A, P); \
        else if (what == 6) hipLaunchKernelGGL((k_ns_temps<S, NCV>), grid, block, 0, st, A, P);      \
        else hipLaunchKernelGGL((k_ns_visc<S, NCV, true>), grid, block, 0, st, A, P);               \
        break;
    switch (nc) {
    CHARM_NS_CASE(1)
    CHARM_NS_CASE(2)
    CHARM_NS_CASE(3)
    CHARM_NS_CASE(4)
    default: return hipErrorInvalidValue;
    }
#undef CHARM_NS_CASE
    return hipGetLastError();
}

hipError_t launch_ns(int prec, int what, int nc, const NsArgs &A, const DevParams &P, hipStream_t st)
{
    if (A.n <= 0) return hipSuccess;
    return prec ? launch_ns_r<float>(what, nc, A, P, st) : launch_ns_r<double>(what, nc, A, P, st);
}

} // namespace charm
