// charm_adj.hip -- gfx950 kernels of the discrete adjoint of the SSP-RK3
// time step (Euler terms, constant cp, no limiter): the transposed
// linearisation of one stage S(U) = U - dt A^-1 int(U), the gradient of the
// objective Q(U) = sum over the leaves of a box of int p, and Q itself.
//
// A stage adjoint is element-centric: one thread per leaf forms
//   T = sb + (d int / d U)^T W,  W = -dt A^-1 sb,  sb = c_in X,
// from the volume term and every point of its six face slots.  At each point
// the pointwise numerics (cons_to_prim, boundary state, HLLC / LF) run on
// forward-mode dual numbers whose tangents are this leaf's point values;
// the contraction s = sum_v q_v G_v with the adjoint weights G of both sides
// then carries d s / d u(point) in its tangents.  Every interior point is
// evaluated from both sides, each side keeping its own derivative (no
// atomics).  The inlet parameters (velocity x, T, p of one boundary face)
// get their own dual pass.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "charm_device.hpp"
#include "charm_dual.hpp"
#include "charm_elemfaces.hpp"
#include "charm_vjp.hpp"
#include "charm_kernels.h"

namespace charm {

namespace {

// adjoint weights of leaf e: W[4v + k] = -dt ainv_k(level) c_in X[4v + k]
template <typename R, int NV>
__device__ __forceinline__ void adj_weights(const AdjArgs &A, const DevParams &P, const R *__restrict__ X,
                                            int e, R *w)
{
    const int lev = A.level[e];
    const R sc = ldexp(R(1), 3 * lev);
    #pragma unroll
    for (int v = 0; v < NV; v++)
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            const R ai = cst<R>(P.ainv0[k], P.ainv0_f[k]) * sc;
            w[4 * v + k] = R(-A.dt) * ai * (R(A.c_in) * (X + (4 * v + k) * A.n)[e]);
        }
}

template <typename R, int NV>
__device__ __forceinline__ R wdot(const R *w, int v, const R ph[3])
{
    return w[4 * v] + ph[0] * w[4 * v + 1] + ph[1] * w[4 * v + 2] + ph[2] * w[4 * v + 3];
}

// flux in the frame of n = sg e_a (global-order momentum components out)
template <typename D, int NC, int FLUX>
__device__ __forceinline__ void flux_global(const Prim<D, NC> &pl, const Prim<D, NC> &pr, int a, D sg, D *qg)
{
    constexpr int NV = 4 + NC;
    // the axis may differ between lanes: components are picked with selects,
    // not by indexing arrays of duals (which go to private memory)
    Frame<D, NC> fl, fr;
    fl.r = pl.r; fl.e_tot = pl.e_tot; fl.p = pl.p; fl.cz = pl.cz; fl.gam = pl.gam;
    fr.r = pr.r; fr.e_tot = pr.e_tot; fr.p = pr.p; fr.cz = pr.cz; fr.gam = pr.gam;
    fl.un = sg * sel(a == 0, pl.u, sel(a == 1, pl.v, pl.w));
    fr.un = sg * sel(a == 0, pr.u, sel(a == 1, pr.v, pr.w));
    fl.ut1 = sel(a == 0, pl.v, pl.u); fr.ut1 = sel(a == 0, pr.v, pr.u);     // t1 = 1 for a = 0, else 0
    fl.ut2 = sel(a == 2, pl.v, pl.w); fr.ut2 = sel(a == 2, pr.v, pr.w);     // t2 = 1 for a = 2, else 2
    #pragma unroll
    for (int i = 0; i < NC; i++) { fl.c[i] = pl.c[i]; fr.c[i] = pr.c[i]; }
    D q[NV];
    flux_x<D, NC, FLUX>(fl, fr, q);
    const D qn = sg * q[0];
    qg[0] = sel(a == 0, qn, q[1]);
    qg[1] = sel(a == 1, qn, sel(a == 0, q[1], q[2]));
    qg[2] = sel(a == 2, qn, q[2]);
    #pragma unroll
    for (int v = 3; v < NV; v++) qg[v] = q[v];
}

// inlet exterior state with the parameters as duals (u = theta0, T = theta1,
// p = theta2; v, w fixed), as bnd_state for BOUND_INLET (src/charm_bnd_cond.c:
// 36-47) followed by EOS flag 3
template <typename R, int NC>
__device__ __forceinline__ void inlet_dual(const DevParams &P, int face, const Prim<Dual<R, 3>, NC> &in,
                                           Prim<Dual<R, 3>, NC> &out)
{
    using D = Dual<R, 3>;
    out = in;
    out.u = D::var(cst<R>(P.bnd_param[face][0], P.bnd_param_f[face][0]), 0);
    out.v = D(cst<R>(P.bnd_param[face][1], P.bnd_param_f[face][1]));
    out.w = D(cst<R>(P.bnd_param[face][2], P.bnd_param_f[face][2]));
    const D t = D::var(cst<R>(P.bnd_param[face][3], P.bnd_param_f[face][3]), 1);
    D pr = D::var(cst<R>(P.bnd_param[face][4], P.bnd_param_f[face][4]), 2);
    eos_coeffs<D, NC>(P, out);
    if (pr < D(kEps)) pr = D(kEps);
    out.p = pr;
    out.r = dv(pr * out.m, t * D(kGasR));
    out.cz = dsqrt(dv(out.gam * pr, out.r));
    out.e = dv(pr, out.r * (out.gam - D(1)));
    out.e_tot = out.e + D(0.5) * (out.u * out.u + out.v * out.v + out.w * out.w);
}

// Volume term of a stage adjoint at leaf e (coefficients u, adjoint weights
// w): ub += (d / d u) of sum_d sum_v cw_dv F_d,v(t) - sum_{v<3} g_v wJ w(ph) rho
// over the 8 Gauss points, cw_dv = -(wJ / h) w[v][1 + d]
template <typename R, int NC>
__device__ __forceinline__ void adj_volume(const DevParams &P, const R *u, const R *w, R h, R hinv, R *ub)
{
    constexpr int NV = 4 + NC;
    const R gph = R(0.28867513459481288225);
    const R wj = h * h * h * R(0.125);
    const R f = -wj * hinv;
    #pragma unroll 1
    for (int g = 0; g < 8; g++) {
        const R ph[3] = {(g & 1) ? gph : -gph, (g & 2) ? gph : -gph, (g & 4) ? gph : -gph};
        R t[NV], tb[NV];
        #pragma unroll
        for (int v = 0; v < NV; v++) t[v] = at<R>(&u[4 * v], ph[0], ph[1], ph[2]);
        Prim<R, NC> p;
        cons_to_prim<R, NC>(P, t, p);
        PrimBar<R, NC> pb;
        pb.r = pb.u = pb.v = pb.w = pb.e_tot = pb.p = pb.cz = pb.gam = R(0);
        #pragma unroll
        for (int i = 0; i < NC; i++) pb.c[i] = R(0);
        const R vel[3] = {p.u, p.v, p.w};
        R velb[3] = {R(0), R(0), R(0)};
        #pragma unroll
        for (int d = 0; d < 3; d++) {
            R cw[NV];
            #pragma unroll
            for (int v = 0; v < NV; v++) cw[v] = f * w[4 * v + 1 + d];
            const R md = t[d];
            R td = cw[0] * p.u + cw[1] * p.v + cw[2] * p.w + cw[3] * p.e_tot;
            #pragma unroll
            for (int i = 0; i < NC; i++) { td += cw[4 + i] * p.c[i]; pb.c[i] += cw[4 + i] * md; }
            tb[d] = td;
            pb.u += cw[0] * md; pb.v += cw[1] * md; pb.w += cw[2] * md; pb.e_tot += cw[3] * md;
            pb.p += cw[d] + cw[3] * vel[d];
            velb[d] += cw[3] * p.p;
        }
        pb.u += velb[0]; pb.v += velb[1]; pb.w += velb[2];
        #pragma unroll
        for (int v = 0; v < 3; v++) pb.r -= cst<R>(P.grav[v], P.grav_f[v]) * wj * wdot<R, NV>(w, v, ph);
        tb[3] = R(0);
        #pragma unroll
        for (int v = 4; v < NV; v++) tb[v] = R(0);
        R tp[NV];
        cons_to_prim_vjp<R, NC>(P, t, p, pb, tp);
        const R phi[4] = {R(1), ph[0], ph[1], ph[2]};
        #pragma unroll
        for (int v = 0; v < NV; v++)
            #pragma unroll
            for (int k = 0; k < 4; k++) ub[4 * v + k] += phi[k] * (tb[v] + tp[v]);
    }

}

} // namespace

// Interior part of one stage adjoint: the volume term and every interior
// face point of leaf e, each through the hand-written reverse sweep (the
// primal at the point, then the flux and cons_to_prim VJPs for e's side).
// Boundary faces are added by k_adj_bnd_pts / k_adj_bnd_sum.
template <typename R, int NC, int FLUX>
__global__ __launch_bounds__(128)
void k_adj_stage(AdjArgs A, DevParams P)
{
    constexpr int NV = 4 + NC;
    const int64_t n = A.n;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const R *__restrict__ U = reinterpret_cast<const R *>(A.U);
    const R *__restrict__ X = reinterpret_cast<const R *>(A.X);
    const int lev = A.level[e];
    const R h = ldexp(R(1), -lev), hinv = ldexp(R(1), lev);
    R u[4 * NV], w[4 * NV], ub[4 * NV];
    load_coeffs<R, NV>(U, n, e, u);
    adj_weights<R, NV>(A, P, X, e, w);
    #pragma unroll
    for (int i = 0; i < 4 * NV; i++) ub[i] = R(0);

    adj_volume<R, NC>(P, u, w, h, hinv, ub);

    // interior face points: the neighbour's state and weights are loaded
    // once per neighbour; e's trace gradient at each point from one
    // reverse sweep through the flux
    auto point = [&](const R *un, const R *wn, int a, const R ps[3], const R pn[3], R wjf, R sign) {
        R ts[NV], tn[NV];
        #pragma unroll
        for (int v = 0; v < NV; v++) {
            ts[v] = at<R>(&u[4 * v], ps[0], ps[1], ps[2]);
            tn[v] = at<R>(&un[4 * v], pn[0], pn[1], pn[2]);
        }
        Prim<R, NC> pS, pN;
        cons_to_prim<R, NC>(P, ts, pS);
        cons_to_prim<R, NC>(P, tn, pN);
        Frame<R, NC> fS, fN;
        frame_of<R, NC>(pS, a, fS);
        frame_of<R, NC>(pN, a, fN);
        R G[NV];
        #pragma unroll
        for (int v = 0; v < NV; v++) G[v] = sign * wjf * (wdot<R, NV>(w, v, ps) - wdot<R, NV>(wn, v, pn));
        const int t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
        R qb[NV];
        qb[0] = G[a]; qb[1] = G[t1]; qb[2] = G[t2];
        #pragma unroll
        for (int v = 3; v < NV; v++) qb[v] = G[v];
        FrameBar<R, NC> b0, b1;
        zero_bar<R, NC>(b0);
        zero_bar<R, NC>(b1);
        // q = flux(side 0, side 1) with n = +e_a; e is side 0 when sign > 0
        if (sign > R(0)) flux_vjp<R, NC, FLUX>(fS, fN, qb, b0, b1);
        else flux_vjp<R, NC, FLUX>(fN, fS, qb, b0, b1);
        PrimBar<R, NC> pb;
        prim_bar_of<R, NC>(sign > R(0) ? b0 : b1, a, pb);
        R tb[NV];
        cons_to_prim_vjp<R, NC>(P, ts, pS, pb, tb);
        const R phi[4] = {R(1), ps[0], ps[1], ps[2]};
        #pragma unroll
        for (int v = 0; v < NV; v++)
            #pragma unroll
            for (int k = 0; k < 4; k++) ub[4 * v + k] += phi[k] * tb[v];
    };
    #pragma unroll 1
    for (int d = 0; d < 6; d++) {
        const int nb = A.nbr[6 * e + d];
        const R sign = (d & 1) ? R(1) : R(-1);
        const int a = d >> 1;
        if (nb >= 0) {
            R o1 = R(0), o2 = R(0), m = R(1);
            if (A.level[nb] != lev) {
                // e is sub-face j of the coarser neighbour's face
                const int4 grp = A.aux[-kNbrAuxBase - A.nbr[6 * nb + (d ^ 1)]];
                const int j = (grp.x == e) ? 0 : ((grp.y == e) ? 1 : ((grp.z == e) ? 2 : 3));
                o1 = (j & 1) ? R(0.25) : R(-0.25);
                o2 = (j & 2) ? R(0.25) : R(-0.25);
                m = R(0.5);
            }
            R un[4 * NV], wn[4 * NV];
            load_coeffs<R, NV>(U, n, nb, un);
            adj_weights<R, NV>(A, P, X, nb, wn);
            const R wjf = h * h * R(1.0 / 6.0);
            #pragma unroll 1
            for (int k = 0; k < 6; k++) {
                R ps[3], pn[3];
                face_point<R>(d, k, R(0), R(0), R(1), ps);
                face_point<R>(d ^ 1, k, o1, o2, m, pn);
                point(un, wn, a, ps, pn, wjf, sign);
            }
        } else if (nb <= -kNbrAuxBase) {
            const int4 grp = A.aux[-kNbrAuxBase - nb];
            const int q4[4] = {grp.x, grp.y, grp.z, grp.w};
            const R wjf = R(0.25) * h * h * R(1.0 / 6.0);
            #pragma unroll 1
            for (int j = 0; j < 4; j++) {
                R un[4 * NV], wn[4 * NV];
                load_coeffs<R, NV>(U, n, q4[j], un);
                adj_weights<R, NV>(A, P, X, q4[j], wn);
                const R o1 = (j & 1) ? R(0.25) : R(-0.25), o2 = (j & 2) ? R(0.25) : R(-0.25);
                #pragma unroll 1
                for (int k = 0; k < 6; k++) {
                    R ps[3], pn[3];
                    face_point<R>(d, k, o1, o2, R(0.5), ps);
                    face_point<R>(d ^ 1, k, R(0), R(0), R(1), pn);
                    point(un, wn, a, ps, pn, wjf, sign);
                }
            }
        }
    }

    // not __restrict__: the last stage adjoint accumulates into its own output
    R *out = reinterpret_cast<R *>(A.OUT);
    const R *acc = reinterpret_cast<const R *>(A.ACC);
    const R *acc2 = reinterpret_cast<const R *>(A.ACC2);
    #pragma unroll
    for (int i = 0; i < 4 * NV; i++) {
        const R sb = R(A.c_in) * (X + i * n)[e];
        R o = R(A.c_out) * (sb + ub[i]);
        if (acc) o += R(A.c_acc) * (acc + i * n)[e];
        if (acc2) o += R(A.c_acc2) * (acc2 + i * n)[e];
        (out + i * n)[e] = o;
    }
}

// One interior face of a chunk, both sides at once: at each of the six
// points one primal evaluation and one reverse sweep of
//   s = wJ sum_v q_v(t0, t1) (w0_v(p0) - w1_v(p1))
// (q the flux along side 0's outward normal, global component order; w_s
// the adjoint weights of side s at its point) give d s / d t0 and d s / d t1;
// out[side][m][v] = sum over the points of (d s / d t_side)_v g_m,
// g = (1, t1 coordinate, t2 coordinate) of the small face's point.
// The face's normal axis AX is a template argument, so that every
// coefficient and velocity component is picked at compile time.
template <typename R, int NC, int FLUX, int AX>
__device__ __forceinline__ void adj_face(const AdjArgs &A, const DevParams &P, const R *__restrict__ U,
                                         const R *__restrict__ X, int4 rec, R *out)
{
    constexpr int NV = 4 + NC;
    constexpr int a = AX, t1 = (AX == 0) ? 1 : 0, t2 = (AX == 2) ? 1 : 2;
    const int64_t n = A.n;
    const int info = rec.z;
    const int dir0 = info & 7, kind = (info >> 3) & 3, sub = (info >> 5) & 3;
    const R s = (dir0 & 1) ? R(1) : R(-1);
    const int lown = A.level[rec.x] + (kind == FACE_HANG1 ? 1 : 0);
    const R h = ldexp(R(1), -lown);
    const R wj = h * h * R(1.0 / 6.0);
    const R pa0 = R(0.5) * s, pa1 = -pa0;
    const R d1 = (sub & 1) ? R(0.25) : R(-0.25), d2 = (sub & 2) ? R(0.25) : R(-0.25);
    const bool big0 = kind == FACE_HANG1, big1 = kind == FACE_HANG0;
    const R m0 = big0 ? R(0.5) : R(1), m1 = big1 ? R(0.5) : R(1);
    const R o01 = big0 ? d1 : R(0), o02 = big0 ? d2 : R(0);
    const R o11 = big1 ? d1 : R(0), o12 = big1 ? d2 : R(0);
    R u0[4 * NV], u1[4 * NV], w0[4 * NV], w1[4 * NV];
    load_coeffs<R, NV>(U, n, rec.x, u0);
    load_coeffs<R, NV>(U, n, rec.y, u1);
    adj_weights<R, NV>(A, P, X, rec.x, w0);
    adj_weights<R, NV>(A, P, X, rec.y, w1);
    R M0[3][NV], M1[3][NV];
    #pragma unroll
    for (int m = 0; m < 3; m++)
        #pragma unroll
        for (int v = 0; v < NV; v++) { M0[m][v] = R(0); M1[m][v] = R(0); }
    #pragma unroll 1
    for (int k = 0; k < 6; k++) {
        const R g1 = gp_t1<R>(k), g2 = gp_t2<R>(k);
        const R q01 = o01 + m0 * g1, q02 = o02 + m0 * g2;     // side 0 tangential coordinates
        const R q11 = o11 + m1 * g1, q12 = o12 + m1 * g2;     // side 1
        R tL[NV], tR[NV], G[NV];
        #pragma unroll
        for (int v = 0; v < NV; v++) {
            tL[v] = u0[4*v] + pa0 * u0[4*v+1+a] + q01 * u0[4*v+1+t1] + q02 * u0[4*v+1+t2];
            tR[v] = u1[4*v] + pa1 * u1[4*v+1+a] + q11 * u1[4*v+1+t1] + q12 * u1[4*v+1+t2];
            const R wl = w0[4*v] + pa0 * w0[4*v+1+a] + q01 * w0[4*v+1+t1] + q02 * w0[4*v+1+t2];
            const R wr = w1[4*v] + pa1 * w1[4*v+1+a] + q11 * w1[4*v+1+t1] + q12 * w1[4*v+1+t2];
            G[v] = wj * (wl - wr);
        }
        Prim<R, NC> pL, pR;
        cons_to_prim<R, NC>(P, tL, pL);
        cons_to_prim<R, NC>(P, tR, pR);
        Frame<R, NC> fL, fR;
        frame_of<R, NC>(pL, a, fL);
        frame_of<R, NC>(pR, a, fR);
        fL.un *= s; fR.un *= s;
        R qb[NV];
        qb[0] = s * G[a]; qb[1] = G[t1]; qb[2] = G[t2];
        #pragma unroll
        for (int v = 3; v < NV; v++) qb[v] = G[v];
        FrameBar<R, NC> b0, b1;
        zero_bar<R, NC>(b0);
        zero_bar<R, NC>(b1);
        flux_vjp<R, NC, FLUX>(fL, fR, qb, b0, b1);
        b0.un *= s; b1.un *= s;
        PrimBar<R, NC> pb;
        R tb[NV];
        prim_bar_of<R, NC>(b0, a, pb);
        cons_to_prim_vjp<R, NC>(P, tL, pL, pb, tb);
        #pragma unroll
        for (int v = 0; v < NV; v++) { M0[0][v] += tb[v]; M0[1][v] += tb[v] * g1; M0[2][v] += tb[v] * g2; }
        prim_bar_of<R, NC>(b1, a, pb);
        cons_to_prim_vjp<R, NC>(P, tR, pR, pb, tb);
        #pragma unroll
        for (int v = 0; v < NV; v++) { M1[0][v] += tb[v]; M1[1][v] += tb[v] * g1; M1[2][v] += tb[v] * g2; }
    }
    #pragma unroll
    for (int m = 0; m < 3; m++)
        #pragma unroll
        for (int v = 0; v < NV; v++) {
            out[m * NV + v] = M0[m][v];
            out[(3 + m) * NV + v] = M1[m][v];
        }
}

// Chunked face-centric stage adjoint (interior faces and volume; the
// boundary faces' per-point parts come from k_adj_bnd_pts, launched first,
// and phase 2 adds each leaf's parts): the chunks of the stage kernel; phase 1
// one thread per face of the chunk writes both sides' trace-gradient
// moments to LDS, phase 2 one thread per element adds its volume term and
// gathers its face slots (a fixed slot list, no atomics).  Each interior
// face is evaluated once per chunk that holds one of its sides.
template <typename R, int NC, int FLUX, int BS>
__global__ __launch_bounds__(BS)
void k_adj_chunk(AdjArgs A, DevParams P)
{
    constexpr int NV = 4 + NC;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    R *mom = reinterpret_cast<R *>(smem_raw);                           // [maxf][2][3][NV]
    int *finfo = reinterpret_cast<int *>(smem_raw + (size_t) A.max_faces * 6 * NV * sizeof(R));
    const R *__restrict__ U = reinterpret_cast<const R *>(A.U);
    const R *__restrict__ X = reinterpret_cast<const R *>(A.X);
    const int c = xcd_block(A.n_xcd);
    if (c >= A.n_chunks) return;                 // whole workgroup exits together
    const int fb = A.ch_fbeg[c], fe = A.ch_fbeg[c + 1];
    const int eb = A.ch_ebeg[c], ee = A.ch_ebeg[c + 1];
    for (int f = fb + (int) threadIdx.x; f < fe; f += BS) {
        const int4 rec = A.cf4[f];
        if (((rec.z >> 3) & 3) != FACE_BND) {
            // faces of a chunk are sorted by axis, so the branch is mostly wave-uniform
            R *dst = mom + (size_t) (f - fb) * 6 * NV;
            const int ax = (rec.z & 7) >> 1;
            if (ax == 0) adj_face<R, NC, FLUX, 0>(A, P, U, X, rec, dst);
            else if (ax == 1) adj_face<R, NC, FLUX, 1>(A, P, U, X, rec, dst);
            else adj_face<R, NC, FLUX, 2>(A, P, U, X, rec, dst);
        }
        finfo[f - fb] = rec.z;
    }
    __syncthreads();
    const int e = eb + (int) threadIdx.x;
    if (e >= ee) return;
    const int64_t n = A.n;
    const int lev = A.level[e];
    const R h = ldexp(R(1), -lev), hinv = ldexp(R(1), lev);
    R u[4 * NV], w[4 * NV], ub[4 * NV];
    load_coeffs<R, NV>(U, n, e, u);
    adj_weights<R, NV>(A, P, X, e, w);
    #pragma unroll
    for (int i = 0; i < 4 * NV; i++) ub[i] = R(0);
    adj_volume<R, NC>(P, u, w, h, hinv, ub);
    const int s_end = A.slot_ptr[e + 1];
    for (int si = A.slot_ptr[e]; si < s_end; si++) {
        const int sl = A.slots[si];
        const int fl = sl >> 1, side = sl & 1;
        const int info = finfo[fl];
        const int dir0 = info & 7, kind = (info >> 3) & 3, sub = (info >> 5) & 3;
        if (kind == FACE_BND) continue;
        const int ax = dir0 >> 1;
        const R pa = ((dir0 & 1) ^ side) ? R(0.5) : R(-0.5);
        const bool big = (kind == FACE_HANG0 && side == 1) || (kind == FACE_HANG1 && side == 0);
        const R mu = big ? R(0.5) : R(1);
        const R d1 = big ? ((sub & 1) ? R(0.25) : R(-0.25)) : R(0);
        const R d2 = big ? ((sub & 2) ? R(0.25) : R(-0.25)) : R(0);
        const R *m = mom + (size_t) fl * 6 * NV + (size_t) side * 3 * NV;
        #pragma unroll
        for (int v = 0; v < NV; v++) {
            const R S0 = m[v], S1 = m[NV + v], S2 = m[2 * NV + v];
            // (normal, t1, t2) parts spread onto (x, y, z)
            const R cn = pa * S0, c1 = mu * S1 + d1 * S0, c2 = mu * S2 + d2 * S0;
            ub[4*v+0] += S0;
            ub[4*v+1] += (ax == 0) ? cn : c1;
            ub[4*v+2] += (ax == 1) ? cn : ((ax == 0) ? c1 : c2);
            ub[4*v+3] += (ax == 2) ? cn : c2;
        }
    }
    // boundary faces: the per-point parts of k_adj_bnd_pts, in list order
    const int bi = A.bnd_of ? A.bnd_of[e] : -1;
    if (bi >= 0) {
        const R *__restrict__ part = reinterpret_cast<const R *>(A.bnd_part);
        double th[3] = {0.0, 0.0, 0.0};
        for (int p = 6 * A.bnd_ptr[bi]; p < 6 * A.bnd_ptr[bi + 1]; p++) {
            #pragma unroll
            for (int j = 0; j < 4 * NV; j++) ub[j] += part[(size_t) p * 4 * NV + j];
            if (A.theta_bar) {
                #pragma unroll
                for (int j = 0; j < 3; j++) th[j] += A.bnd_theta[3 * (size_t) p + j];
            }
        }
        if (A.theta_bar) {
            #pragma unroll
            for (int j = 0; j < 3; j++) A.theta_bar[3 * (size_t) e + j] += th[j];
        }
    }
    // not __restrict__: the last stage adjoint accumulates into its own output
    R *out = reinterpret_cast<R *>(A.OUT);
    const R *acc = reinterpret_cast<const R *>(A.ACC);
    const R *acc2 = reinterpret_cast<const R *>(A.ACC2);
    #pragma unroll
    for (int i = 0; i < 4 * NV; i++) {
        const R sb = R(A.c_in) * (X + i * n)[e];
        R o = R(A.c_out) * (sb + ub[i]);
        if (acc) o += R(A.c_acc) * (acc + i * n)[e];
        if (acc2) o += R(A.c_acc2) * (acc2 + i * n)[e];
        (out + i * n)[e] = o;
    }
}

// Boundary faces, one thread per (boundary face, point): the boundary state
// functions run on forward-mode dual numbers (tangents = the leaf's trace);
// the point's contribution phi_k (d s / d t) to the leaf's coefficient
// gradient (and, on the inlet face of the parameters, d s / d theta) is
// written to its own slot of a scratch array, summed per leaf in a fixed
// order by k_adj_bnd_sum (element-centric path) or by phase 2 of
// k_adj_chunk.
template <typename R, int NC, int FLUX>
__global__ __launch_bounds__(64)
void k_adj_bnd_pts(AdjArgs A, DevParams P)
{
    constexpr int NV = 4 + NC;
    using D = Dual<R, NV>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 6 * A.n_bface) return;
    const int f = i / 6, k = i - 6 * f;
    const int code = A.bnd_faces[f];
    const int e = code >> 3, d = code & 7;
    const int face = -1 - A.nbr[6 * e + d], a = d >> 1;
    const int64_t n = A.n;
    const R *__restrict__ U = reinterpret_cast<const R *>(A.U);
    const R *__restrict__ X = reinterpret_cast<const R *>(A.X);
    const R h = ldexp(R(1), -(int) A.level[e]);
    const R wjf = h * h * R(1.0 / 6.0);
    const R sg = (d & 1) ? R(1) : R(-1);
    R u[4 * NV], w[4 * NV];
    load_coeffs<R, NV>(U, n, e, u);
    adj_weights<R, NV>(A, P, X, e, w);
    R ps[3];
    face_point<R>(d, k, R(0), R(0), R(1), ps);
    R *__restrict__ part = reinterpret_cast<R *>(A.bnd_part) + (size_t) i * 4 * NV;
    {
        D ts[NV];
        #pragma unroll
        for (int v = 0; v < NV; v++) ts[v] = D::var(at<R>(&u[4 * v], ps[0], ps[1], ps[2]), v);
        Prim<D, NC> pi, po;
        cons_to_prim<D, NC>(P, ts, pi);
        bnd_state<D, NC>(P, face, a, D(sg), pi, po, -1);
        D qg[NV];
        flux_global<D, NC, FLUX>(pi, po, a, D(sg), qg);
        D s(0.0);
        #pragma unroll
        for (int v = 0; v < NV; v++) s += qg[v] * D(wjf * wdot<R, NV>(w, v, ps));
        const R phi[4] = {R(1), ps[0], ps[1], ps[2]};
        #pragma unroll
        for (int v = 0; v < NV; v++)
            #pragma unroll
            for (int q = 0; q < 4; q++) part[4 * v + q] = phi[q] * s.d[v];
    }
    if (A.theta_bar) {
        double th[3] = {0.0, 0.0, 0.0};
        if (face == A.theta_face && P.bnd_type[face] == BND_INLET) {
            using T3 = Dual<R, 3>;
            T3 ts[NV];
            #pragma unroll
            for (int v = 0; v < NV; v++) ts[v] = T3(at<R>(&u[4 * v], ps[0], ps[1], ps[2]));
            Prim<T3, NC> pi, po;
            cons_to_prim<T3, NC>(P, ts, pi);
            inlet_dual<R, NC>(P, face, pi, po);
            T3 qg[NV];
            flux_global<T3, NC, FLUX>(pi, po, a, T3(sg), qg);
            T3 s(0.0);
            #pragma unroll
            for (int v = 0; v < NV; v++) s += qg[v] * T3(wjf * wdot<R, NV>(w, v, ps));
            th[0] = (double) s.d[0]; th[1] = (double) s.d[1]; th[2] = (double) s.d[2];
        }
        #pragma unroll
        for (int j = 0; j < 3; j++) A.bnd_theta[3 * (size_t) i + j] = th[j];
    }
}

// per boundary leaf: OUT += c_out * (sum over its boundary faces and their
// points, in list order), theta_bar += the same sum of the parameter parts
template <typename R, int NV>
__global__ __launch_bounds__(128)
void k_adj_bnd_sum(AdjArgs A)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n_bnd) return;
    const int e = A.bnd_elems[i];
    const int f0 = A.bnd_ptr[i], f1 = A.bnd_ptr[i + 1];
    const R *__restrict__ part = reinterpret_cast<const R *>(A.bnd_part);
    R acc[4 * NV];
    #pragma unroll
    for (int j = 0; j < 4 * NV; j++) acc[j] = R(0);
    double th[3] = {0.0, 0.0, 0.0};
    for (int p = 6 * f0; p < 6 * f1; p++) {
        #pragma unroll
        for (int j = 0; j < 4 * NV; j++) acc[j] += part[(size_t) p * 4 * NV + j];
        if (A.theta_bar) {
            #pragma unroll
            for (int j = 0; j < 3; j++) th[j] += A.bnd_theta[3 * (size_t) p + j];
        }
    }
    R *out = reinterpret_cast<R *>(A.OUT);
    const int64_t n = A.n;
    #pragma unroll
    for (int j = 0; j < 4 * NV; j++) (out + j * n)[e] += R(A.c_out) * acc[j];
    if (A.theta_bar) {
        A.theta_bar[3 * e + 0] += th[0];
        A.theta_bar[3 * e + 1] += th[1];
        A.theta_bar[3 * e + 2] += th[2];
    }
}

template <typename R, int NC>
static hipError_t launch_adj_bnd_pts(int flux, const AdjArgs &A, const DevParams &P, hipStream_t st)
{
    if (A.n_bnd <= 0) return hipSuccess;
    const dim3 gp((unsigned) ((6 * (int64_t) A.n_bface + 63) / 64)), bp(64);
    switch (flux) {
    case FLUX_LF: hipLaunchKernelGGL((k_adj_bnd_pts<R, NC, FLUX_LF>), gp, bp, 0, st, A, P); break;
    case FLUX_HLLC: hipLaunchKernelGGL((k_adj_bnd_pts<R, NC, FLUX_HLLC>), gp, bp, 0, st, A, P); break;
    default: hipLaunchKernelGGL((k_adj_bnd_pts<R, NC, FLUX_CD>), gp, bp, 0, st, A, P); break;
    }
    return hipGetLastError();
}

// element-centric path: boundary points after the interior kernel, then the
// per-leaf sum into OUT
template <typename R, int NC>
static hipError_t launch_adj_bnd(int flux, const AdjArgs &A, const DevParams &P, hipStream_t st)
{
    if (A.n_bnd <= 0) return hipSuccess;
    hipError_t err = launch_adj_bnd_pts<R, NC>(flux, A, P, st);
    if (err != hipSuccess) return err;
    const dim3 gs((unsigned) ((A.n_bnd + 127) / 128)), bs(128);
    hipLaunchKernelGGL((k_adj_bnd_sum<R, 4 + NC>), gs, bs, 0, st, A);
    return hipGetLastError();
}

// objective Q(U) = sum over leaves with centre in the box of sum_gp wJ p:
// per-block FP64 partial sums (mode 0), or lambda += coef dQ/dU (mode 1)
template <typename R, int NC>
__global__ __launch_bounds__(256)
void k_adj_objective(ObjArgs A, DevParams P)
{
    constexpr int NV = 4 + NC;
    using D = Dual<R, NV>;
    __shared__ double red[256];
    const int64_t n = A.n;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    double val = 0.0;
    if (e < n) {
        const int4 q = A.q[e];
        const int len = (1 << 19) >> q.w;
        const double inv = 1.0 / (double) (1 << 19);
        const double c[3] = {(q.x + 0.5 * len) * inv, (q.y + 0.5 * len) * inv, (q.z + 0.5 * len) * inv};
        bool in = true;
        #pragma unroll
        for (int d = 0; d < 3; d++) in = in && c[d] >= A.lo[d] && c[d] <= A.hi[d];
        if (in) {
            const R *__restrict__ U = reinterpret_cast<const R *>(A.U);
            R u[4 * NV], g[4 * NV];
            load_coeffs<R, NV>(U, n, e, u);
            #pragma unroll
            for (int i = 0; i < 4 * NV; i++) g[i] = R(0);
            const R h = ldexp(R(1), -q.w), wj = h * h * h * R(0.125);
            const R gph = R(0.28867513459481288225);
            for (int gp = 0; gp < 8; gp++) {
                const R ph[3] = {(gp & 1) ? gph : -gph, (gp & 2) ? gph : -gph, (gp & 4) ? gph : -gph};
                D t[NV];
                #pragma unroll
                for (int v = 0; v < NV; v++) t[v] = D::var(at<R>(&u[4 * v], ph[0], ph[1], ph[2]), v);
                Prim<D, NC> p;
                cons_to_prim<D, NC>(P, t, p);
                val += (double) (wj * p.p.v);
                const R phi[4] = {R(1), ph[0], ph[1], ph[2]};
                #pragma unroll
                for (int v = 0; v < NV; v++)
                    #pragma unroll
                    for (int k = 0; k < 4; k++) g[4 * v + k] += wj * phi[k] * p.p.d[v];
            }
            if (A.mode == 1) {
                R *__restrict__ L = reinterpret_cast<R *>(A.lambda);
                #pragma unroll
                for (int i = 0; i < 4 * NV; i++) (L + i * n)[e] += R(A.coef) * g[i];
            }
        }
    }
    if (A.mode == 0) {
        red[threadIdx.x] = val;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int) threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) A.partial[blockIdx.x] = red[0];
    }
}

template <typename R, int NC, int FLUX, int BS>
static hipError_t launch_adj_chunk_bs(const AdjArgs &A, const DevParams &P, hipStream_t st)
{
    const size_t sh = adj_chunk_smem_bytes(sizeof(R) == 4, NC, A.max_faces);
    const void *fn = reinterpret_cast<const void *>(&k_adj_chunk<R, NC, FLUX, BS>);
    hipError_t err = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int) sh);
    if (err != hipSuccess) return err;
    const dim3 grid((unsigned) (((A.n_chunks + A.n_xcd - 1) / A.n_xcd) * A.n_xcd)), block(BS);
    hipLaunchKernelGGL((k_adj_chunk<R, NC, FLUX, BS>), grid, block, sh, st, A, P);
    return hipGetLastError();
}

template <typename R, int NC, int FLUX>
static hipError_t launch_adj_chunk(const AdjArgs &A, const DevParams &P, hipStream_t st)
{
    // FP32 only: the FP64 face phase does not fit the registers
    if constexpr (sizeof(R) != 4) return hipErrorInvalidValue;
    if (A.chunk_elems > 256) return hipErrorInvalidValue;
    return A.chunk_elems > 128 ? launch_adj_chunk_bs<R, NC, FLUX, 256>(A, P, st)
                               : launch_adj_chunk_bs<R, NC, FLUX, 128>(A, P, st);
}

template <typename R, int NC>
static hipError_t launch_adj_nc(int flux, const AdjArgs &A, const DevParams &P, hipStream_t st)
{
    if (sizeof(R) == 4 && A.cf4) {
        hipError_t err = launch_adj_bnd_pts<R, NC>(flux, A, P, st);
        if (err != hipSuccess) return err;
        switch (flux) {
        case FLUX_LF: err = launch_adj_chunk<R, NC, FLUX_LF>(A, P, st); break;
        case FLUX_HLLC: err = launch_adj_chunk<R, NC, FLUX_HLLC>(A, P, st); break;
        case FLUX_CD: err = launch_adj_chunk<R, NC, FLUX_CD>(A, P, st); break;
        default: return hipErrorInvalidValue;
        }
        return err;
    }
    const dim3 grid((unsigned) ((A.n + 127) / 128)), block(128);
    switch (flux) {
    case FLUX_LF: hipLaunchKernelGGL((k_adj_stage<R, NC, FLUX_LF>), grid, block, 0, st, A, P); break;
    case FLUX_HLLC: hipLaunchKernelGGL((k_adj_stage<R, NC, FLUX_HLLC>), grid, block, 0, st, A, P); break;
    case FLUX_CD: hipLaunchKernelGGL((k_adj_stage<R, NC, FLUX_CD>), grid, block, 0, st, A, P); break;
    default: return hipErrorInvalidValue;
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    return launch_adj_bnd<R, NC>(flux, A, P, st);
}

template <typename R>
static hipError_t launch_adj_r(int nc, int flux, const AdjArgs &A, const DevParams &P, hipStream_t st)
{
    switch (nc) {
    case 1: return launch_adj_nc<R, 1>(flux, A, P, st);
    case 2: return launch_adj_nc<R, 2>(flux, A, P, st);
    }
    return hipErrorInvalidValue;
}

// Pointwise flux VJP: the computation of k_flux_points (prim rows -> EOS ->
// face frame -> frame flux -> rotation back) on dual numbers carrying the
// 2 (5 + NC) inputs of the point as tangents; bar = J^T qbar.
template <int NC, int FLUX>
__global__ void k_flux_points_vjp(FluxArgs A, DevParams P)
{
    constexpr int NV = 4 + NC, NI = 5 + NC;
    using D = Dual<double, 2 * NI>;
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    Prim<D, NC> pp[2];
    for (int s = 0; s < 2; s++) {
        const double *row = (s ? A.prim_r : A.prim_l) + i * NI;
        Prim<D, NC> &p = pp[s];
        p.r = D::var(row[0], s * NI + 0); p.u = D::var(row[1], s * NI + 1); p.v = D::var(row[2], s * NI + 2);
        p.w = D::var(row[3], s * NI + 3); p.p = D::var(row[4], s * NI + 4);
        #pragma unroll
        for (int c = 0; c < NC; c++) p.c[c] = D::var(row[5 + c], s * NI + 5 + c);
        eos_coeffs<D, NC>(P, p);
        p.e = p.p / (p.r * (p.gam - D(1)));
        p.e_tot = p.e + D(0.5) * (p.u * p.u + p.v * p.v + p.w * p.w);
        p.cz = dsqrt(p.gam * p.p / p.r);
    }
    double nt[3][3];
    ref_face_frame(A.normals + 3 * i, nt);
    Frame<D, NC> f[2];
    for (int s = 0; s < 2; s++)
        f[s] = to_frame<D, NC>(pp[s], D(nt[0][0]), D(nt[0][1]), D(nt[0][2]), D(nt[1][0]), D(nt[1][1]),
                               D(nt[1][2]), D(nt[2][0]), D(nt[2][1]), D(nt[2][2]));
    D q[NV];
    flux_x<D, NC, FLUX>(f[0], f[1], q);
    D o[NV];
    for (int d = 0; d < 3; d++) o[d] = q[0] * D(nt[0][d]) + q[1] * D(nt[1][d]) + q[2] * D(nt[2][d]);
    for (int v = 3; v < NV; v++) o[v] = q[v];
    const double *qb = A.qbar + i * NV;
    for (int j = 0; j < 2 * NI; j++) {
        double sj = 0.0;
        for (int v = 0; v < NV; v++) sj += qb[v] * o[v].d[j];
        if (j < NI) A.bar_l[i * NI + j] = sj;
        else A.bar_r[i * NI + j - NI] = sj;
    }
}

template <int NC>
static hipError_t launch_flux_vjp_nc(int flux, const FluxArgs &A, const DevParams &P, hipStream_t st)
{
    const dim3 grid((unsigned) ((A.n + 63) / 64)), block(64);
    switch (flux) {
    case FLUX_LF: hipLaunchKernelGGL((k_flux_points_vjp<NC, FLUX_LF>), grid, block, 0, st, A, P); break;
    case FLUX_HLLC: hipLaunchKernelGGL((k_flux_points_vjp<NC, FLUX_HLLC>), grid, block, 0, st, A, P); break;
    case FLUX_CD: hipLaunchKernelGGL((k_flux_points_vjp<NC, FLUX_CD>), grid, block, 0, st, A, P); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_flux_points_vjp(int nc, int flux, const FluxArgs &A, const DevParams &P, hipStream_t st)
{
    if (A.n <= 0) return hipSuccess;
    switch (nc) {
    case 1: return launch_flux_vjp_nc<1>(flux, A, P, st);
    case 2: return launch_flux_vjp_nc<2>(flux, A, P, st);
    }
    return hipErrorInvalidValue;
}

size_t adj_chunk_smem_bytes(int prec, int nc, int max_faces)
{
    return (size_t) max_faces * (6 * (4 + nc) * (prec ? 4 : 8) + sizeof(int));
}

hipError_t launch_adj_stage(int prec, int nc, int flux, const AdjArgs &A, const DevParams &P, hipStream_t st)
{
    if (A.n <= 0) return hipSuccess;
    return prec ? launch_adj_r<float>(nc, flux, A, P, st) : launch_adj_r<double>(nc, flux, A, P, st);
}

template <typename R>
static hipError_t launch_obj_r(int nc, const ObjArgs &A, const DevParams &P, hipStream_t st)
{
    const dim3 grid((unsigned) ((A.n + 255) / 256)), block(256);
    switch (nc) {
    case 1: hipLaunchKernelGGL((k_adj_objective<R, 1>), grid, block, 0, st, A, P); break;
    case 2: hipLaunchKernelGGL((k_adj_objective<R, 2>), grid, block, 0, st, A, P); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_adj_objective(int prec, int nc, const ObjArgs &A, const DevParams &P, hipStream_t st)
{
    if (A.n <= 0) return hipSuccess;
    return prec ? launch_obj_r<float>(nc, A, P, st) : launch_obj_r<double>(nc, A, P, st);
}


} // namespace charm
