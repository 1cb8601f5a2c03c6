// charm_vjp.hpp -- hand-written reverse-mode (vector-Jacobian) products of
// the pointwise numerics of the Euler terms (constant cp per species):
// conservative -> primitive variables, the x-split HLLC and Lax-Friedrichs /
// central fluxes in the face frame, and the volume flux contraction.
//
// Each *_vjp function takes the forward inputs, recomputes the primal
// quantities it needs (in the operation order of the forward code, so that
// every branch -- floors, clamps, wave-speed ordering, flux regions --
// follows the forward evaluation) and returns the gradient of a scalar
// contraction of the outputs.  One reverse sweep gives the gradient with
// respect to both sides of a face for the cost of a few primal
// evaluations, where forward-mode dual numbers need one tangent per input.
#pragma once
#include <hip/hip_runtime.h>
#include "charm_device.hpp"

namespace charm {

// adjoints of the primitive fields the fluxes read (r is the floored density)
template <typename R, int NC>
struct PrimBar {
    R r, u, v, w, e_tot, p, cz, gam;
    R c[NC];
};

template <typename R, int NC>
struct FrameBar {
    R r, un, ut1, ut2, e_tot, p, cz, gam;
    R c[NC];
};

template <typename R, int NC>
__device__ __forceinline__ void zero_bar(FrameBar<R, NC> &b)
{
    b.r = b.un = b.ut1 = b.ut2 = b.e_tot = b.p = b.cz = b.gam = R(0);
    #pragma unroll
    for (int i = 0; i < NC; i++) b.c[i] = R(0);
}

// d gamma / d c_i of the constant-cp mixture, gamma = cp / (cp - R M_),
// cp = sum c_i cp_i, M_ = sum c_i / M_i (eos_coeffs, PH = 0)
template <typename R, int NC>
__device__ __forceinline__ void mix_gamma_grad(const DevParams &P, const R *c, R *dg)
{
    R cp = R(0), M_ = R(0);
    #pragma unroll
    for (int i = 0; i < NC; i++) {
        M_ += dv(c[i], cst<R>(P.comp_m[i], P.comp_m_f[i]));
        cp += c[i] * cst<R>(P.comp_cp[i], P.comp_cp_f[i]);
    }
    const R m = rcp(M_);
    const R cv = cp - dv(R(kGasR), m);
    const R icv2 = rcp(cv * cv);
    #pragma unroll
    for (int i = 0; i < NC; i++) {
        const R cpi = cst<R>(P.comp_cp[i], P.comp_cp_f[i]);
        const R dcv = cpi - dv(R(kGasR), cst<R>(P.comp_m[i], P.comp_m_f[i]));
        dg[i] = (cpi * cv - cp * dcv) * icv2;
    }
}

// Reverse of cons_to_prim<R, NC, 0> at the trace t (NV conservative values)
// with primal result p: tb = (d / d t) of sum (pb . prim fields).
template <typename R, int NC>
__device__ __forceinline__ void cons_to_prim_vjp(const DevParams &P, const R *t, const Prim<R, NC> &p,
                                                 const PrimBar<R, NC> &pb, R *tb)
{
    constexpr int NV = 4 + NC;
    R r = t[4];
    #pragma unroll
    for (int i = 1; i < NC; i++) r += t[4 + i];
    const R ri = rcp(r);
    const R gm1 = p.gam - R(1);
    R eb = R(0), gamb = pb.gam;
    // cz = sqrt(gam (gam - 1) e)
    if (p.cz > R(0)) {
        const R g2b = pb.cz * R(0.5) * rcp(p.cz);
        eb += g2b * p.gam * gm1;
        gamb += g2b * p.e * (p.gam + gm1);
    }
    // p = r_floored e (gam - 1)
    const R rfb = pb.r + pb.p * p.e * gm1;
    eb += pb.p * p.r * gm1;
    gamb += pb.p * p.r * p.e;
    R rb = (r < R(kEps)) ? R(0) : rfb;
    R cb[NC];
    #pragma unroll
    for (int i = 0; i < NC; i++) cb[i] = pb.c[i];
    if constexpr (NC > 1) {
        if (P.eos != EOS_IDEAL) {
            R dg[NC];
            mix_gamma_grad<R, NC>(P, p.c, dg);
            #pragma unroll
            for (int i = 0; i < NC; i++) cb[i] += gamb * dg[i];
        }
    }
    // e = e_tot - (u^2 + v^2 + w^2) / 2
    const R etb = pb.e_tot + eb;
    const R ub = pb.u - eb * p.u, vb = pb.v - eb * p.v, wb = pb.w - eb * p.w;
    tb[0] = ub * ri; tb[1] = vb * ri; tb[2] = wb * ri; tb[3] = etb * ri;
    R rib = ub * t[0] + vb * t[1] + wb * t[2] + etb * t[3];
    // c_i = clamp(rho_i / r, 0, 1): the gradient passes inside [0, 1]
    #pragma unroll
    for (int i = 0; i < NC; i++) {
        const R raw = t[4 + i] * ri;
        const bool in = raw >= R(0) && raw <= R(1);
        tb[4 + i] = in ? cb[i] * ri : R(0);
        rib += in ? cb[i] * t[4 + i] : R(0);
    }
    rb -= rib * ri * ri;
    #pragma unroll
    for (int v = 4; v < NV; v++) tb[v] += rb;
}

// Reverse of hllc_x: given qb (frame order), accumulate into Lb and Rb.
template <typename R, int NC>
__device__ __forceinline__ void hllc_vjp(const Frame<R, NC> &L, const Frame<R, NC> &Rt, const R *qb,
                                         FrameBar<R, NC> &Lb, FrameBar<R, NC> &Rb)
{
    // primal, as hllc_x
    const R du = Rt.un - L.un, rs = L.r + Rt.r, cs = L.cz + Rt.cz;
    const R p_pvrs = R(0.5) * (L.p + Rt.p) - R(0.5) * du * R(0.25) * rs * cs;
    const bool pos = p_pvrs > R(0);
    const R p_star = pos ? p_pvrs : R(0);
    const R gl = dv(L.gam + R(1), R(2) * L.gam), gr = dv(Rt.gam + R(1), R(2) * Rt.gam);
    const bool shl = !(p_star <= L.p), shr = !(p_star <= Rt.p);
    const R xl = dv(p_star, L.p), xr = dv(p_star, Rt.p);
    const R ql = shl ? dsqrt(R(1) + gl * (xl - R(1))) : R(1);
    const R qr = shr ? dsqrt(R(1) + gr * (xr - R(1))) : R(1);
    const R sl0 = L.un - L.cz * ql, sr0 = Rt.un + Rt.cz * qr;
    const bool sw = sl0 > sr0;
    const R sl = sw ? sr0 : sl0, sr = sw ? sl0 : sr0;
    const R aL = sl - L.un, aR = sr - Rt.un;
    R num = Rt.p - L.p;
    num += L.r * L.un * aL;
    num -= Rt.r * Rt.un * aR;
    const R den = L.r * aL - Rt.r * aR;
    const R sst0 = dv(num, den);
    int clamp = 0;
    R sst = sst0;
    if (sst < sl) { sst = sl; clamp = 1; }
    if (sst > sr) { sst = sr; clamp = 2; }

    R slb = R(0), srb = R(0), sstb = R(0);
    if (sl >= R(0) || sr <= R(0)) {
        const bool left = sl >= R(0);
        const Frame<R, NC> &K = left ? L : Rt;
        FrameBar<R, NC> &Kb = left ? Lb : Rb;
        Kb.r += qb[0] * K.un * K.un + qb[1] * K.ut1 * K.un + qb[2] * K.ut2 * K.un + qb[3] * K.e_tot * K.un;
        Kb.un += qb[0] * R(2) * K.r * K.un + qb[1] * K.r * K.ut1 + qb[2] * K.r * K.ut2
               + qb[3] * (K.r * K.e_tot + K.p);
        Kb.ut1 += qb[1] * K.r * K.un;
        Kb.ut2 += qb[2] * K.r * K.un;
        Kb.e_tot += qb[3] * K.r * K.un;
        Kb.p += qb[0] + qb[3] * K.un;
        #pragma unroll
        for (int i = 0; i < NC; i++) {
            Kb.r += qb[4 + i] * K.un * K.c[i];
            Kb.un += qb[4 + i] * K.r * K.c[i];
            Kb.c[i] += qb[4 + i] * K.r * K.un;
        }
    } else {
        const bool left = sst >= R(0);
        const Frame<R, NC> &K = left ? L : Rt;
        FrameBar<R, NC> &Kb = left ? Lb : Rb;
        const R sk = left ? sl : sr;
        const R A = sk - K.un, B = sst - K.un, D = sk - sst;
        const R inv = rcp(D);
        const R pk = K.p + K.r * A * B;
        const R N0 = sst * (K.r * K.un * A - K.p) + sk * pk;
        const R N1 = sst * K.r * K.ut1 * A, N2 = sst * K.r * K.ut2 * A;
        const R T3 = K.r * K.e_tot * A - K.p * K.un + sk * pk;
        const R N3 = sst * T3;
        R invb = qb[0] * N0 + qb[1] * N1 + qb[2] * N2 + qb[3] * N3;
        const R N0b = qb[0] * inv, N1b = qb[1] * inv, N2b = qb[2] * inv, N3b = qb[3] * inv;
        R skb = R(0), Ab = R(0), Bb = R(0), pkb = R(0);
        R rb = R(0), unb = R(0), pb = R(0);
        // N0
        sstb += N0b * (K.r * K.un * A - K.p);
        rb += N0b * sst * K.un * A;
        unb += N0b * sst * K.r * A;
        Ab += N0b * sst * K.r * K.un;
        pb -= N0b * sst;
        skb += N0b * pk;
        pkb += N0b * sk;
        // N1, N2
        sstb += N1b * K.r * K.ut1 * A + N2b * K.r * K.ut2 * A;
        rb += N1b * sst * K.ut1 * A + N2b * sst * K.ut2 * A;
        Kb.ut1 += N1b * sst * K.r * A;
        Kb.ut2 += N2b * sst * K.r * A;
        Ab += N1b * sst * K.r * K.ut1 + N2b * sst * K.r * K.ut2;
        // N3 = sst T3
        sstb += N3b * T3;
        const R T3b = N3b * sst;
        rb += T3b * K.e_tot * A;
        Kb.e_tot += T3b * K.r * A;
        Ab += T3b * K.r * K.e_tot;
        pb -= T3b * K.un;
        unb -= T3b * K.p;
        skb += T3b * pk;
        pkb += T3b * sk;
        // species
        #pragma unroll
        for (int i = 0; i < NC; i++) {
            const R Nb = qb[4 + i] * inv;
            const R Ni = sst * K.r * K.c[i] * A;
            invb += qb[4 + i] * Ni;
            sstb += Nb * K.r * K.c[i] * A;
            rb += Nb * sst * K.c[i] * A;
            Kb.c[i] += Nb * sst * K.r * A;
            Ab += Nb * sst * K.r * K.c[i];
        }
        // pk = p + r A B
        pb += pkb;
        rb += pkb * A * B;
        Ab += pkb * K.r * B;
        Bb += pkb * K.r * A;
        // inv = 1 / D, D = sk - sst
        const R Db = -invb * inv * inv;
        skb += Db;
        sstb -= Db;
        // A = sk - un, B = sst - un
        skb += Ab;
        unb -= Ab;
        sstb += Bb;
        unb -= Bb;
        Kb.r += rb; Kb.un += unb; Kb.p += pb;
        if (left) slb += skb; else srb += skb;
    }
    // s* = clamp(num / den, sl, sr)
    if (clamp == 1) slb += sstb;
    else if (clamp == 2) srb += sstb;
    else {
        const R idn = rcp(den);
        const R numb = sstb * idn, denb = -sstb * sst0 * idn;
        Rb.p += numb;
        Lb.p -= numb;
        Lb.r += numb * L.un * aL + denb * aL;
        Lb.un += numb * L.r * aL;
        Rb.r -= numb * Rt.un * aR + denb * aR;
        Rb.un -= numb * Rt.r * aR;
        const R aLb = numb * L.r * L.un + denb * L.r;
        const R aRb = -numb * Rt.r * Rt.un - denb * Rt.r;
        slb += aLb; Lb.un -= aLb;
        srb += aRb; Rb.un -= aRb;
    }
    // wave-speed ordering
    const R sl0b = sw ? srb : slb, sr0b = sw ? slb : srb;
    Lb.un += sl0b;
    Lb.cz -= sl0b * ql;
    const R qlb = -sl0b * L.cz;
    Rb.un += sr0b;
    Rb.cz += sr0b * qr;
    const R qrb = sr0b * Rt.cz;
    R pstb = R(0);
    if (shl) {
        const R zb = qlb * R(0.5) * rcp(ql);
        Lb.gam += zb * (xl - R(1)) * (R(-0.5) * rcp(L.gam * L.gam));
        const R xb = zb * gl;
        const R ip = rcp(L.p);
        pstb += xb * ip;
        Lb.p -= xb * xl * ip;
    }
    if (shr) {
        const R zb = qrb * R(0.5) * rcp(qr);
        Rb.gam += zb * (xr - R(1)) * (R(-0.5) * rcp(Rt.gam * Rt.gam));
        const R xb = zb * gr;
        const R ip = rcp(Rt.p);
        pstb += xb * ip;
        Rb.p -= xb * xr * ip;
    }
    // p* = max(p_pvrs, 0), p_pvrs = (pL + pR) / 2 - (uR - uL)(rL + rR)(cL + cR) / 8
    if (pos) {
        Lb.p += R(0.5) * pstb;
        Rb.p += R(0.5) * pstb;
        const R k = R(-0.125) * pstb;
        const R dub = k * rs * cs, rsb = k * du * cs, csb = k * du * rs;
        Rb.un += dub; Lb.un -= dub;
        Lb.r += rsb; Rb.r += rsb;
        Lb.cz += csb; Rb.cz += csb;
    }
}

// Reverse of lf_x (central = true: the central flux, no dissipation)
template <typename R, int NC>
__device__ __forceinline__ void lf_vjp(const Frame<R, NC> &L, const Frame<R, NC> &Rt, const R *qb,
                                       FrameBar<R, NC> &Lb, FrameBar<R, NC> &Rb, bool central)
{
    const R aLs = (L.un < R(0) ? -L.un : L.un) + L.cz, aRs = (Rt.un < R(0) ? -Rt.un : Rt.un) + Rt.cz;
    const R alpha = central ? R(0) : (aRs > aLs ? aRs : aLs);
    const R h0 = R(0.5) * qb[0], h1 = R(0.5) * qb[1], h2 = R(0.5) * qb[2], h3 = R(0.5) * qb[3];
    // alpha's coefficient: -(1/2) sum qb_v (R_v - L_v)
    R alb = -(h0 * (Rt.r * Rt.un - L.r * L.un) + h1 * (Rt.r * Rt.ut1 - L.r * L.ut1)
              + h2 * (Rt.r * Rt.ut2 - L.r * L.ut2) + h3 * (Rt.r * Rt.e_tot - L.r * L.e_tot));
    // side terms: sgn = +1 for the right state, -1 for the left in the
    // dissipation; the physical flux enters with +1 on both sides
    #pragma unroll
    for (int s = 0; s < 2; s++) {
        const Frame<R, NC> &K = s ? Rt : L;
        FrameBar<R, NC> &Kb = s ? Rb : Lb;
        const R sg = s ? R(1) : R(-1);
        Kb.r += h0 * (K.un * K.un - alpha * sg * K.un) + h1 * (K.un * K.ut1 - alpha * sg * K.ut1)
              + h2 * (K.un * K.ut2 - alpha * sg * K.ut2) + h3 * (K.e_tot * K.un - alpha * sg * K.e_tot);
        Kb.un += h0 * (R(2) * K.r * K.un - alpha * sg * K.r) + h1 * K.r * K.ut1 + h2 * K.r * K.ut2
               + h3 * (K.r * K.e_tot + K.p);
        Kb.ut1 += h1 * (K.r * K.un - alpha * sg * K.r);
        Kb.ut2 += h2 * (K.r * K.un - alpha * sg * K.r);
        Kb.e_tot += h3 * (K.r * K.un - alpha * sg * K.r);
        Kb.p += h0 + h3 * K.un;
        #pragma unroll
        for (int i = 0; i < NC; i++) {
            const R hc = R(0.5) * qb[4 + i];
            if (s == 1) alb -= hc * (Rt.r * Rt.c[i] - L.r * L.c[i]);
            Kb.r += hc * (K.un * K.c[i] - alpha * sg * K.c[i]);
            Kb.un += hc * K.r * K.c[i];
            Kb.c[i] += hc * (K.r * K.un - alpha * sg * K.r);
        }
    }
    if (!central) {
        // alpha = max(|uL| + cL, |uR| + cR); the right side when strictly larger
        const bool right = aRs > aLs;
        const Frame<R, NC> &K = right ? Rt : L;
        FrameBar<R, NC> &Kb = right ? Rb : Lb;
        Kb.un += alb * ((K.un < R(0)) ? R(-1) : R(1));
        Kb.cz += alb;
    }
}

template <typename R, int NC, int FLUX>
__device__ __forceinline__ void flux_vjp(const Frame<R, NC> &L, const Frame<R, NC> &Rt, const R *qb,
                                         FrameBar<R, NC> &Lb, FrameBar<R, NC> &Rb)
{
    if constexpr (FLUX == FLUX_HLLC) hllc_vjp<R, NC>(L, Rt, qb, Lb, Rb);
    else if constexpr (FLUX == FLUX_LF) lf_vjp<R, NC>(L, Rt, qb, Lb, Rb, false);
    else lf_vjp<R, NC>(L, Rt, qb, Lb, Rb, true);
}

// frame of the axis-aligned normal +e_a (sign +1) from a primitive state
template <typename R, int NC>
__device__ __forceinline__ void frame_of(const Prim<R, NC> &p, int a, Frame<R, NC> &f)
{
    const R vel[3] = {p.u, p.v, p.w};
    const int t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
    f.r = p.r; f.e_tot = p.e_tot; f.p = p.p; f.cz = p.cz; f.gam = p.gam;
    f.un = vel[a]; f.ut1 = vel[t1]; f.ut2 = vel[t2];
    #pragma unroll
    for (int i = 0; i < NC; i++) f.c[i] = p.c[i];
}

// frame bar -> primitive bar (the inverse permutation of frame_of)
template <typename R, int NC>
__device__ __forceinline__ void prim_bar_of(const FrameBar<R, NC> &fb, int a, PrimBar<R, NC> &pb)
{
    const int t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
    R vel[3];
    vel[a] = fb.un; vel[t1] = fb.ut1; vel[t2] = fb.ut2;
    pb.u = vel[0]; pb.v = vel[1]; pb.w = vel[2];
    pb.r = fb.r; pb.e_tot = fb.e_tot; pb.p = fb.p; pb.cz = fb.cz; pb.gam = fb.gam;
    #pragma unroll
    for (int i = 0; i < NC; i++) pb.c[i] = fb.c[i];
}

} // namespace charm
