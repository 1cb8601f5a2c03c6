// charm_device.hpp -- pointwise device numerics of the DG path (gfx950).
//
// Templated on the arithmetic type R (float: the bulk path, double: the
// plain FP64 path) and on the species count NC.  Each function states the
// reference function whose arithmetic it reproduces.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace charm {

constexpr int kMaxComp = 8;
// rows of the point pass per leaf: 24 (sub)face slots x 6 points, then the
// 8 volume points; planes m, cv, gamma (the EOS closure), mu, k (transport)
// of kMemoRows rows each
constexpr int kMemoVolRow = 24 * 6;
constexpr int kMemoRows = kMemoVolRow + 8;
constexpr int kMemoPlanes = 5;
constexpr double kGasR = 8.31446261815324;       // src/charm_eos.c:8
constexpr double kEps = 1.e-12;                   // CHARM_EPS
enum { FLUX_LF = 0, FLUX_HLLC = 1, FLUX_GODUNOV = 2, FLUX_CD = 3 };
enum { BND_INLET = 0, BND_OUTLET = 1, BND_WALL_SLIP = 2, BND_WALL_NO_SLIP = 3, BND_MASS_FLOW = 4 };
// boundary parameters per face: inlet (u, v, w, T, p); no-slip wall (T); mass
// flow (flux, static p, total T, 3 direction cosines, outflow p, composition)
constexpr int kBndParams = 7 + 4;
enum { EOS_IDEAL = 0, EOS_MIX = 1 };
enum { FACE_CONF = 0, FACE_HANG0 = 1, FACE_HANG1 = 2, FACE_BND = 3 };

// Physical constants shared by every kernel of a context (passed by value),
// in FP64 and, for the FP32 path, pre-rounded to FP32 on the host.
struct DevParams {
    int    nc;
    int    eos;
    double gam, cv, m;                 // single-component ideal gas
    double comp_m[kMaxComp];           // mixture: molar masses
    double comp_minv[kMaxComp];        //          1 / M_i (FP64, for the FP32 path's closure)
    double comp_cp[kMaxComp];          //          constant cp per species
    int    bnd_type[6];
    double bnd_param[6][kBndParams];
    double grav[3];
    double cfl;
    double ainv0[4];                   // diagonal of A^-1 at level 0 (FP64 Gauss-Jordan)
    float  ainv0_f[4];
    float  gam_f, cv_f, m_f;
    float  comp_m_f[kMaxComp], comp_cp_f[kMaxComp];
    float  bnd_param_f[6][kBndParams];
    float  grav_f[3];
    // Navier-Stokes / real-gas extensions, read by the PH = 1 kernels only
    int    newton_t;                   // comp 0 has a polynomial cp: T by FP64 Newton
    int    cp_poly[kMaxComp];          // cp(T) is the power series
    int    cp_dt_rest;                 // a species with cp_poly 0 has a series of n_cp > 1
    int    n_cp[kMaxComp];
    double cp_c[kMaxComp][8];
    int    ml_type[kMaxComp], kp_type[kMaxComp];
    double ml0[kMaxComp], kp0[kMaxComp], t0[kMaxComp], ts[kMaxComp], h0[kMaxComp];
    double t0inv[kMaxComp];            // 1 / t0 (Sutherland)
    double hc[kMaxComp][8];            // enthalpy as a polynomial: h(T) = hH0 + T (hc0 + T (hc1 + ...)),
    double hH0[kMaxComp];              // hc_k = cp_k / (k + 1), hH0 = h0 - sum_k hc_k T_ref^(k+1)
    double sig[kMaxComp], ek[kMaxComp];
    double t_ref;
    int    use_visc, use_diff;
    int    n_react, chem_parity, chem_energy_molar;
    int    r_left[4][3], r_right[4][3];
    double r_a[4], r_e[4], r_n[4];
    double br2_eta;                    // BR2 penalty factor (0: BR1)
    float  br2_eta_f;
};

// c ? a : b on values (the dual-number overload selects component-wise, so
// that a select of two structs never becomes a load through a selected
// address in private memory)
template <typename T>
__device__ __forceinline__ T sel(bool c, const T &a, const T &b) { return c ? a : b; }

template <typename R> __device__ __forceinline__ R cst(double d, float) { return R(d); }
template <> __device__ __forceinline__ double cst<double>(double d, float) { return d; }
template <> __device__ __forceinline__ float cst<float>(double, float f) { return f; }

// Workgroups are dispatched round-robin over the XCDs (each with its own
// L2): workgroup b runs on XCD b % n_xcd.  Renumber so that every XCD walks a
// contiguous range of Morton-ordered chunks / element blocks and neighbours
// share its L2.  The grid is n_xcd * ceil(n / n_xcd) workgroups.
__device__ __forceinline__ int xcd_block(int n_xcd)
{
    const int per = (int) gridDim.x / n_xcd;
    return ((int) blockIdx.x % n_xcd) * per + (int) blockIdx.x / n_xcd;
}

// Coefficient (v, k) of a state U[(4 v + k) n + e] through a buffer
// resource per variable v: its base and size sit in scalar registers, the
// coefficient's plane offset k n sizeof(S) is a scalar offset, and a lane
// supplies only its 32-bit byte offset, so a load or store costs no 64-bit
// address arithmetic on the VALU.  The host keeps 4 n sizeof(S) below 2^31.
template <typename S>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t var_rsrc(const S *U, int64_t n, int v)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *) (U + 4 * v * n), 0, (int) (4 * n * (int64_t) sizeof(S)),
                                             0x00020000);
}

// i = 4 v + k
template <typename S>
__device__ __forceinline__ S ldp(const S *U, int64_t n, int i, int e)
{
    const __amdgpu_buffer_rsrc_t r = var_rsrc(U, n, i >> 2);
    const int so = (int) ((i & 3) * n * (int64_t) sizeof(S));
    if constexpr (sizeof(S) == 8)
        return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b64(r, e * 8, so, 0));
    else
        return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b32(r, e * 4, so, 0));
}

template <typename S>
__device__ __forceinline__ void stp(S *U, int64_t n, int i, int e, S x)
{
    const __amdgpu_buffer_rsrc_t r = var_rsrc<S>(U, n, i >> 2);
    const int so = (int) ((i & 3) * n * (int64_t) sizeof(S));
    if constexpr (sizeof(S) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, x), r,
                                              e * 8, so, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, e * 4, so, 0);
}

// Reciprocal, quotient and square root.  The FP32 bulk path uses the
// hardware v_rcp_f32 / v_sqrt_f32 (about 1 ulp; operands here are normal
// numbers), the FP64 path IEEE division and sqrt.
template <typename R> __device__ __forceinline__ R rcp(R x) { return R(1) / x; }
template <> __device__ __forceinline__ float rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <typename R> __device__ __forceinline__ R dv(R a, R b) { return a / b; }
template <> __device__ __forceinline__ float dv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename R> __device__ __forceinline__ R dsqrt(R x) { return sqrt(x); }
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
// clamp to [0, 1] (mass fractions); x < 0 -> 0, x > 1 -> 1 as two compares,
// one median-of-three on the FP32 path
template <typename R> __device__ __forceinline__ R clamp01(R x)
{
    x = x < R(0) ? R(0) : x;
    return x > R(1) ? R(1) : x;
}
template <> __device__ __forceinline__ float clamp01<float>(float x) { return __builtin_amdgcn_fmed3f(x, 0.f, 1.f); }

// cp of species i at T, charm_comp_calc_cp (src/charm_globals.c:603-625):
// the power series when cp_poly (cp_type polynomial with ml_type 1, as the
// reference's test reads), else cp[0]
__device__ __forceinline__ double comp_cp(const DevParams &P, int i, double t)
{
    if (!P.cp_poly[i]) return P.comp_cp[i];
    double r = 0.0, tt = 1.0;
    for (int k = 0; k < P.n_cp[i]; k++) { r += tt * P.cp_c[i][k]; tt *= t; }
    return r;
}

// d cp / dT, charm_comp_calc_cp_dt (src/charm_globals.c:628-649)
__device__ __forceinline__ double comp_cp_dt(const DevParams &P, int i, double t)
{
    double r = 0.0, tt = 1.0;
    for (int k = 1; k < P.n_cp[i]; k++) { r += tt * P.cp_c[i][k] * k; tt *= t; }
    return r;
}

// temperature of the polynomial-cp mixture by Newton from T = 1 K,
// (cp(T) - R sum c_i / M_i) T = e, at most 100 steps, stop at dT^2 < 1e-12
// (_calc_t, src/charm_eos.c:12-54); FP64 whatever the storage precision.
// The mixture's cp(T) = sum_i c_i cp_i(T) is one power series whose
// coefficients sum_i c_i a_ik are formed once, so that an iteration
// evaluates one polynomial and its derivative (Horner) instead of two per
// species; the iterates are the reference's up to rounding.  As
// charm_comp_calc_cp_dt, the derivative takes every species' stored series:
// a species whose cp is cp[0] (cp_poly false) adds its series' derivative
// in a branch uniform over the launch (P.cp_dt_rest).
template <int NC>
__device__ __forceinline__ double newton_t(const DevParams &P, double e, const double *c)
{
    double tt = 1.0, m_ = 0.0;
    #pragma unroll
    for (int i = 0; i < NC; i++) m_ += c[i] * P.comp_minv[i];
    const double rm = kGasR * m_;
    double a[8];
    #pragma unroll
    for (int k = 0; k < 8; k++) a[k] = 0.0;
    #pragma unroll
    for (int i = 0; i < NC; i++) {
        if (P.cp_poly[i]) {
            #pragma unroll
            for (int k = 0; k < 8; k++) a[k] = fma(c[i], P.cp_c[i][k], a[k]);
        } else {
            a[0] = fma(c[i], P.comp_cp[i], a[0]);
        }
    }
    for (int j = 0; j < 100; j++) {
        double cp = a[7], cp_dt = 7.0 * a[7];
        #pragma unroll
        for (int k = 6; k >= 0; k--) cp = fma(cp, tt, a[k]);
        #pragma unroll
        for (int k = 6; k >= 1; k--) cp_dt = fma(cp_dt, tt, k * a[k]);
        if (P.cp_dt_rest) {
            #pragma unroll
            for (int i = 0; i < NC; i++)
                if (!P.cp_poly[i]) cp_dt += c[i] * comp_cp_dt(P, i, tt);
        }
        const double ft = (cp - rm) * tt - e, ft_dt = (cp - rm) + cp_dt * tt;
        const double tt1 = tt - ft / ft_dt;
        const double d = tt1 - tt;
        tt = tt1;
        if (d * d < kEps) break;
    }
    return tt;
}

template <typename R, int NC>
struct Prim {
    R r, u, v, w, e_tot, p, cz, gam, e, cv, m;
    R c[NC];
};

// EOS closure at a given temperature (FP64): cp, M, cv, gamma of the ideal
// gas (component 0) or the mixture, src/charm_eos.c:98-148
template <typename R, int NC>
__device__ __forceinline__ void eos_coeffs_at(const DevParams &P, Prim<R, NC> &p, double t)
{
    double cp, m;
    if (P.eos == EOS_IDEAL) {
        cp = comp_cp(P, 0, t);
        m = P.comp_m[0];
    } else {
        double M_ = 0.0;
        cp = 0.0;
        #pragma unroll
        for (int i = 0; i < NC; i++) {
            M_ += (double) p.c[i] * P.comp_minv[i];
            cp += (double) p.c[i] * comp_cp(P, i, t);
        }
        m = 1.0 / M_;
    }
    const double cv = cp - kGasR / m;
    p.m = (R) m; p.cv = (R) cv; p.gam = (R) (cp / cv);
}

// EOS closure: gamma, cv, molar mass of the state (src/charm_eos.c:98-148).
// PH = 0: constant cp per species.  PH = 1: any cp; with a polynomial cp the
// temperature comes from the FP64 Newton solve on e (flags 1 and 4).
template <typename R, int NC, int PH = 0>
__device__ __forceinline__ void eos_coeffs(const DevParams &P, Prim<R, NC> &p)
{
    if constexpr (PH == 1) {
        double t = 0.0;
        if (P.newton_t) {
            double c[NC];
            #pragma unroll
            for (int i = 0; i < NC; i++) c[i] = (double) p.c[i];
            t = newton_t<NC>(P, (double) p.e, c);
        }
        eos_coeffs_at<R, NC>(P, p, t);
        return;
    }
    if constexpr (NC == 1) {
        p.gam = cst<R>(P.gam, P.gam_f); p.cv = cst<R>(P.cv, P.cv_f); p.m = cst<R>(P.m, P.m_f);
    } else if constexpr (sizeof(R) == 4) {
        // FP32 bulk path: the mixture closure in FP64 arithmetic (FP64
        // EOS), only its results rounded: M^-1 = sum c_i / M_i with the
        // host's FP64 reciprocals, R / M = R M^-1
        double cp = 0.0, minv = 0.0;
        #pragma unroll
        for (int i = 0; i < NC; i++) {
            minv += (double) p.c[i] * P.comp_minv[i];
            cp += (double) p.c[i] * P.comp_cp[i];
        }
        const double cv = cp - kGasR * minv;
        const bool ideal = P.eos == EOS_IDEAL;
        p.gam = sel(ideal, P.gam_f, (float) (cp / cv));
        p.cv = sel(ideal, P.cv_f, (float) cv);
        p.m = sel(ideal, P.m_f, (float) (1.0 / minv));
    } else {
        // the mixture values, then a select on the (uniform) EOS type: with
        // two branches writing the record the compiler kept both candidates
        // in private memory
        R cp = R(0), M_ = R(0);
        #pragma unroll
        for (int i = 0; i < NC; i++) {
            M_ += dv(p.c[i], cst<R>(P.comp_m[i], P.comp_m_f[i]));
            cp += p.c[i] * cst<R>(P.comp_cp[i], P.comp_cp_f[i]);
        }
        const R m = rcp(M_);
        const R cv = cp - dv(R(kGasR), m);
        const R gam = dv(cp, cv);
        const bool ideal = P.eos == EOS_IDEAL;
        p.gam = sel(ideal, cst<R>(P.gam, P.gam_f), gam);
        p.cv = sel(ideal, cst<R>(P.cv, P.cv_f), cv);
        p.m = sel(ideal, cst<R>(P.m, P.m_f), m);
    }
}

// Tangential Gauss-point offsets of the six-point face rule (two triangles
// (v0,v1,v2), (v1,v2,v3) of the face, barycentric (2/3,1/6,1/6)
// permutations, src/charm_geom.c:443-494) relative to the face centre in
// units of the element edge; t1 < t2 are the two tangential axes.  Literal
// switches so that the unrolled loops fold them into immediates.
template <typename R>
__device__ __forceinline__ R gp_t1(int k)
{
    switch (k) { case 0: return R(-1.0/3); case 1: return R(1.0/6); case 2: return R(-1.0/3);
                 case 3: return R(1.0/3); case 4: return R(-1.0/6); default: return R(1.0/3); }
}
template <typename R>
__device__ __forceinline__ R gp_t2(int k)
{
    switch (k) { case 0: return R(-1.0/3); case 1: return R(-1.0/3); case 2: return R(1.0/6);
                 case 3: return R(-1.0/6); case 4: return R(1.0/3); default: return R(1.0/3); }
}

// charm_param_cons_to_prim + EOS flag 4, src/charm_globals.c:183-208 and
// src/charm_eos.c:86-91: velocities and energy use the unfloored density,
// pressure and sound speed the floored one.
// tk (PH = 1 only, may be null): this state's EOS closure from the point
// pass, m, cv and gamma at tk[0], tk[qs], tk[2 qs], as this function's own
// closure computes them from the same trace (the FP64 Newton temperature
// and the mixture's cp at it: the same values, evaluated once per point)
template <typename R, int NC, int PH = 0>
__device__ __forceinline__ void cons_to_prim(const DevParams &P, const R *u, Prim<R, NC> &p,
                                             const R *tk = nullptr, int64_t qs = 0)
{
    R r = u[4];
    #pragma unroll
    for (int i = 1; i < NC; i++) r += u[4 + i];
    const R ri = rcp(r);
    p.u = u[0] * ri; p.v = u[1] * ri; p.w = u[2] * ri;
    p.e_tot = u[3] * ri;
    p.e = p.e_tot - R(0.5) * (p.u * p.u + p.v * p.v + p.w * p.w);
    #pragma unroll
    for (int i = 0; i < NC; i++) p.c[i] = clamp01(u[4 + i] * ri);
    if constexpr (PH == 1) {
        if (tk) { p.m = tk[0]; p.cv = tk[qs]; p.gam = tk[2 * qs]; }
        else eos_coeffs<R, NC, PH>(P, p);
    } else {
        eos_coeffs<R, NC, PH>(P, p);
    }
    if (r < R(kEps)) r = R(kEps);
    p.r = r;
    const R gm1 = p.gam - R(1);
    p.p = r * p.e * gm1;
    // c^2 = gamma p / r = gamma (gamma - 1) e with the same (floored) r
    p.cz = dsqrt(p.gam * gm1 * p.e);
}

// Exterior state of a boundary face, src/charm_bnd_cond.c:21-93: the BC
// function edits (u, v, w, T, p), then EOS flag 3 rebuilds (r, cz, e) from
// (T, p) and e_tot is recomputed.  n = s * e_axis (axis-aligned octree).
// With frame_axis >= 0 the velocity components are stored in the rotated
// order (frame_axis, t1, t2) and the inlet velocity is rotated alike.
template <typename R, int NC, int PH = 0>
__device__ __forceinline__ void bnd_state(const DevParams &P, int face, int axis, R s,
                                          const Prim<R, NC> &in, Prim<R, NC> &out, int frame_axis)
{
    out = in;
    R t = dv(in.e, in.cv);               // T from flag 4 (src/charm_eos.c:90)
    R pr = in.p;
    R vel[3] = {in.u, in.v, in.w};
    const int type = P.bnd_type[face];
    if (type == BND_INLET) {
        R pv[3];
        #pragma unroll
        for (int i = 0; i < 3; i++) pv[i] = cst<R>(P.bnd_param[face][i], P.bnd_param_f[face][i]);
        if (frame_axis >= 0) {
            const int a = frame_axis;
            vel[0] = sel(a == 0, pv[0], sel(a == 1, pv[1], pv[2]));
            vel[1] = sel(a == 0, pv[1], pv[0]);
            vel[2] = sel(a == 2, pv[1], pv[2]);
        } else {
            vel[0] = pv[0]; vel[1] = pv[1]; vel[2] = pv[2];
        }
        t = cst<R>(P.bnd_param[face][3], P.bnd_param_f[face][3]);
        pr = cst<R>(P.bnd_param[face][4], P.bnd_param_f[face][4]);
    } else if (type == BND_WALL_SLIP) {
        R va = sel(axis == 0, vel[0], sel(axis == 1, vel[1], vel[2]));
        R vn = va * s;
        va = (va - s * vn) - s * vn;
        vel[0] = sel(axis == 0, va, vel[0]);
        vel[1] = sel(axis == 1, va, vel[1]);
        vel[2] = sel(axis == 2, va, vel[2]);
    } else if (type == BND_WALL_NO_SLIP) {
        R va = sel(axis == 0, vel[0], sel(axis == 1, vel[1], vel[2]));
        R vn = va * s;
        va = va - s * vn;
        vel[0] = sel(axis == 0, va, vel[0]);
        vel[1] = sel(axis == 1, va, vel[1]);
        vel[2] = sel(axis == 2, va, vel[2]);
        t = cst<R>(P.bnd_param[face][0], P.bnd_param_f[face][0]);
    } else if (type == BND_MASS_FLOW) {
        // src/charm_bnd_cond.c:101-146: param = (mass flux, static p, total T,
        // direction cosines, outflow p, composition).  Outflow (the direction
        // leaves through the face): the pressure param[6] where the normal
        // velocity is subsonic.  The reference also writes the interior
        // pressure of a supersonic face back into param[6], which later faces
        // read; that depends on the face-visit order, so param[6] stays fixed
        // here.  Inflow: the composition of the parameters, EOS flag 0 at the
        // interior T, the static pressure if supersonic, T from
        // T + (m (cp - cv) T / (p f))^2 / (2 cp) = T_tot, speed m / (rho f).
        R cs[3];
        #pragma unroll
        for (int i = 0; i < 3; i++) cs[i] = cst<R>(P.bnd_param[face][3 + i], P.bnd_param_f[face][3 + i]);
        R cf0 = cs[0], cf1 = cs[1], cf2 = cs[2];          // in the state's component order
        if (frame_axis >= 0) {
            const int a = frame_axis;
            cf0 = sel(a == 0, cs[0], sel(a == 1, cs[1], cs[2]));
            cf1 = sel(a == 0, cs[1], cs[0]);
            cf2 = sel(a == 2, cs[1], cs[2]);
        }
        const R fd = s * sel(axis == 0, cf0, sel(axis == 1, cf1, cf2));      // n . direction
        const R vn = s * sel(axis == 0, vel[0], sel(axis == 1, vel[1], vel[2]));
        const R mflow = cst<R>(P.bnd_param[face][0], P.bnd_param_f[face][0]);
        if (fd > R(0)) {
            if (vn < in.cz) pr = cst<R>(P.bnd_param[face][6], P.bnd_param_f[face][6]);
        } else {
            Prim<R, NC> q = in;
            #pragma unroll
            for (int i = 0; i < NC; i++) q.c[i] = cst<R>(P.bnd_param[face][7 + i], P.bnd_param_f[face][7 + i]);
            if constexpr (PH == 1) eos_coeffs_at<R, NC>(P, q, (double) t);
            else eos_coeffs<R, NC>(P, q);
            const R r = in.r < R(kEps) ? R(kEps) : in.r;
            const R p0 = r * in.e * (q.gam - R(1));
            const R v2 = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
            pr = v2 > dv(q.gam * p0, r) ? cst<R>(P.bnd_param[face][1], P.bnd_param_f[face][1]) : p0;
            const R cv = q.cv, cp = q.gam * cv;
            const R x = dv(mflow * (cp - cv), pr * fd);
            const R k = dv(x * x, R(2) * cp);
            const R ttot = cst<R>(P.bnd_param[face][2], P.bnd_param_f[face][2]);
            t = dv(R(-1) + dsqrt(R(1) + R(4) * k * ttot), R(2) * k);
            const R vm = fabs(dv(mflow, r * fd));
            vel[0] = cf0 * vm; vel[1] = cf1 * vm; vel[2] = cf2 * vm;
            #pragma unroll
            for (int i = 0; i < NC; i++) out.c[i] = q.c[i];
        }
    }
    out.u = vel[0]; out.v = vel[1]; out.w = vel[2];
    if constexpr (PH == 1) eos_coeffs_at<R, NC>(P, out, (double) t);   // flag 3: T given
    else eos_coeffs<R, NC>(P, out);
    if (pr < R(kEps)) pr = R(kEps);
    out.p = pr;
    out.r = dv(pr * out.m, t * R(kGasR));
    out.cz = dsqrt(dv(out.gam * pr, out.r));
    out.e = dv(pr, out.r * (out.gam - R(1)));
    out.e_tot = out.e + R(0.5) * (out.u * out.u + out.v * out.v + out.w * out.w);
}

// ----------------------------------------------------------------- fluxes
// A flux evaluation works in the face frame (un, ut1, ut2): for the
// axis-aligned normal n = s e_a the reference's tangent choice
// (src/charm_fluxes.c:458-488) is a signed permutation, and every
// tangential flux is linear in its tangential velocity, so evaluating with
// the unsigned tangential components gives the reference's result exactly.
template <typename R, int NC>
struct Frame {
    R r, un, ut1, ut2, e_tot, p, cz, gam;
    R c[NC];
};

template <typename R, int NC>
__device__ __forceinline__ Frame<R, NC> to_frame(const Prim<R, NC> &p, R n0, R n1, R n2,
                                                 R t10, R t11, R t12, R t20, R t21, R t22)
{
    Frame<R, NC> f;
    f.r = p.r; f.e_tot = p.e_tot; f.p = p.p; f.cz = p.cz; f.gam = p.gam;
    f.un = p.u * n0 + p.v * n1 + p.w * n2;
    f.ut1 = p.u * t10 + p.v * t11 + p.w * t12;
    f.ut2 = p.u * t20 + p.v * t21 + p.w * t22;
    #pragma unroll
    for (int i = 0; i < NC; i++) f.c[i] = p.c[i];
    return f;
}

// field-wise select of two face-frame states (a select of the two records'
// addresses would send them to private memory)
template <typename R, int NC>
__device__ __forceinline__ Frame<R, NC> sel_frame(bool c, const Frame<R, NC> &a, const Frame<R, NC> &b)
{
    Frame<R, NC> o;
    o.r = sel(c, a.r, b.r); o.un = sel(c, a.un, b.un); o.ut1 = sel(c, a.ut1, b.ut1); o.ut2 = sel(c, a.ut2, b.ut2);
    o.e_tot = sel(c, a.e_tot, b.e_tot); o.p = sel(c, a.p, b.p); o.cz = sel(c, a.cz, b.cz); o.gam = sel(c, a.gam, b.gam);
    #pragma unroll
    for (int i = 0; i < NC; i++) o.c[i] = sel(c, a.c[i], b.c[i]);
    return o;
}

// the reference's face frame for an arbitrary unit normal (src/charm_fluxes.c:
// 458-488): rows n, t1 (the first non-degenerate of the three in-plane
// choices), t2 = n x t1
__device__ __forceinline__ void ref_face_frame(const double *nn, double nt[3][3])
{
    nt[0][0] = nn[0]; nt[0][1] = nn[1]; nt[0][2] = nn[2];
    nt[1][0] = nt[1][1] = nt[1][2] = 0.0;
    double ri = sqrt(nn[0] * nn[0] + nn[1] * nn[1]);
    if (ri > kEps) { nt[1][0] = -nn[1] / ri; nt[1][1] = nn[0] / ri; }
    else {
        ri = sqrt(nn[1] * nn[1] + nn[2] * nn[2]);
        if (ri > kEps) { nt[1][1] = -nn[2] / ri; nt[1][2] = nn[1] / ri; }
        else { ri = sqrt(nn[0] * nn[0] + nn[2] * nn[2]); if (ri > kEps) { nt[1][0] = -nn[2] / ri; nt[1][2] = nn[0] / ri; } }
    }
    nt[2][0] = nt[0][1] * nt[1][2] - nt[0][2] * nt[1][1];
    nt[2][1] = -nt[0][0] * nt[1][2] + nt[0][2] * nt[1][0];
    nt[2][2] = nt[0][0] * nt[1][1] - nt[0][1] * nt[1][0];
}

// x-split HLLC, src/charm_fluxes.c:319-437; q = (qn, qt1, qt2, qe, qc[]).
// Returns false where the reference aborts (:354-357).
template <typename R, int NC>
__device__ __forceinline__ bool hllc_x(const Frame<R, NC> &L, const Frame<R, NC> &Rt, R *q)
{
    R p_pvrs = R(0.5) * (L.p + Rt.p) - R(0.5) * (Rt.un - L.un) * R(0.25) * (L.r + Rt.r) * (L.cz + Rt.cz);
    R p_star = p_pvrs > R(0) ? p_pvrs : R(0);
    // (gamma + 1) / (2 gamma) folded to one factor per side
    const R gl = dv(L.gam + R(1), R(2) * L.gam), gr = dv(Rt.gam + R(1), R(2) * Rt.gam);
    R ql = (p_star <= L.p) ? R(1) : dsqrt(R(1) + gl * (dv(p_star, L.p) - R(1)));
    R qr = (p_star <= Rt.p) ? R(1) : dsqrt(R(1) + gr * (dv(p_star, Rt.p) - R(1)));
    R sl = L.un - L.cz * ql;
    R sr = Rt.un + Rt.cz * qr;
    if (sl > sr) { R t = sl; sl = sr; sr = t; }
    R s_star = Rt.p - L.p;
    s_star += L.r * L.un * (sl - L.un);
    s_star -= Rt.r * Rt.un * (sr - Rt.un);
    s_star = dv(s_star, L.r * (sl - L.un) - Rt.r * (sr - Rt.un));
    if (s_star < sl) s_star = sl;
    if (s_star > sr) s_star = sr;
    const bool ok = (sl <= s_star) && (s_star <= sr);
    if (sl >= R(0) || sr <= R(0)) {
        const Frame<R, NC> K = sel_frame<R, NC>(sl >= R(0), L, Rt);
        q[0] = K.r * K.un * K.un + K.p;
        q[1] = K.r * K.ut1 * K.un;
        q[2] = K.r * K.ut2 * K.un;
        q[3] = (K.r * K.e_tot + K.p) * K.un;
        #pragma unroll
        for (int i = 0; i < NC; i++) q[4 + i] = K.r * K.un * K.c[i];
    } else {
        const bool left = s_star >= R(0);
        const Frame<R, NC> K = sel_frame<R, NC>(left, L, Rt);
        const R sk = left ? sl : sr;
        // F_HLLC_U / _V / _E macros, src/charm_fluxes.c:319-321
        const R inv = rcp(sk - s_star);
        const R pk = K.p + K.r * (sk - K.un) * (s_star - K.un);
        q[0] = (s_star * (sk * (K.r * K.un) - (K.r * K.un * K.un + K.p)) + sk * pk) * inv;
        q[1] = (s_star * (sk * (K.r * K.ut1) - K.r * K.un * K.ut1)) * inv;
        q[2] = (s_star * (sk * (K.r * K.ut2) - K.r * K.un * K.ut2)) * inv;
        q[3] = (s_star * (sk * (K.r * K.e_tot) - (K.r * K.e_tot + K.p) * K.un) + sk * pk * s_star) * inv;
        #pragma unroll
        for (int i = 0; i < NC; i++)
            q[4 + i] = (s_star * (sk * (K.r * K.c[i]) - K.r * K.c[i] * K.un)) * inv;
    }
    return ok;
}

// Lax-Friedrichs in the face frame, src/charm_fluxes.c:255-315
template <typename R, int NC>
__device__ __forceinline__ void lf_x(const Frame<R, NC> &L, const Frame<R, NC> &Rt, R *q, bool central)
{
    const R alpha = central ? R(0) : fmax(fabs(L.un) + L.cz, fabs(Rt.un) + Rt.cz);
    const R fL0 = L.r * L.un * L.un + L.p, fR0 = Rt.r * Rt.un * Rt.un + Rt.p;
    q[0] = R(0.5) * (fR0 + fL0 - alpha * (Rt.r * Rt.un - L.r * L.un));
    q[1] = R(0.5) * (Rt.r * Rt.un * Rt.ut1 + L.r * L.un * L.ut1 - alpha * (Rt.r * Rt.ut1 - L.r * L.ut1));
    q[2] = R(0.5) * (Rt.r * Rt.un * Rt.ut2 + L.r * L.un * L.ut2 - alpha * (Rt.r * Rt.ut2 - L.r * L.ut2));
    q[3] = R(0.5) * ((Rt.r * Rt.e_tot + Rt.p) * Rt.un + (L.r * L.e_tot + L.p) * L.un
                     - alpha * (Rt.r * Rt.e_tot - L.r * L.e_tot));
    #pragma unroll
    for (int i = 0; i < NC; i++)
        q[4 + i] = R(0.5) * (Rt.r * Rt.un * Rt.c[i] + L.r * L.un * L.c[i]
                             - alpha * (Rt.r * Rt.c[i] - L.r * L.c[i]));
}

// Exact Riemann solver (Newton on p* and sampling at xi = 0 of rim_orig,
// src/charm_fluxes.c:8-180) with one gamma per side; the flux is built as in
// the disabled body of charm_calc_flux_godunov (:184-252).
template <typename R, int NC>
__device__ __forceinline__ void godunov_x(const Frame<R, NC> &Lf, const Frame<R, NC> &Rf, R *q,
                                          int max_steps)
{
    const R RB = Lf.r, PB = Lf.p, UB = Lf.un, RE = Rf.r, PE = Rf.p, UE = Rf.un;
    const R gb = Lf.gam, ge = Rf.gam;
    const R AB = gb - R(1), AE = ge - R(1);
    const R DB = dv(R(2), AB), DE = dv(R(2), AE);
    const R GB = dsqrt(gb * AB), GE = dsqrt(ge * AE);
    const R HB = (AB * R(0.5)), HE = (AE * R(0.5));
    const R FB = R(3) * gb - R(1), FE = R(3) * ge - R(1);
    const R OB = dv(AB, R(2) * gb), OE = dv(AE, R(2) * ge);
    const R QB = gb + R(1), QE = ge + R(1);
    const R PGB = dv(QB, R(2) * gb), PGE = dv(QE, R(2) * ge);
    const R RGB = R(4) * gb, RGE = R(4) * ge;
    const R SB = gb * AB, SE = ge * AE;
    const R TB = (QB * R(0.5)), TE = (QE * R(0.5));
    const R eps = R(1.e-5);
    const R CB = dsqrt(dv(gb * PB, RB)), CE = dsqrt(dv(ge * PE, RE));
    // The reference stops Newton when |DU - F1 - F2| <= 1e-5 (absolute, in
    // velocity units).  FP64 keeps that test.  In FP32 the residual of
    // states with speeds ~1e3 cannot get below a few ulp of the speed scale,
    // so the test is max(1e-5, 8 ulp of (cB + cE + |uB| + |uE|)); Newton's
    // quadratic convergence then ends in a few steps instead of the bound.
    const R tol = sizeof(R) == 8 ? eps
                                 : fmax(eps, R(8 * 1.1920929e-7) * (CB + CE + fabs(UB) + fabs(UE)));
    const R EB = dv(CB * CB, SB), EE = dv(CE * CE, SE);
    const R RCB = RB * CB, RCE = RE * CE;
    const R DU = UB - UE;
    R RF = 0, RS = 0, EF = 0, ES = 0, UF = 0, US = 0, SBL, SFL, SSL, SEL;
    if (DU < -(dv(R(2) * CB, AB) + dv(R(2) * CE, AE))) {
        SBL = UB - CB; SFL = UB + dv(R(2) * CB, AB); SSL = UE - dv(R(2) * CE, AE); SEL = UE + CE;
    } else {
        R P = dv(PB * RCE + PE * RCB + DU * RCB * RCE, RCB + RCE);
        R F1, F2, FS1, FS2;
        int step = 0;
        do {
            if (P < eps) P = eps;
            const R PPB = dv(P, PB), PPE = dv(P, PE);
            if (PB <= P) {
                const R PKB = PGB * PPB + OB, ZNB = RCB * dsqrt(PKB);
                F1 = dv(P - PB, ZNB);
                FS1 = dv(QB * PPB + FB, RGB * ZNB * PKB);
            } else {
                const R ZFB = CB * exp(log(PPB) * OB);
                F1 = DB * (ZFB - CB);
                FS1 = dv(ZFB, gb * P);
            }
            if (PE <= P) {
                const R PKE = PGE * PPE + OE, ZNE = RCE * dsqrt(PKE);
                F2 = dv(P - PE, ZNE);
                FS2 = dv(QE * PPE + FE, RGE * ZNE * PKE);
            } else {
                const R ZFE = CE * exp(log(PPE) * OE);
                F2 = DE * (ZFE - CE);
                FS2 = dv(ZFE, ge * P);
            }
            P = P + dv(DU - F1 - F2, FS1 + FS2);
        } while ((fabs(DU - F1 - F2) > tol) && (++step < max_steps));
        const R PPB = dv(P, PB), PPE = dv(P, PE);
        const R ZFB = CB * exp(log(PPB) * OB), ZFE = CE * exp(log(PPE) * OE);
        if (PB <= P) {
            const R D = UB - dsqrt(dv(TB * P + HB * PB, RB));
            const R UBD = UB - D, RUBD = RB * UBD;
            RF = dv(RUBD * RUBD, PB - P + RUBD * UBD);
            UF = D + dv(RUBD, RF);
            EF = dv(P, AB * RF);
            SBL = D; SFL = D;
        } else {
            EF = dv(ZFB * ZFB, SB);
            UF = UB + DB * (CB - ZFB);
            RF = dv(P, AB * EF);
            SBL = UB - CB; SFL = UF - ZFB;
        }
        if (PE <= P) {
            const R D = UE + dsqrt(dv(TE * P + HE * PE, RE));
            const R UED = UE - D, RUED = RE * UED;
            RS = dv(RUED * RUED, PE - P + RUED * UED);
            US = D + dv(RUED, RS);
            ES = dv(P, AE * RS);
            SEL = D; SSL = D;
        } else {
            ES = dv(ZFE * ZFE, SE);
            US = UE - DE * (CE - ZFE);
            RS = dv(P, AE * ES);
            SSL = US + ZFE; SEL = UE + CE;
        }
    }
    R RI, EI, PI, UI, VI, WI;
    int side;
    if (SEL <= R(0)) {
        RI = RE; EI = EE; UI = UE; VI = Rf.ut1; WI = Rf.ut2; PI = AE * EI * RI; side = 1;
    } else if (SBL >= R(0)) {
        RI = RB; EI = EB; UI = UB; VI = Lf.ut1; WI = Lf.ut2; PI = AB * EI * RI; side = 0;
    } else if ((SSL >= R(0)) && (SFL <= R(0))) {
        if (US >= R(0)) { RI = RF; EI = EF; UI = UF; VI = Lf.ut1; WI = Lf.ut2; PI = AB * EI * RI; side = 0; }
        else { RI = RS; EI = ES; UI = US; VI = Rf.ut1; WI = Rf.ut2; PI = AE * EI * RI; side = 1; }
    } else if (SFL > R(0)) {
        UI = dv(UB + DB * GB * dsqrt(EB), R(1) + DB);
        VI = Lf.ut1; WI = Lf.ut2;
        EI = dv(UI * UI, SB);
        RI = RB * exp(log(dv(EI, EB)) * rcp(AB));
        PI = AB * EI * RI; side = 0;
    } else {
        UI = dv(UE - DE * GE * dsqrt(EE), R(1) + DE);
        VI = Rf.ut1; WI = Rf.ut2;
        EI = dv(UI * UI, SE);
        RI = RE * exp(log(dv(EI, EE)) * rcp(AE));
        PI = AE * EI * RI; side = 1;
    }
    const R qr = RI * UI;
    q[0] = qr * UI + PI;
    q[1] = qr * VI;
    q[2] = qr * WI;
    q[3] = (RI * (EI + R(0.5) * (UI * UI + VI * VI + WI * WI)) + PI) * UI;
    #pragma unroll
    for (int i = 0; i < NC; i++) q[4 + i] = qr * (side ? Rf.c[i] : Lf.c[i]);
}

template <typename R, int NC, int FLUX>
__device__ __forceinline__ bool flux_x(const Frame<R, NC> &L, const Frame<R, NC> &Rt, R *q)
{
    if constexpr (FLUX == FLUX_HLLC) {
        return hllc_x<R, NC>(L, Rt, q);
    } else if constexpr (FLUX == FLUX_GODUNOV) {
        godunov_x<R, NC>(L, Rt, q, sizeof(R) == 8 ? 5000 : 64);
        return true;
    } else {
        lf_x<R, NC>(L, Rt, q, FLUX == FLUX_CD);
        return true;
    }
}

} // namespace charm
