// charm_elemfaces.hpp -- element-centric access to a leaf's face points
// through the neighbour table (shared by the Navier-Stokes and adjoint
// kernels): one thread per leaf walks its six face slots (conforming, the
// small side of a hanging face, the big side with its four small
// neighbours, boundary) and evaluates every face point itself.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "charm_device.hpp"

namespace charm {

constexpr int kNbrAuxBase = 8;        // nbr <= -8: group of four small neighbours

template <typename R, int NV>
__device__ __forceinline__ void load_coeffs(const R *__restrict__ U, int64_t n, int e, R *u)
{
    #pragma unroll
    for (int i = 0; i < 4 * NV; i++) u[i] = (U + i * n)[e];
}

// value of a P1 Taylor field (4 coefficients) at reference coordinates
template <typename R>
__device__ __forceinline__ R at(const R *c, R px, R py, R pz)
{
    return c[0] + px * c[1] + py * c[2] + pz * c[3];
}

// component a (0, 1, 2; known at run time) of three values, by selects
template <typename R>
__device__ __forceinline__ R pick3(int a, R v0, R v1, R v2)
{
    return a == 0 ? v0 : (a == 1 ? v1 : v2);
}

// reference coordinates of face point k on the face of direction d of a leaf
// (normal coordinate +-1/2, tangential offsets off + scale * gp); xyz order
template <typename R>
__device__ __forceinline__ void face_point(int d, int k, R o1, R o2, R scale, R ph[3])
{
    // d may differ between lanes: selects, not a store through ph[a]
    const int a = d >> 1;
    const R pn = (d & 1) ? R(0.5) : R(-0.5);
    const R p1 = o1 + scale * gp_t1<R>(k), p2 = o2 + scale * gp_t2<R>(k);
    ph[0] = (a == 0) ? pn : p1;                           // t1 = 0 unless a = 0
    ph[1] = (a == 1) ? pn : ((a == 0) ? p1 : p2);
    ph[2] = (a == 2) ? pn : p2;                           // t2 = 2 unless a = 2
}

// Enumerates the face points of leaf e: fn(d, nb, self_ph, nb_ph, wj, sign,
// k, js, jn) for interior points (d the face direction of e; sign +1 when e
// is side 0 of the face, i.e. its face lies on the upper side of axis d / 2;
// point k of the (sub)face numbered js among the sub-faces of e's face d and
// jn among those of nb's face d ^ 1, 0 on a face that is not split on that
// side) and bfn(d, bc_face, k, self_ph, wj) for boundary faces.  Hanging
// faces are visited at the small face's points.
template <typename R, typename F, typename B>
__device__ __forceinline__ void face_points(const int *__restrict__ nbr, const int4 *__restrict__ aux,
                                            const uint8_t *__restrict__ level, int e, R h, F fn, B bfn)
{
    const int lev = level[e];
    for (int d = 0; d < 6; d++) {
        const int nb = nbr[6 * e + d];
        const R sign = (d & 1) ? R(1) : R(-1);
        if (nb >= 0) {
            R o1 = R(0), o2 = R(0), m = R(1);
            int jn = 0;
            if (level[nb] != lev) {
                // e is sub-face jn of the coarser neighbour's face
                const int4 grp = aux[-kNbrAuxBase - nbr[6 * nb + (d ^ 1)]];
                jn = (grp.x == e) ? 0 : ((grp.y == e) ? 1 : ((grp.z == e) ? 2 : 3));
                o1 = (jn & 1) ? R(0.25) : R(-0.25);
                o2 = (jn & 2) ? R(0.25) : R(-0.25);
                m = R(0.5);
            }
            const R wj = h * h * R(1.0 / 6.0);
            for (int k = 0; k < 6; k++) {
                R ps[3], pn[3];
                face_point<R>(d, k, R(0), R(0), R(1), ps);
                face_point<R>(d ^ 1, k, o1, o2, m, pn);
                fn(d, nb, ps, pn, wj, sign, k, 0, jn);
            }
        } else if (nb > -kNbrAuxBase) {
            const R wj = h * h * R(1.0 / 6.0);
            for (int k = 0; k < 6; k++) {
                R ps[3];
                face_point<R>(d, k, R(0), R(0), R(1), ps);
                bfn(d, -1 - nb, k, ps, wj);
            }
        } else {
            const int4 grp = aux[-kNbrAuxBase - nb];
            const int q4[4] = {grp.x, grp.y, grp.z, grp.w};
            const R wj = R(0.25) * h * h * R(1.0 / 6.0);
            for (int j = 0; j < 4; j++) {
                const R o1 = (j & 1) ? R(0.25) : R(-0.25), o2 = (j & 2) ? R(0.25) : R(-0.25);
                for (int k = 0; k < 6; k++) {
                    R ps[3], pn[3];
                    face_point<R>(d, k, o1, o2, R(0.5), ps);
                    face_point<R>(d ^ 1, k, R(0), R(0), R(1), pn);
                    fn(d, q4[j], ps, pn, wj, sign, k, j, 0);
                }
            }
        }
    }
}

// Enumerates the interior (sub)faces of leaf e, one call per (sub)face:
// gfn(d, nb, so1, so2, sm, no1, no2, nm, wj, sign, js, jn).  Point k of the
// (sub)face is face_point(d, k, so1, so2, sm) on e and face_point(d ^ 1, k,
// no1, no2, nm) on nb; wj and sign as in face_points.  js / jn number the
// (sub)face among the sub-faces of e's face d / nb's face d ^ 1 (0 on a face
// that is not split on that side).  Boundary faces are skipped.  The
// (sub)faces and their points are those face_points visits.
template <typename R, typename G>
__device__ __forceinline__ void face_subfaces(const int *__restrict__ nbr, const int4 *__restrict__ aux,
                                              const uint8_t *__restrict__ level, int e, R h, G gfn)
{
    const int lev = level[e];
    for (int d = 0; d < 6; d++) {
        const int nb = nbr[6 * e + d];
        const R sign = (d & 1) ? R(1) : R(-1);
        if (nb >= 0) {
            R o1 = R(0), o2 = R(0), m = R(1);
            int j = 0;
            if (level[nb] != lev) {
                const int4 grp = aux[-kNbrAuxBase - nbr[6 * nb + (d ^ 1)]];
                j = (grp.x == e) ? 0 : ((grp.y == e) ? 1 : ((grp.z == e) ? 2 : 3));
                o1 = (j & 1) ? R(0.25) : R(-0.25);
                o2 = (j & 2) ? R(0.25) : R(-0.25);
                m = R(0.5);
            }
            gfn(d, nb, R(0), R(0), R(1), o1, o2, m, h * h * R(1.0 / 6.0), sign, 0, j);
        } else if (nb <= -kNbrAuxBase) {
            const int4 grp = aux[-kNbrAuxBase - nb];
            for (int j = 0; j < 4; j++) {
                // by selects: the loop need not unroll around a large body
                const int qj = j == 0 ? grp.x : (j == 1 ? grp.y : (j == 2 ? grp.z : grp.w));
                const R o1 = (j & 1) ? R(0.25) : R(-0.25), o2 = (j & 2) ? R(0.25) : R(-0.25);
                gfn(d, qj, o1, o2, R(0.5), R(0), R(0), R(1), R(0.25) * h * h * R(1.0 / 6.0), sign, j, 0);
            }
        }
    }
}

} // namespace charm
