// charm_dual.hpp -- forward-mode dual numbers for the adjoint kernels.
//
// Dual<R, N> carries a value and N tangents.  The pointwise device numerics
// (cons_to_prim, the boundary states, the HLLC / LF fluxes) are templates on
// the arithmetic type, so instantiating them with Dual gives the exact
// derivative of the same branch the forward code takes at that point.  The
// adjoint kernels seed the tangents with the inputs of one point (its state
// values, or the inlet parameters) and read the gradient of a scalar
// contraction of the outputs.
#pragma once
#include <hip/hip_runtime.h>
#include "charm_device.hpp"

namespace charm {

template <typename R, int N>
struct Dual {
    R v;
    R d[N];
    __device__ __forceinline__ Dual() {}
    __device__ __forceinline__ Dual(double x) : v(R(x))
    {
        #pragma unroll
        for (int i = 0; i < N; i++) d[i] = R(0);
    }
    __device__ __forceinline__ Dual(float x) : v(R(x))
    {
        #pragma unroll
        for (int i = 0; i < N; i++) d[i] = R(0);
    }
    __device__ __forceinline__ Dual(int x) : v(R(x))
    {
        #pragma unroll
        for (int i = 0; i < N; i++) d[i] = R(0);
    }
    // a variable: value x, tangent e_k
    static __device__ __forceinline__ Dual var(R x, int k)
    {
        Dual r(x);
        #pragma unroll
        for (int i = 0; i < N; i++) r.d[i] = (i == k) ? R(1) : R(0);
        return r;
    }
    __device__ __forceinline__ Dual &operator+=(const Dual &b)
    {
        v += b.v;
        #pragma unroll
        for (int i = 0; i < N; i++) d[i] += b.d[i];
        return *this;
    }
    __device__ __forceinline__ Dual &operator-=(const Dual &b)
    {
        v -= b.v;
        #pragma unroll
        for (int i = 0; i < N; i++) d[i] -= b.d[i];
        return *this;
    }
    __device__ __forceinline__ Dual &operator*=(const Dual &b)
    {
        #pragma unroll
        for (int i = 0; i < N; i++) d[i] = d[i] * b.v + v * b.d[i];
        v *= b.v;
        return *this;
    }
};

// component-wise select (see sel in charm_device.hpp)
template <typename R, int N>
__device__ __forceinline__ Dual<R, N> sel(bool c, const Dual<R, N> &a, const Dual<R, N> &b)
{
    Dual<R, N> o;
    o.v = c ? a.v : b.v;
    #pragma unroll
    for (int i = 0; i < N; i++) o.d[i] = c ? a.d[i] : b.d[i];
    return o;
}


#define CHARM_DUAL template <typename R, int N> __device__ __forceinline__

CHARM_DUAL Dual<R, N> operator-(const Dual<R, N> &a)
{
    Dual<R, N> r;
    r.v = -a.v;
    #pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
    return r;
}
CHARM_DUAL Dual<R, N> operator+(const Dual<R, N> &a, const Dual<R, N> &b)
{
    Dual<R, N> r;
    r.v = a.v + b.v;
    #pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i];
    return r;
}
CHARM_DUAL Dual<R, N> operator-(const Dual<R, N> &a, const Dual<R, N> &b)
{
    Dual<R, N> r;
    r.v = a.v - b.v;
    #pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i];
    return r;
}
CHARM_DUAL Dual<R, N> operator*(const Dual<R, N> &a, const Dual<R, N> &b)
{
    Dual<R, N> r;
    r.v = a.v * b.v;
    #pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
CHARM_DUAL Dual<R, N> operator/(const Dual<R, N> &a, const Dual<R, N> &b)
{
    Dual<R, N> r;
    const R inv = R(1) / b.v;
    r.v = a.v * inv;
    #pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
    return r;
}
CHARM_DUAL bool operator<(const Dual<R, N> &a, const Dual<R, N> &b) { return a.v < b.v; }
CHARM_DUAL bool operator>(const Dual<R, N> &a, const Dual<R, N> &b) { return a.v > b.v; }
CHARM_DUAL bool operator<=(const Dual<R, N> &a, const Dual<R, N> &b) { return a.v <= b.v; }
CHARM_DUAL bool operator>=(const Dual<R, N> &a, const Dual<R, N> &b) { return a.v >= b.v; }

CHARM_DUAL Dual<R, N> sqrt(const Dual<R, N> &a)
{
    Dual<R, N> r;
    r.v = dsqrt(a.v);
    const R h = R(0.5) / r.v;
    #pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * h;
    return r;
}
CHARM_DUAL Dual<R, N> fabs(const Dual<R, N> &a) { return a.v < R(0) ? -a : a; }
CHARM_DUAL Dual<R, N> fmax(const Dual<R, N> &a, const Dual<R, N> &b) { return (b.v > a.v) ? b : a; }
CHARM_DUAL Dual<R, N> fmin(const Dual<R, N> &a, const Dual<R, N> &b) { return (b.v < a.v) ? b : a; }

#undef CHARM_DUAL

} // namespace charm
