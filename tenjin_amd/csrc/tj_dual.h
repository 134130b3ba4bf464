// Forward-mode dual numbers for the tangent-linear (differentiated) form of
// the solid path: a value and K directional derivatives.  The value part of
// every operation is the same IEEE double operation the FP64 form performs,
// in the same order, so a tangent run reproduces the FP64 trajectory bit for
// bit; the derivative parts follow the chain rule of each operation.
//
// The derivative parts are not bound to the value's rounding: they use fused
// multiply-adds and a reciprocal refined from the hardware estimate (two
// Newton steps, a few ulp) instead of a second IEEE division.  The MUSCL
// slope takes its derivative from the limiter's closed form (muscl1 below),
// not through the ratio of differences, which ahead of a wave is a ratio of
// exponentially small precursor values whose derivative overflows.
//
// Non-smooth points take one one-sided derivative: |x| and max/min follow the
// branch the value takes, sqrt at 0 has derivative 0 (it is reached only at a
// minimum of its argument, HJ2 = 0 of an undeformed cell), and the MUSCL
// limiter's inf/NaN reset zeroes the derivative with the value.
#pragma once
#include "tj_device.h"
#include "tj_fluid.h"

namespace tj {

template <int K>
struct Dual {
    double v;
    double d[K];
    Dual() = default;
    __host__ __device__ __forceinline__ Dual(double x) : v(x) {
#pragma unroll
        for (int t = 0; t < K; ++t) d[t] = 0.0;
    }
};

// 1/b for derivative parts: v_rcp_f64's estimate and two Newton steps
__device__ __forceinline__ double drcp(double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    return __builtin_fma(r, e, r);
}

#define TJ_DUAL_FN template <int K> __device__ __forceinline__

TJ_DUAL_FN Dual<K> operator-(const Dual<K>& a) {
    Dual<K> r;
    r.v = -a.v;
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = -a.d[t];
    return r;
}
TJ_DUAL_FN Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = a.d[t] + b.d[t];
    return r;
}
TJ_DUAL_FN Dual<K> operator+(const Dual<K>& a, double b) {
    Dual<K> r = a;
    r.v = a.v + b;
    return r;
}
TJ_DUAL_FN Dual<K> operator+(double a, const Dual<K>& b) {
    Dual<K> r = b;
    r.v = a + b.v;
    return r;
}
TJ_DUAL_FN Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = a.d[t] - b.d[t];
    return r;
}
TJ_DUAL_FN Dual<K> operator-(const Dual<K>& a, double b) {
    Dual<K> r = a;
    r.v = a.v - b;
    return r;
}
TJ_DUAL_FN Dual<K> operator-(double a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a - b.v;
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = -b.d[t];
    return r;
}
TJ_DUAL_FN Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = __builtin_fma(a.d[t], b.v, a.v * b.d[t]);
    return r;
}
TJ_DUAL_FN Dual<K> operator*(const Dual<K>& a, double b) {
    Dual<K> r;
    r.v = a.v * b;
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = a.d[t] * b;
    return r;
}
TJ_DUAL_FN Dual<K> operator*(double a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a * b.v;
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = a * b.d[t];
    return r;
}
TJ_DUAL_FN Dual<K> tdiv(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v / b.v;
    const double rb = drcp(b.v);
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = __builtin_fma(-r.v, b.d[t], a.d[t]) * rb;
    return r;
}
TJ_DUAL_FN Dual<K> tdiv(const Dual<K>& a, double b) {
    Dual<K> r;
    r.v = a.v / b;
    const double rb = drcp(b);
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = a.d[t] * rb;
    return r;
}
TJ_DUAL_FN Dual<K> tdiv(double a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a / b.v;
    const double g = -r.v * drcp(b.v);
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = g * b.d[t];
    return r;
}
TJ_DUAL_FN Dual<K> tdiv_cr(const Dual<K>& a, const Dual<K>& b) { return tdiv(a, b); }
TJ_DUAL_FN Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) { return tdiv(a, b); }
TJ_DUAL_FN Dual<K> operator/(const Dual<K>& a, double b) { return tdiv(a, b); }
TJ_DUAL_FN Dual<K> operator/(double a, const Dual<K>& b) { return tdiv(a, b); }
TJ_DUAL_FN Dual<K>& operator+=(Dual<K>& a, const Dual<K>& b) { return a = a + b; }
TJ_DUAL_FN Dual<K>& operator-=(Dual<K>& a, const Dual<K>& b) { return a = a - b; }
TJ_DUAL_FN Dual<K>& operator*=(Dual<K>& a, const Dual<K>& b) { return a = a * b; }

#define TJ_DUAL_CMP(OP)                                                                          \
    TJ_DUAL_FN bool operator OP(const Dual<K>& a, const Dual<K>& b) { return a.v OP b.v; }       \
    TJ_DUAL_FN bool operator OP(const Dual<K>& a, double b) { return a.v OP b; }                 \
    TJ_DUAL_FN bool operator OP(double a, const Dual<K>& b) { return a OP b.v; }
TJ_DUAL_CMP(<)
TJ_DUAL_CMP(<=)
TJ_DUAL_CMP(>)
TJ_DUAL_CMP(>=)
TJ_DUAL_CMP(==)
TJ_DUAL_CMP(!=)
#undef TJ_DUAL_CMP

TJ_DUAL_FN Dual<K> tsqrt(const Dual<K>& a) {
    Dual<K> r;
    r.v = __builtin_sqrt(a.v);
    const double h = r.v > 0.0 ? 0.5 * drcp(r.v) : 0.0;
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = a.d[t] * h;
    return r;
}
TJ_DUAL_FN Dual<K> tabs(const Dual<K>& a) {
    Dual<K> r;
    r.v = __builtin_fabs(a.v);
    const double s = a.v < 0.0 ? -1.0 : 1.0;
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = s * a.d[t];
    return r;
}
TJ_DUAL_FN bool tisnan(const Dual<K>& a) { return __builtin_isnan(a.v); }
TJ_DUAL_FN bool tisinf(const Dual<K>& a) { return __builtin_isinf(a.v); }
// x^(-1/3): d = -(1/3) x^(-1/3) / x dx
TJ_DUAL_FN Dual<K> trcbrt(const Dual<K>& a) {
    Dual<K> r;
    r.v = pm_rcbrt(a.v);
    const double g = -(1.0 / 3.0) * r.v * drcp(a.v);
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = g * a.d[t];
    return r;
}
// MUSCL with van Leer's limiter (hllc.cpp:527-575) on dual numbers.  The
// value parts are muscl1 of tj_fluid.h, operation for operation.  With
// a = uc - um, b = up - uc and r = a / b, the limited slope
// phi(r) (a + b) / 2 with phi = min(2r, 2) / (1 + r) for r > 0 is exactly a
// for 0 < r < 1 and b for r >= 1 (0 for r <= 0): its derivative is da or db
// along the branch the value takes.
TJ_DUAL_FN void muscl1(Dual<K> um, Dual<K> uc, Dual<K> up, Dual<K>& L, Dual<K>& Rt) {
    double r = tdiv(uc.v - um.v, up.v - uc.v);
    if (tisinf(r) || tisnan(r)) r = 0.0;
    const double g = vanleer(r) * 0.5 * (up.v - um.v);
    L.v = uc.v - 0.5 * g;
    Rt.v = uc.v + 0.5 * g;
#pragma unroll
    for (int t = 0; t < K; ++t) {
        const double gd = r <= 0.0 ? 0.0 : (r < 1.0 ? uc.d[t] - um.d[t] : up.d[t] - uc.d[t]);
        L.d[t] = uc.d[t] - 0.5 * gd;
        Rt.d[t] = uc.d[t] + 0.5 * gd;
    }
}
// THINC's exp / tanh / cosh: values by the project routines the FP64 form
// uses; exp' = exp, tanh' = 1 - tanh^2, cosh' = sinh
TJ_DUAL_FN Dual<K> texp(const Dual<K>& a) {
    Dual<K> r;
    r.v = pm_exp(a.v);
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = r.v * a.d[t];
    return r;
}
TJ_DUAL_FN Dual<K> ttanh(const Dual<K>& a) {
    Dual<K> r;
    r.v = pm_tanh(a.v);
    const double g = 1.0 - r.v * r.v;
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = g * a.d[t];
    return r;
}
TJ_DUAL_FN Dual<K> tcosh(const Dual<K>& a) {
    Dual<K> r;
    r.v = pm_cosh(a.v);
    const double g = 0.5 * (pm_exp(a.v) - pm_exp(-a.v));
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = g * a.d[t];
    return r;
}
// lane src's value of x within the wave (__shfl), for plain and dual numbers
__device__ __forceinline__ double shfl_any(double x, int src) { return __shfl(x, src, 64); }
__device__ __forceinline__ float shfl_any(float x, int src) { return __shfl(x, src, 64); }
TJ_DUAL_FN Dual<K> shfl_any(const Dual<K>& x, int src) {
    Dual<K> r;
    r.v = __shfl(x.v, src, 64);
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = __shfl(x.d[t], src, 64);
    return r;
}
// true when any derivative part is NaN (the value part is checked separately)
TJ_DUAL_FN bool tan_nan(const Dual<K>& a) {
    bool b = false;
#pragma unroll
    for (int t = 0; t < K; ++t) b |= __builtin_isnan(a.d[t]);
    return b;
}
#undef TJ_DUAL_FN

}  // namespace tj
