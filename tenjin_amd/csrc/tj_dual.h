// Forward-mode dual numbers for the tangent-linear (differentiated) form of
// the solid path: a value and K directional derivatives.  The value part of
// every operation is the same IEEE double operation the FP64 form performs,
// in the same order, so a tangent run reproduces the FP64 trajectory bit for
// bit; the derivative parts follow the chain rule of each operation.
//
// The derivative parts are not bound to the value's rounding: they use fused
// multiply-adds and a reciprocal refined from the hardware estimate (two
// Newton steps, a few ulp) instead of a second IEEE division.  The one
// quotient whose denominator reaches the subnormals, the MUSCL ratio of
// differences of exponentially small precursor values ahead of a wave (where
// the estimate overflows), divides (muscl1 below).
//
// Non-smooth points take one one-sided derivative: |x| and max/min follow the
// branch the value takes, sqrt at 0 has derivative 0 (it is reached only at a
// minimum of its argument, HJ2 = 0 of an undeformed cell), and the MUSCL
// limiter's inf/NaN reset zeroes the derivative with the value.
#pragma once
#include "tj_device.h"

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
// MUSCL with van Leer's limiter (hllc.cpp:527-575) on dual numbers: the
// value parts as muscl1 of tj_fluid.h; the ratio's derivative parts by IEEE
// division (its denominator can be subnormal)
TJ_DUAL_FN void muscl1(Dual<K> um, Dual<K> uc, Dual<K> up, Dual<K>& L, Dual<K>& Rt) {
    const Dual<K> num = uc - um, den = up - uc;
    Dual<K> r;
    r.v = num.v / den.v;
#pragma unroll
    for (int t = 0; t < K; ++t) r.d[t] = (num.d[t] - r.v * den.d[t]) / den.v;
    if (tisinf(r) || tisnan(r)) r = Dual<K>(0.0);
    Dual<K> g = vanleer(r) * 0.5 * (up - um);
    L = uc - 0.5 * g;
    Rt = uc + 0.5 * g;
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
