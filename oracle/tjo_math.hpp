// Test infrastructure only (the oracle). Not product code.
//
// Project-written elementary functions used by the oracle where the reference
// calls libm (std::exp / cosh / tanh in thinc.cpp:166-169, std::pow(x,-1/3) in
// boxaccesscellarray.cpp:514).  The GPU kernels evaluate the same published
// algorithms with the same constants and the same operation order, so the
// oracle and the device agree bit for bit instead of relying on glibc and
// the ROCm device library agreeing to the last ulp.  The libm variants are
// kept beside them (math_mode = 1) so tests can show that the two differ only
// at the ulp level.
//
//  exp   : Cody-Waite reduction x = k ln2 + r with the fdlibm split of ln2
//          (ln2_hi has 32 trailing zero bits, so k*ln2_hi is exact), then a
//          degree-13 Taylor polynomial for e^r on |r| <= ln2/2 evaluated by
//          Horner's rule, then scaling by 2^k.
//  expm1 : degree-14 Taylor polynomial for |y| < 0.3, exp(y)-1 otherwise.
//  tanh  : expm1(2|x|) / (expm1(2|x|) + 2), sign restored.
//  cosh  : (e + 1/e)/2 with e = exp(|x|).
//  rcbrt : x^(-1/3) by bit-level initial guess plus Newton steps on
//          y <- y (4 - x y^3)/3, finished with one step of the cube-root
//          Newton iteration on 1/y.
//  log   : x = 2^e m with m in [sqrt(2)/2, sqrt(2)), s = (m - 1)/(m + 1) and
//          log m = 2 atanh(s) = 2s + s z (2/3 + 2/5 z + ... + 2/25 z^11),
//          z = s^2 (Horner), then e ln2_hi + (log m + e ln2_lo).
//  pow   : exp(y log x) (std::pow in reactive.cpp:17): x < 0 gives NaN
//          through the log, x = 0 gives 0 for y > 0.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace tjo {

inline double pm_exp(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return HUGE_VAL;
    if (x < -745.1332191019412) return 0.0;
    const double log2e = 1.4426950408889634;
    const double ln2hi = 6.93147180369123816490e-01;
    const double ln2lo = 1.90821492927058770002e-10;
    double kd = std::nearbyint(x * log2e);
    double r = (x - kd * ln2hi) - kd * ln2lo;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return std::ldexp(p, (int)kd);
}

inline double pm_expm1(double y) {
    if (std::fabs(y) < 0.3) {
        double p = 1.0 / 87178291200.0;
        p = p * y + 1.0 / 6227020800.0;
        p = p * y + 1.0 / 479001600.0;
        p = p * y + 1.0 / 39916800.0;
        p = p * y + 1.0 / 3628800.0;
        p = p * y + 1.0 / 362880.0;
        p = p * y + 1.0 / 40320.0;
        p = p * y + 1.0 / 5040.0;
        p = p * y + 1.0 / 720.0;
        p = p * y + 1.0 / 120.0;
        p = p * y + 1.0 / 24.0;
        p = p * y + 1.0 / 6.0;
        p = p * y + 0.5;
        p = p * y + 1.0;
        return p * y;
    }
    return pm_exp(y) - 1.0;
}

inline double pm_tanh(double x) {
    double ax = std::fabs(x);
    if (ax > 20.0) return x > 0.0 ? 1.0 : -1.0;
    double t = pm_expm1(2.0 * ax);
    double r = t / (t + 2.0);
    return x < 0.0 ? -r : r;
}

inline double pm_cosh(double x) {
    double e = pm_exp(std::fabs(x));
    return (e + 1.0 / e) * 0.5;
}

// x^(-1/3) for finite x > 0.
inline double pm_rcbrt(double x) {
    uint64_t bits;
    std::memcpy(&bits, &x, 8);
    // initial guess: exponent scaled by -1/3 around the bias
    uint64_t g = 0x553ef0ff289dd796ULL - bits / 3;
    double y;
    std::memcpy(&y, &g, 8);
    for (int it = 0; it < 5; ++it) {
        double y3 = (y * y) * y;
        y = y * ((4.0 - x * y3) / 3.0);
    }
    double c = 1.0 / y;                       // cube root estimate
    c = c - (c - x / (c * c)) / 3.0;          // Newton step for c^3 = x
    return 1.0 / c;
}

inline double pm_log(double x) {
    if (x != x || x < 0.0) return std::nan("");
    if (x == 0.0) return -HUGE_VAL;
    if (x == HUGE_VAL) return x;
    uint64_t bits;
    std::memcpy(&bits, &x, 8);
    int e = 0;
    if ((bits >> 52) == 0) {  // subnormal: scale by 2^54
        x *= 18014398509481984.0;
        std::memcpy(&bits, &x, 8);
        e = -54;
    }
    e += (int)((bits >> 52) & 0x7ff) - 1023;
    bits = (bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m;
    std::memcpy(&m, &bits, 8);
    if (m > 1.4142135623730951) {
        m *= 0.5;
        e += 1;
    }
    const double ln2hi = 6.93147180369123816490e-01;
    const double ln2lo = 1.90821492927058770002e-10;
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double q = 2.0 / 25.0;
    q = q * z + 2.0 / 23.0;
    q = q * z + 2.0 / 21.0;
    q = q * z + 2.0 / 19.0;
    q = q * z + 2.0 / 17.0;
    q = q * z + 2.0 / 15.0;
    q = q * z + 2.0 / 13.0;
    q = q * z + 2.0 / 11.0;
    q = q * z + 2.0 / 9.0;
    q = q * z + 2.0 / 7.0;
    q = q * z + 2.0 / 5.0;
    q = q * z + 2.0 / 3.0;
    const double lm = 2.0 * s + s * (z * q);
    const double ed = (double)e;
    return ed * ln2hi + (lm + ed * ln2lo);
}

inline double pm_pow(double x, double y) {
    if (y == 0.0) return 1.0;
    if (x == 0.0) return y > 0.0 ? 0.0 : HUGE_VAL;
    return pm_exp(y * pm_log(x));
}

// math_mode 0: project functions (bit-identical with the device);
// math_mode 1: libm, as the reference calls it.
struct Math {
    int mode = 0;
    double exp(double x) const { return mode ? std::exp(x) : pm_exp(x); }
    double cosh(double x) const { return mode ? std::cosh(x) : pm_cosh(x); }
    double tanh(double x) const { return mode ? std::tanh(x) : pm_tanh(x); }
    double pow_m13(double x) const { return mode ? std::pow(x, -1.0 / 3.0) : pm_rcbrt(x); }
    double log(double x) const { return mode ? std::log(x) : pm_log(x); }
    double pow(double x, double y) const { return mode ? std::pow(x, y) : pm_pow(x, y); }
};

}  // namespace tjo
