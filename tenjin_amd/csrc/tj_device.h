// Device-side data structures and per-cell physics of the gfx950 update.
//
// Every expression keeps the operation order of the reference file cited
// beside it; the library is compiled with -ffp-contract=off, so the FP64 form
// rounds exactly like the reference's CPU code (IEEE division and square root
// are correctly rounded on both sides).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tj {

constexpr int MAXM = 4;

// Per-material constants.  Gg, gp1 and xi are the values the reference
// recomputes on every call (GruneisenGamma, GruneisenGamma + 1.0,
// 1.0 / GruneisenGamma in equationofstate.cpp:47-75); they are evaluated once
// on the host with the same operations, so they are bit-identical.  The
// mixed-precision form reads an FP32 copy (MatT<float>) in its FP32 parts.
template <typename R>
struct MatT {
    int eos, solid, plastic, mix;
    R Gg, gp1, xi;
    R pref, eref, CV;
    R rho0, c0, s, G0, yield;
    // MixtureEOS (equationofstate.cpp:116-376): ga = the reactant's
    // adiabaticIndex as given (the mixture formulas read it, not Gg + 1);
    // g2, Gg2, pref2, eref2, CV2 the products' MieGruneisen EOS
    R ga, g2, Gg2, pref2, eref2, CV2;
    // RomenskiiSolidEOS (K0, EOSalpha, EOSbeta; equationofstate.cpp:403-495)
    // and WilkinsSolidEOS (e1 .. e5; :502-596), with rho0 and G0 above
    R K0, ea, eb, e1, e2, e3, e4, e5;
};
using MatC = MatT<double>;

// Variable numbering of structdefinitions.h:10-32
enum Var {
    ALPHA, ALPHARHO, RHO_K, RHO, RHOU, TOTAL_E, VELOCITY, PRESSURE, SOUNDSPEED, USTAR,
    RHO_MIX, ALPHARHOLAMBDA, LAMBDA, SIGMA, V_TENSOR, VSTAR, DEVH, HJ2, EPSILON,
    ALPHARHOEPSILON, NVAR
};

struct Level {
    int dim;
    int n[3];          // cells per dimension (1 for unused dimensions)
    double dx[3];
    double yn[2];      // -1/(8 dx), -1/(8 dy): Youngs-normal prefactors (thinc.cpp:67-68)
    int nmat;
    int solid, plastic, thinc;
    double beta0, cfl;
    int bc_lo[3], bc_hi[3];
    const uint8_t* edge[3];  // per coordinate: bit0 = first cell of a box, bit1 = last cell of a box
    MatC m[MAXM];
    MatT<float> mf[MAXM];  // FP32 copy of m for the mixed-precision form
    int cs[NVAR];      // first component of each Variable (AccessPattern), -1 if absent
    long long ncells;
    int mixm;          // the mixture material (MixtureEOS), -1 if none
    // materials with derivative parts (MatT<Dual<K>>[MAXM], device memory) for
    // the tangent-linear form; null elsewhere
    const void* md;
    // partial slots of the CFL minimum (DT_PARTS words; null: the clock's word)
    unsigned long long* dtpart;
};

template <typename R>
__device__ __forceinline__ const MatT<R>& mat(const Level& lv, int m) {
    return static_cast<const MatT<R>*>(lv.md)[m];
}
template <>
__device__ __forceinline__ const MatT<double>& mat<double>(const Level& lv, int m) { return lv.m[m]; }
template <>
__device__ __forceinline__ const MatT<float>& mat<float>(const Level& lv, int m) { return lv.mf[m]; }

// One FAB of a per-box MultiFab: p0 is the address level cell (0,0,0) would
// have in this FAB's strides (an address outside the FAB when the box does
// not start at the origin; only cells of the box itself are ever formed).
struct FabRec {
    double* p0;
    long long js, ks, ns;
};

// The kernels reach the state only through MF::at(c, i, j, k) with level cell
// indices.  The library is compiled twice, once per form of MF: a level array,
// or one FAB per box.  Each compilation puts MF and everything that takes it
// (kernels, device bodies, launch code) in its own inline namespace, so the
// two sets of symbols stay apart; the types above and below are shared.
#ifdef TJ_FAB_VIEWS
#define TJ_VIEW_NS fab_view
#else
#define TJ_VIEW_NS level_view
#endif
inline namespace TJ_VIEW_NS {
#ifdef TJ_FAB_VIEWS
// Caller-owned MultiFab of one FAB per box (tj_fabset): the cell's box from the
// per-dimension box maps (cell index -> box coordinate), then its FAB.
struct MF {
    const FabRec* rec;                // boxes in box-grid order, x fastest
    const int* map0;                  // per x index: box column
    const int* map1;                  // per y index: box row * boxes along x
    const int* map2;                  // per z index: box layer * boxes per layer
    __device__ __forceinline__ double& at(int c, int i, int j, int k) const {
        const FabRec& r = rec[map2[k] + map1[j] + map0[i]];
        return r.p0[(long long)c * r.ns + (long long)k * r.ks + (long long)j * r.js + i];
    }
};
__host__ __device__ __forceinline__ bool same_view(const MF& a, const MF& b) { return a.rec == b.rec; }
#else
// Caller-owned CellArray view (component-major, i fastest).
struct MF {
    double* p;
    long long js, ks, ns;
    __device__ __forceinline__ double& at(int c, int i, int j, int k) const {
        return p[(long long)c * ns + (long long)k * ks + (long long)j * js + i];
    }
};
__host__ __device__ __forceinline__ bool same_view(const MF& a, const MF& b) { return a.p == b.p; }
#endif
}  // inline namespace TJ_VIEW_NS

// Face-centred flux planes of one direction: lower face of cell (i,j,k),
// extent n_d + 1 along d; component-major.  FP64, or FP32 in the
// mixed-precision form.
template <typename T>
struct FacesT {
    T* p;
    int nx, ny, nz;     // face-grid extents
    long long cs;       // component stride
    __device__ __forceinline__ T& at(int c, int i, int j, int k) const {
        return p[(long long)c * cs + ((long long)k * ny + j) * nx + i];
    }
};
using Faces = FacesT<double>;

// THINC/BVD history planes (faithful form of the reference's stale
// ULStar/URStar reads, thinc.cpp:57-63 with hllc.cpp:404-406):
//   urstar  : last THINC right-face alpha written for each cell and material
//   star_in : star-state alpha of each cell's lower face from the previous
//             directional flux pass (what ULStar holds in valid cells)
//   edge_in : star-state alpha of box-high faces from the previous pass in the
//             same direction (what ULStar holds in a box's high ghost layer)
struct Hist {
    double* urstar;
    const double* star_in;
    double* star_out;
    const double* edge_in;
    double* edge_out;
    long long pstride;  // plane stride of urstar/star planes (= ncells)
    long long estride;  // plane stride of the edge planes (= face count of direction d)
};

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ double sgnd(double v) { return (double)((0.0 < v) - (v < 0.0)); }

// Real-type generic forms used by the reconstruction and flux code, which runs
// in FP64 (R = double: the same operations as above) or FP32 (mixed form).
template <typename R>
__device__ __forceinline__ R tmin(R a, R b) { return (b < a) ? b : a; }
template <typename R>
__device__ __forceinline__ R tmax(R a, R b) { return (a < b) ? b : a; }
template <typename R>
__device__ __forceinline__ R tsgn(R v) { return (R)((R(0) < v) - (v < R(0))); }
// Division and square root: IEEE (correctly rounded) in FP64; in the FP32
// parts of the mixed form the hardware reciprocal and square root (1 ulp),
// whose error is far below that form's accuracy bar.
__device__ __forceinline__ double tdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float tdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
// Correctly rounded in both forms: the HLLC star-state factor (S_K - u)/(S_K - S*)
// (hllc.cpp:40-44), which must be exactly 1 for two identical states at rest.
// With the 1-ulp reciprocal it is not, and every face of a quiet region then
// carries a flux S_K (mult - 1) U: it cancels inside the region, but at an open
// domain end it is a boundary flux that moves mass and energy in or out.
__device__ __forceinline__ double tdiv_cr(double a, double b) { return a / b; }
__device__ __forceinline__ float tdiv_cr(float a, float b) { return a / b; }
__device__ __forceinline__ double tsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float tsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double tabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float tabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ bool tisnan(double x) { return __builtin_isnan(x); }
__device__ __forceinline__ bool tisnan(float x) { return __builtin_isnan(x); }
__device__ __forceinline__ bool tisinf(double x) { return __builtin_isinf(x); }
__device__ __forceinline__ bool tisinf(float x) { return __builtin_isinf(x); }

// ------------------------------------------------------------------ wave index
// The wave's index within its block, marked wave-uniform for the compiler
// (readfirstlane): what is derived from it (rows, segments, row addresses,
// per-wave LDS bases) is then scalar work in SGPRs, not per-lane VALU work.
__device__ __forceinline__ int wave_in_block() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// ------------------------------------------------------------------ neighbour-lane moves
// x of lane - 1 (lane_dn) or lane + 1 (lane_up) within the wave, lane 0 (63)
// keeping its own: DPP wave_shr:1 / wave_shl:1 row moves (a VALU mov on the
// register file) in place of __shfl's ds_bpermute through the LDS crossbar.
// Every lane of the wave must be active at the call.  bound_ctrl off: the lane
// with no source keeps `old`, which is x itself.
__device__ __forceinline__ int lane_dn(int x) { return __builtin_amdgcn_update_dpp(x, x, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int lane_up(int x) { return __builtin_amdgcn_update_dpp(x, x, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ float lane_dn(float x) {
    return __builtin_bit_cast(float, lane_dn(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ float lane_up(float x) {
    return __builtin_bit_cast(float, lane_up(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double lane_dn(double x) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)lane_dn((int)b), hi = (unsigned)lane_dn((int)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double lane_up(double x) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)lane_up((int)b), hi = (unsigned)lane_up((int)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// ------------------------------------------------------------------ math
// Project-written exp/expm1/tanh/cosh/x^(-1/3): Cody-Waite reduction with the
// fdlibm split of ln2 and Taylor/Horner polynomials; the oracle evaluates the
// same algorithm with the same constants in the same order.
__device__ inline double pm_exp(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return __builtin_huge_val();
    if (x < -745.1332191019412) return 0.0;
    const double log2e = 1.4426950408889634;
    const double ln2hi = 6.93147180369123816490e-01;
    const double ln2lo = 1.90821492927058770002e-10;
    double kd = __builtin_rint(x * log2e);
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
    return __builtin_ldexp(p, (int)kd);
}
__device__ inline double pm_expm1(double y) {
    if (__builtin_fabs(y) < 0.3) {
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
__device__ inline double pm_tanh(double x) {
    double ax = __builtin_fabs(x);
    if (ax > 20.0) return x > 0.0 ? 1.0 : -1.0;
    double t = pm_expm1(2.0 * ax);
    double r = t / (t + 2.0);
    return x < 0.0 ? -r : r;
}
__device__ inline double pm_cosh(double x) {
    double e = pm_exp(__builtin_fabs(x));
    return (e + 1.0 / e) * 0.5;
}
__device__ inline double pm_rcbrt(double x) {
    uint64_t bits = __builtin_bit_cast(uint64_t, x);
    uint64_t g = 0x553ef0ff289dd796ULL - bits / 3;
    double y = __builtin_bit_cast(double, g);
    for (int it = 0; it < 5; ++it) {
        double y3 = (y * y) * y;
        y = y * ((4.0 - x * y3) / 3.0);
    }
    double c = 1.0 / y;
    c = c - (c - x / (c * c)) / 3.0;
    return 1.0 / c;
}

// log (x = 2^e m, m in [sqrt(2)/2, sqrt(2)), the atanh series of
// s = (m - 1)/(m + 1) to s^25) and pow = exp(y log x), for std::pow of
// reactive.cpp:17; the oracle evaluates the same steps.
__device__ inline double pm_log(double x) {
    if (x != x || x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_huge_val();
    if (x == __builtin_huge_val()) return x;
    uint64_t bits = __builtin_bit_cast(uint64_t, x);
    int e = 0;
    if ((bits >> 52) == 0) {
        x *= 18014398509481984.0;
        bits = __builtin_bit_cast(uint64_t, x);
        e = -54;
    }
    e += (int)((bits >> 52) & 0x7ff) - 1023;
    bits = (bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m = __builtin_bit_cast(double, bits);
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
__device__ inline double pm_pow(double x, double y) {
    if (y == 0.0) return 1.0;
    if (x == 0.0) return y > 0.0 ? 0.0 : __builtin_huge_val();
    return pm_exp(y * pm_log(x));
}

// exp/tanh/cosh by real type: the project-written FP64 routines above, and
// the device library's single-precision functions for the mixed form.
// x^y and log x of the solid EOSs (std::pow, std::log in equationofstate.cpp:
// 403-596): the project routines in FP64, the device library's in FP32.
__device__ __forceinline__ double tpow(double x, double y) { return pm_pow(x, y); }
__device__ __forceinline__ float tpow(float x, float y) { return powf(x, y); }
__device__ __forceinline__ double tlog(double x) { return pm_log(x); }
__device__ __forceinline__ float tlog(float x) { return logf(x); }
__device__ __forceinline__ double texp(double x) { return pm_exp(x); }
__device__ __forceinline__ double ttanh(double x) { return pm_tanh(x); }
__device__ __forceinline__ double tcosh(double x) { return pm_cosh(x); }
__device__ __forceinline__ float texp(float x) { return expf(x); }
__device__ __forceinline__ float ttanh(float x) { return tanhf(x); }
__device__ __forceinline__ float tcosh(float x) { return coshf(x); }

// ------------------------------------------------------------------ boundaries
// boundary.cpp:3-42: transmissive -> foextrap (nearest valid cell),
// reflective -> reflect_even, with reflect_odd (sign flip) for the normal
// velocity and momentum.  Each coordinate is mapped independently.
__device__ __forceinline__ int bcmap(int c, int n, int rlo, int rhi, bool& flip) {
    flip = false;
    int s = c;
    if (c < 0) {
        if (rlo) { flip = true; s = -1 - c; }
        else s = 0;
    } else if (c > n - 1) {
        if (rhi) { flip = true; s = 2 * n - 1 - c; }
        else s = n - 1;
    }
    // lanes far outside the domain (unused results) still load in bounds
    return s < 0 ? 0 : (s > n - 1 ? n - 1 : s);
}

}  // namespace tj
