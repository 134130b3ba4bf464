// Elastoplastic (solid) per-cell physics: 3x3 algebra, Hencky strain and
// stress, the Us-up Mie-Gruneisen solid EOS, primitive/conservative maps,
// sound and transverse speeds, the 5-wave HLLC-type flux, the perfect-plastic
// radial return and the V clean-up.  Operation order follows the reference
// files cited beside each function.
#pragma once
#include "tj_fluid.h"
#include "tj_dual.h"

namespace tj {
inline namespace TJ_VIEW_NS {

// ------------------------------------------------------------------ 3x3 algebra
// tensor.cpp:81-115 (products: k = 0,1,2 summed left to right), 140-150
// (trace), 187-233 (adjugate inverse), 296-333 (cofactor determinant).
// R: double (FP64 form and every per-cell update), float (face states and
// fluxes of the mixed-precision form) or Dual<K> (tangent-linear form).
template <typename R>
__device__ __forceinline__ R det2(R a, R b, R c, R d) { return a * d - b * c; }
template <typename R>
__device__ __forceinline__ R det3(const R* M) {
    R det = R(0);
    det += M[0] * det2(M[4], M[5], M[7], M[8]);
    det += (-M[1]) * det2(M[3], M[5], M[6], M[8]);
    det += M[2] * det2(M[3], M[4], M[6], M[7]);
    return det;
}
template <typename R>
__device__ __forceinline__ void inv3(const R* M, R* Mi) {
    const R D = det3(M);
    R cof[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int k0 = i == 0 ? 1 : 0, k1 = i == 2 ? 1 : 2;
            const int m0 = j == 0 ? 1 : 0, m1 = j == 2 ? 1 : 2;
            const R sign = ((i + j) % 2) ? R(-1) : R(1);
            cof[i * 3 + j] = sign * det2(M[k0 * 3 + m0], M[k0 * 3 + m1], M[k1 * 3 + m0], M[k1 * 3 + m1]);
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Mi[i * 3 + j] = tdiv(cof[j * 3 + i], D);
}
template <typename R>
__device__ __forceinline__ void mm3(const R* A, const R* B, R* C) {  // C = A B
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            C[r * 3 + c] = (A[r * 3] * B[c] + A[r * 3 + 1] * B[3 + c]) + A[r * 3 + 2] * B[6 + c];
}
template <typename R>
__device__ __forceinline__ void mmT3(const R* A, const R* B, R* C) {  // C = A B^T
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            C[r * 3 + c] = (A[r * 3] * B[c * 3] + A[r * 3 + 1] * B[c * 3 + 1]) + A[r * 3 + 2] * B[c * 3 + 2];
}

// getDeviatoricHenckyStrain + getHenckyJ2 (boxaccesscellarray.cpp:326-380):
// devH = (G - G^-1)/4 with G = V V^T, HJ2 = tr(devH devH^T).
template <typename R>
__device__ __forceinline__ R hencky(const R* V, R* dH) {
    R G[9], Gi[9];
    mmT3(V, V, G);
    inv3(G, Gi);
#pragma unroll
    for (int q = 0; q < 9; ++q) dH[q] = R(0.5) * R(0.5) * (G[q] - Gi[q]);
    R t = R(0);
#pragma unroll
    for (int r = 0; r < 3; ++r) t += (dH[r * 3] * dH[r * 3] + dH[r * 3 + 1] * dH[r * 3 + 1]) + dH[r * 3 + 2] * dH[r * 3 + 2];
    return t;
}

// Principal square root of an SPD 3x3 matrix by cyclic Jacobi rotations
// (Golub & Van Loan, Matrix Computations, 8.5), six sweeps; the oracle
// evaluates the same sequence.
// jacobi_spd leaves A = Q diag(a) Q^T (a[0], a[4], a[8]) with the rotations
// accumulated in q; sqrtm_spd assembles Q diag(sqrt(a)) Q^T.
template <typename R>
__device__ __forceinline__ void jacobi_spd(const R* Ain, R* a, R* q) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        a[i] = Ain[i];
        q[i] = (i / 3 == i % 3) ? R(1) : R(0);
    }
    for (int sweep = 0; sweep < 6; ++sweep) {
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) {
            const int p = rr == 2 ? 1 : 0, qq = rr == 0 ? 1 : 2;
            const R apq = a[p * 3 + qq];
            if (apq == R(0)) continue;
            const R app = a[p * 3 + p], aqq = a[qq * 3 + qq];
            // FP32: an off-diagonal below the diagonal's resolution rotates by
            // nothing representable, and its reciprocal can overflow
            if (sizeof(R) == 4 && tabs(apq) < R(1e-12) * (tabs(app) + tabs(aqq))) continue;
            const R theta = tdiv(aqq - app, R(2) * apq);
            const R t = tdiv(theta >= R(0) ? R(1) : R(-1), tabs(theta) + tsqrt(theta * theta + R(1)));
            const R c = tdiv(R(1), tsqrt(t * t + R(1)));
            const R s = t * c;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const R akp = a[k * 3 + p], akq = a[k * 3 + qq];
                a[k * 3 + p] = c * akp - s * akq;
                a[k * 3 + qq] = s * akp + c * akq;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const R apk = a[p * 3 + k], aqk = a[qq * 3 + k];
                a[p * 3 + k] = c * apk - s * aqk;
                a[qq * 3 + k] = s * apk + c * aqk;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const R qkp = q[k * 3 + p], qkq = q[k * 3 + qq];
                q[k * 3 + p] = c * qkp - s * qkq;
                q[k * 3 + qq] = s * qkp + c * qkq;
            }
        }
    }
}
template <typename R = double>
__device__ inline void sqrtm_spd(const R* Ain, R* S) {
    R a[9], q[9];
    jacobi_spd<R>(Ain, a, q);
    const R sl0 = tsqrt(a[0]), sl1 = tsqrt(a[4]), sl2 = tsqrt(a[8]);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            S[r * 3 + c] = (q[r * 3] * sl0 * q[c * 3] + q[r * 3 + 1] * sl1 * q[c * 3 + 1]) + q[r * 3 + 2] * sl2 * q[c * 3 + 2];
}
// Tangent form: the value by the FP64 sweeps above; the derivative from
// S dS + dS S = dA (differentiating S S = A), solved in the eigenbasis of the
// sweeps: dS = Q [ (Q^T dA Q)_rc / (sqrt(a_r) + sqrt(a_c)) ] Q^T.  Rotating
// the derivative parts through the sweeps instead loses them to cancellation
// when eigenvalues nearly coincide (V close to I).
template <int K>
__device__ inline void sqrtm_spd(const Dual<K>* Ain, Dual<K>* S) {
    double Av[9], a[9], q[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Av[i] = Ain[i].v;
    jacobi_spd<double>(Av, a, q);
    const double sl[3] = {__builtin_sqrt(a[0]), __builtin_sqrt(a[4]), __builtin_sqrt(a[8])};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            S[r * 3 + c].v = (q[r * 3] * sl[0] * q[c * 3] + q[r * 3 + 1] * sl[1] * q[c * 3 + 1]) + q[r * 3 + 2] * sl[2] * q[c * 3 + 2];
#pragma unroll
    for (int t = 0; t < K; ++t) {
        double T1[9], M[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                T1[r * 3 + c] = Ain[r * 3].d[t] * q[c] + Ain[r * 3 + 1].d[t] * q[3 + c] + Ain[r * 3 + 2].d[t] * q[6 + c];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                M[r * 3 + c] = (q[r] * T1[c] + q[3 + r] * T1[3 + c] + q[6 + r] * T1[6 + c]) / (sl[r] + sl[c]);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                T1[r * 3 + c] = q[r * 3] * M[c] + q[r * 3 + 1] * M[3 + c] + q[r * 3 + 2] * M[6 + c];  // Q M
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                S[r * 3 + c].d[t] = T1[r * 3] * q[c * 3] + T1[r * 3 + 1] * q[c * 3 + 1] + T1[r * 3 + 2] * q[c * 3 + 2];  // (Q M) Q^T
    }
}

// ------------------------------------------------------------------ Us-up solid EOS
// Hugoniot reference curve p_H = rho0 c0^2 eta/(1 - s eta)^2, e_H = p_H eta/(2 rho0),
// eta = 1 - rho0/rho, with the Mie-Gruneisen form p = p_H + Gamma rho (e - e_H).
template <typename R>
__device__ __forceinline__ R pH(const MatT<R>& M, R rho) {
    const R eta = R(1) - tdiv(M.rho0, rho);
    const R den = R(1) - M.s * eta;
    return tdiv(M.rho0 * M.c0 * M.c0 * eta, den * den);
}
template <typename R>
__device__ __forceinline__ R eH(const MatT<R>& M, R rho) {
    const R eta = R(1) - tdiv(M.rho0, rho);
    return tdiv(pH(M, rho) * eta, R(2) * M.rho0);
}
template <typename R>
__device__ __forceinline__ R dpH(const MatT<R>& M, R rho) {
    const R eta = R(1) - tdiv(M.rho0, rho);
    const R den = R(1) - M.s * eta;
    const R detadrho = tdiv(M.rho0, rho * rho);
    return tdiv(M.rho0 * M.c0 * M.c0 * (R(1) + M.s * eta), den * den * den) * detadrho;
}
template <typename R>
__device__ __forceinline__ R deH(const MatT<R>& M, R rho) {
    const R eta = R(1) - tdiv(M.rho0, rho);
    const R detadrho = tdiv(M.rho0, rho * rho);
    return tdiv(dpH(M, rho) * eta + pH(M, rho) * detadrho, R(2) * M.rho0);
}
// componentShearModulus (equationofstate.cpp:77-80, 447-450, 548-551): G0 of
// the Us-up and Wilkins solids, G0 (rho_k/rho0)^(beta + 1) of the Romenskii
// solid, 0 for a MieGruneisen material.  RW: the level holds a Romenskii or
// Wilkins material.  The kernels are instantiated both ways and the launch
// picks one, so that the Us-up kernels carry none of the pow/log code (their
// registers and spills double with it).
template <bool RW, typename R>
__device__ __forceinline__ R shearG(const MatT<R>& M, R rk) {
    if (M.eos == 1 || M.eos == 3) return M.G0;
    if constexpr (RW) {
        if (M.eos == 2) return M.G0 * tpow(tdiv(rk, M.rho0), M.eb + R(1.0));
    }
    return R(0);
}

// RomenskiiSolidEOS (equationofstate.cpp:420-495) and WilkinsSolidEOS
// (:519-596): cold energy and pressure, dG/drho, d2G/drho2, dp_c/drho and
// dp_s/drho of material M at density rk; a = alpha, ar = alpha rho.
template <typename R>
__device__ __forceinline__ R cold_e(const MatT<R>& M, R ar, R rk) {
    const R x = tdiv(rk, M.rho0);
    if (M.eos == 2) {
        const R xa = tpow(x, M.ea);
        return ar * (tdiv(M.K0, R(2.0) * M.rho0 * M.ea * M.ea) * (xa - R(1.0)) * (xa - R(1.0)));
    }
    return ar * tdiv(M.e1 + M.e2 * x * x + M.e3 * x + tdiv(M.e4, x) - M.e5 * tlog(x), M.rho0);
}
template <typename R>
__device__ __forceinline__ R cold_p(const MatT<R>& M, R a, R rk) {
    const R x = tdiv(rk, M.rho0);
    if (M.eos == 2) return tdiv(a, M.Gg) * (tdiv(M.K0, M.ea) * tpow(x, M.ea + R(1.0)) * (tpow(x, M.ea) - R(1.0)));
    return tdiv(a, M.Gg) * ((R(2.0) * M.e2 * x * x * x + M.e3 * x * x - M.e4 - M.e5 * x));
}
template <typename R>
__device__ __forceinline__ R dGdrho(const MatT<R>& M, R rk) {
    if (M.eos == 2) return tdiv(M.G0 * (M.eb + R(1.0)) * tpow(tdiv(rk, M.rho0), M.eb), M.rho0);
    return R(0.0);
}
template <typename R>
__device__ __forceinline__ R dG2drho2(const MatT<R>& M, R rk) {
    if (M.eos == 2) return tdiv(M.G0 * (M.eb + R(1.0)) * (M.eb) * tpow(tdiv(rk, M.rho0), M.eb - R(1.0)), M.rho0 * M.rho0);
    return R(0.0);
}
template <typename R>
__device__ __forceinline__ R dpc_drho(const MatT<R>& M, R rk) {
    const R x = tdiv(rk, M.rho0);
    if (M.eos == 2) {
        const R xa = tpow(x, M.ea);
        return tdiv(M.K0, M.ea * M.rho0) * xa * ((R(2.0) * M.ea + R(1.0)) * xa - (M.ea + R(1.0)));
    }
    return tdiv(R(6.0) * M.e2 * x * x + R(2.0) * M.e3 * x - M.e5, M.rho0);
}

// coldCompressionInternalEnergy + shearInternalEnergy and coldCompressionPressure
// + shearPressure of a Romenskii or Wilkins material (equationofstate.cpp:420-495, 519-596)
template <bool RW, typename R>
__device__ __forceinline__ R rw_energy(const MatT<R>& M, R a, R ar, R rk, R hj2) {
    if constexpr (!RW) {
        return R(0);
    } else {
        return (cold_e(M, ar, rk) + a * (shearG<true>(M, rk) * hj2));
    }
}
template <bool RW, typename R>
__device__ __forceinline__ R rw_pressure(const MatT<R>& M, R a, R rk, R hj2) {
    if constexpr (!RW) {
        return R(0);
    } else {
        return cold_p(M, a, rk) + tdiv(a, M.Gg) * ((rk * dGdrho(M, rk) - shearG<true>(M, rk)) * hj2);
    }
}
// getSoundSpeedContribution of a Romenskii or Wilkins material
// (equationofstate.cpp:435-438, 536-539), whose cold and shear pressures carry
// the material's alpha / Gamma factor as the reference's do
template <bool RW, typename R>
__device__ __forceinline__ R rw_c2(const MatT<R>& M, R p, R a, R rk, R hj2) {
    if constexpr (!RW) {
        return R(0);
    } else {
        const R G = shearG<true>(M, rk);
        const R ps = tdiv(a, M.Gg) * ((rk * dGdrho(M, rk) - G) * hj2);
        return tdiv(p * (M.Gg + R(1.0)) - (cold_p(M, a, rk) + ps) * (M.Gg + R(1.0)), rk) + dpc_drho(M, rk) +
               rk * dG2drho2(M, rk) * hj2 + tdiv((R(4.0) / R(3.0)) * G, rk);
    }
}

// effective sums of boxaccesscellarray.cpp:218-252 with the solid terms
template <int NM, bool RW, typename R>
__device__ __forceinline__ void eos_sums_s(const Level& lv, const R* a, const R* ar, const R* rk, R hj2, R& G, R& Ent,
                                           R& Pnt) {
    G = R(0);
    Ent = R(0);
    Pnt = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) G += tdiv(a[m], mat<R>(lv, m).Gg);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatT<R>& M = mat<R>(lv, m);
        if (M.eos == 1) Ent += (ar[m] * eH(M, rk[m]) + a[m] * (shearG<RW>(M, rk[m]) * hj2));
        else if (RW && M.eos >= 2) Ent += rw_energy<RW>(M, a[m], ar[m], rk[m], hj2);
        else Ent += (M.eref * ar[m] + R(0) * a[m]);
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatT<R>& M = mat<R>(lv, m);
        if (M.eos == 1)
            Pnt += tdiv(a[m], M.Gg) * pH(M, rk[m]) + tdiv(a[m], M.Gg) * ((rk[m] * R(0) - shearG<RW>(M, rk[m])) * hj2);
        else if (RW && M.eos >= 2)
            Pnt += rw_pressure<RW>(M, a[m], rk[m], hj2);
        else Pnt += tdiv(M.pref * a[m], M.Gg) + R(0) * a[m];
    }
}
// getSoundSpeedContribution: stiffened gas (equationofstate.cpp:61-64), the
// Us-up solid with c^2 = dp/drho|_s + 4G/(3 rho), or the Romenskii / Wilkins
// form of equationofstate.cpp:435-438, 536-539, whose cold and shear
// pressures carry the material's alpha / Gamma factor as the reference's do.
template <bool RW, typename R>
__device__ __forceinline__ R c2_s(const MatT<R>& M, R p, R a, R rk, R hj2) {
    if (M.eos == 1)
        return tdiv(p * M.gp1 - pH(M, rk) + M.gp1 * M.G0 * hj2, rk) + dpH(M, rk) - M.Gg * rk * deH(M, rk) +
               tdiv((R(4.0) / R(3.0)) * M.G0, rk);
    if (RW && M.eos >= 2) return rw_c2<RW>(M, p, a, rk, hj2);
    return tdiv(p * M.gp1, rk) - tdiv(M.pref, rk);
}
template <int NM, bool RW, typename R>
__device__ __forceinline__ R sound_s(const Level& lv, const R* al, const R* ar, const R* rk, R rho, R p, R hj2) {
    R a = R(0), xt = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatT<R>& M = mat<R>(lv, m);
        a += tmax(R(0), tdiv(M.xi * c2_s<RW>(M, p, al[m], rk[m], hj2) * ar[m], rho));
        xt += M.xi * al[m];
    }
    if (a <= R(0)) a = R(1E-6);
    return tsqrt(tdiv(a, xt));
}
// stressTensor column d (boxaccesscellarray.cpp:176-216)
template <int NM, bool RW, typename R>
__device__ __forceinline__ void stress_col(const Level& lv, const R* al, const R* rk, R p, const R* dH, int d, R* sig) {
    R Gt = R(0), Gi = R(0);
    int solid = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m)
        if (mat<R>(lv, m).solid) {
            solid = 1;
            Gt += shearG<RW>(mat<R>(lv, m), rk[m]) * tdiv(al[m], mat<R>(lv, m).Gg);
        }
#pragma unroll
    for (int m = 0; m < NM; ++m) Gi += tdiv(al[m], mat<R>(lv, m).Gg);
    Gt = tdiv(Gt, Gi);
#pragma unroll
    for (int r = 0; r < 3; ++r)
        sig[r] = solid ? -p * (r == d ? R(1) : R(0)) + R(2) * Gt * dH[r * 3 + d] : -p * (r == d ? R(1) : R(0));
}

// ------------------------------------------------------------------ states
template <int NM, typename R = double>
struct SPrim {
    R a[NM], r[NM], u[3], p, V[9], e[NM];
};
template <int NM, typename R = double>
struct SSt {  // face state along one direction d
    R al[NM], ar[NM], rho, u[3], E, p, a, sig[3], V[9], are[NM];
};

template <int NM, typename R>
__device__ __forceinline__ void load_sprim(const Level& lv, const MF& U, int i, int j, int k, SPrim<NM, R>& P) {
    bool f[3];
    map3(lv, i, j, k, f);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        P.a[m] = (R)U.at(lv.cs[ALPHA] + m, i, j, k);
        P.r[m] = (R)U.at(lv.cs[RHO_K] + m, i, j, k);
        P.e[m] = lv.plastic ? (R)U.at(lv.cs[EPSILON] + m, i, j, k) : R(0);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const R v = (R)U.at(lv.cs[VELOCITY] + r, i, j, k);
        P.u[r] = f[r] ? -v : v;
    }
    P.p = (R)U.at(lv.cs[PRESSURE], i, j, k);
#pragma unroll
    for (int q = 0; q < 9; ++q) P.V[q] = (R)U.at(lv.cs[V_TENSOR] + q, i, j, k);
}

template <int NM, typename Rt>
__device__ __forceinline__ void muscl_s(const SPrim<NM, Rt>& m_, const SPrim<NM, Rt>& c, const SPrim<NM, Rt>& p_,
                                        SPrim<NM, Rt>& L, SPrim<NM, Rt>& R) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        muscl1(m_.a[m], c.a[m], p_.a[m], L.a[m], R.a[m]);
        muscl1(m_.r[m], c.r[m], p_.r[m], L.r[m], R.r[m]);
        muscl1(m_.e[m], c.e[m], p_.e[m], L.e[m], R.e[m]);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) muscl1(m_.u[q], c.u[q], p_.u[q], L.u[q], R.u[q]);
    muscl1(m_.p, c.p, p_.p, L.p, R.p);
#pragma unroll
    for (int q = 0; q < 9; ++q) muscl1(m_.V[q], c.V[q], p_.V[q], L.V[q], R.V[q]);
}

// primitiveToConservative + getSoundSpeed of a reconstructed solid state
// (boxaccesscellarray.cpp:124-174, 254-285), keeping sigma column d.
template <int NM, bool RW, typename R>
__device__ __forceinline__ void p2c_sound_s(const Level& lv, const SPrim<NM, R>& P, int d, SSt<NM, R>& S) {
    R rho = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        S.al[m] = P.a[m];
        S.ar[m] = P.a[m] * P.r[m];
        rho += S.ar[m];
        S.are[m] = mat<R>(lv, m).plastic ? P.e[m] * S.ar[m] : R(0);
    }
    R dH[9];
    const R hj2 = hencky(P.V, dH);
    R ke = R(0);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        S.u[q] = P.u[q];
        ke += R(0.5) * rho * P.u[q] * P.u[q];
    }
    R G, Ent, Pnt;
    eos_sums_s<NM, RW, R>(lv, S.al, S.ar, P.r, hj2, G, Ent, Pnt);
    S.E = P.p * G + Ent - Pnt + ke;
    S.p = P.p;
    S.rho = rho;
    stress_col<NM, RW, R>(lv, S.al, P.r, P.p, dH, d, S.sig);
#pragma unroll
    for (int q = 0; q < 9; ++q) S.V[q] = P.V[q];
    S.a = sound_s<NM, RW, R>(lv, S.al, S.ar, P.r, rho, P.p, hj2);
}

// The phase densities and strains of conservativeToPrimitive alone, formed
// as c2p_s forms them.
template <int NM, typename R>
__device__ __forceinline__ void c2p_rk_e(const Level& lv, const R* al, const R* ar, const R* are, R* rk, R* e) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        rk[m] = tdiv(ar[m], al[m]);
        e[m] = mat<R>(lv, m).plastic ? tdiv(are[m], ar[m]) : R(0);
    }
}

// conservativeToPrimitive (boxaccesscellarray.cpp:67-122) of a solid cell;
// returns HJ2 and the deviatoric Hencky strain in dH.
template <int NM, bool RW, typename R>
__device__ __forceinline__ R c2p_s(const Level& lv, const R* al, const R* ar, const R* are, const R* ru, R E, const R* V,
                                   R* rk, R* e, R& rho, R* u, R& p, R* dH) {
    rho = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        rk[m] = tdiv(ar[m], al[m]);
        rho += ar[m];
        e[m] = mat<R>(lv, m).plastic ? tdiv(are[m], ar[m]) : R(0);
    }
    const R hj2 = hencky(V, dH);
    R ke = R(0);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        u[q] = tdiv(ru[q], rho);
        ke += R(0.5) * rho * u[q] * u[q];
    }
    R G, Ent, Pnt;
    eos_sums_s<NM, RW, R>(lv, al, ar, rk, hj2, G, Ent, Pnt);
    p = tdiv(E - ke - Ent + Pnt, G);
    return hj2;
}

// transverseWaveSpeed of a star state (boxaccesscellarray.cpp:287-324)
template <int NM, bool RW, typename R>
__device__ __forceinline__ R transverse(const Level& lv, const R* al, const R* rk) {
    R b = R(0), g = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m)
        if (mat<R>(lv, m).solid) b += tdiv(tdiv(al[m], mat<R>(lv, m).Gg) * shearG<RW>(mat<R>(lv, m), rk[m]), rk[m]);
    if (tisnan(b) || b < R(0)) b = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) g += tdiv(al[m], mat<R>(lv, m).Gg);
    R t = tsqrt(tdiv(b, g));
    if (tisnan(t)) t = R(0);
    return t;
}

// ------------------------------------------------------------------ 5-wave flux
// Star state of hllc.cpp:12-73 (solid form).
template <int NM, typename R = double>
struct SStar {
    R al[NM], ar[NM], rk[NM], rho, ru[3], u[3], E, V[9], are[NM];
};
template <int NM, typename R>
__device__ __forceinline__ void star_s(const Level& lv, const SSt<NM, R>& K, const R* ruK, R SK, R Ss, int d,
                                       SStar<NM, R>& T) {
    const R mult = (tabs(tdiv(Ss - K.u[d], Ss)) < R(1E-20)) ? R(1) : tdiv_cr(SK - K.u[d], SK - Ss);
    T.rho = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        T.al[m] = K.al[m];
        T.ar[m] = mult * K.ar[m];
        T.rk[m] = tdiv(T.ar[m], T.al[m]);
        T.are[m] = mat<R>(lv, m).plastic ? mult * K.are[m] : R(0);
        T.rho += T.ar[m];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        T.ru[r] = (r == d) ? mult * K.rho * Ss : mult * ruK[r];
        T.u[r] = tdiv(T.ru[r], T.rho);
#pragma unroll
        for (int c = 0; c < 3; ++c) T.V[r * 3 + c] = (r == d) ? K.V[r * 3 + c] : mult * K.V[r * 3 + c];
    }
    T.E = mult * K.rho * (tdiv(K.E, K.rho) + (Ss - K.u[d]) * (Ss - tdiv(K.sig[d], K.rho * (SK - K.u[d]))));
}

// conservative-variable vector of a state (order of the flux layout):
// [0,NM) alpha, [NM,2NM) alphaRho, +3 rhoU, E, +9 V, +NM alphaRhoEpsilon
template <int NM>
struct SCons {
    static constexpr int N = 3 * NM + 13;
};

// Physical flux (flux.cpp:5-47) of the conserved variables of state K.
template <int NM, typename R>
__device__ __forceinline__ void pflux_s(const SSt<NM, R>& K, const R* ruK, int d, R* f) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        f[m] = K.al[m] * K.u[d];
        f[NM + m] = K.ar[m] * K.u[d];
        f[2 * NM + 13 + m] = K.are[m] * K.u[d];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) f[2 * NM + q] = ruK[q] * K.u[d] - K.sig[q];
    R vds = R(0);
#pragma unroll
    for (int q = 0; q < 3; ++q) vds += K.u[q] * K.sig[q];
    f[2 * NM + 3] = K.E * K.u[d] - vds;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) f[2 * NM + 4 + r * 3 + c] = K.V[r * 3 + c] * K.u[d] - K.V[d * 3 + c] * K.u[r];
}

// calc_5Wave_fluxes for one face (hllc.cpp:75-380).  F layout: the conserved
// vector (3 NM + 13), then S*, then VSTAR row d (3).
template <int NM, bool RW, typename Rt>
__device__ inline void wave5(const Level& lv, const SSt<NM, Rt>& L, const SSt<NM, Rt>& R, int d, Rt* F,
                             unsigned& st) {
    constexpr int NC = 3 * NM + 13;
    Rt ruL[3], ruR[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        ruL[q] = L.u[q] * L.rho;
        ruR[q] = R.u[q] * R.rho;
    }
    const Rt SR = tmax(tabs(L.u[d]) + L.a, tabs(R.u[d]) + R.a);
    const Rt SL = -SR;
    const Rt Ss = tdiv(-R.sig[d] + L.sig[d] + ruL[d] * (SL - L.u[d]) - ruR[d] * (SR - R.u[d]),
                       L.rho * (SL - L.u[d]) - R.rho * (SR - R.u[d]));
    if (tisnan(SL) || tisnan(SR) || tisnan(Ss)) st |= 2u;
    F[NC] = Ss;

    SStar<NM, Rt> Ls, Rs;
    Rt SLT = Rt(0), SRT = Rt(0);
    // which: 0 flux of K only, 1 K + K*, 2 K + K* + K**; left/right K
    int kind;
    bool left;
    if (SL >= Rt(0)) {
        star_s<NM, Rt>(lv, L, ruL, SL, Ss, d, Ls);
        kind = 0;
        left = true;
    } else if (Ss >= Rt(0)) {
        star_s<NM, Rt>(lv, L, ruL, SL, Ss, d, Ls);
        SLT = Ss - transverse<NM, RW, Rt>(lv, Ls.al, Ls.rk);
        left = true;
        if (SLT >= Rt(0)) {
            kind = 1;
        } else {
            star_s<NM, Rt>(lv, R, ruR, SR, Ss, d, Rs);
            SRT = Ss + transverse<NM, RW, Rt>(lv, Rs.al, Rs.rk);
            kind = 2;
        }
    } else if (SR >= Rt(0)) {
        star_s<NM, Rt>(lv, R, ruR, SR, Ss, d, Rs);
        SRT = Ss + transverse<NM, RW, Rt>(lv, Rs.al, Rs.rk);
        left = false;
        if (SRT >= Rt(0)) {
            star_s<NM, Rt>(lv, L, ruL, SL, Ss, d, Ls);
            SLT = Ss - transverse<NM, RW, Rt>(lv, Ls.al, Ls.rk);
            kind = 2;
        } else {
            kind = 1;
        }
    } else {
        star_s<NM, Rt>(lv, R, ruR, SR, Ss, d, Rs);
        kind = 0;
        left = false;
    }
    const SSt<NM, Rt>& K = left ? L : R;
    const Rt* ruK = left ? ruL : ruR;
    const SStar<NM, Rt>& Ks = left ? Ls : Rs;
    const Rt SK = left ? SL : SR;
    const Rt SKT = left ? SLT : SRT;

    Rt f[NC];
    pflux_s<NM, Rt>(K, ruK, d, f);
    // alpha: flux of the star state
#pragma unroll
    for (int m = 0; m < NM; ++m) F[m] = Ks.al[m] * Ks.u[d];
    // conserved vectors of K and K*
    Rt uK[NC], uS[NC];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        uK[m] = K.al[m];
        uS[m] = Ks.al[m];
        uK[NM + m] = K.ar[m];
        uS[NM + m] = Ks.ar[m];
        uK[2 * NM + 13 + m] = K.are[m];
        uS[2 * NM + 13 + m] = Ks.are[m];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        uK[2 * NM + q] = ruK[q];
        uS[2 * NM + q] = Ks.ru[q];
    }
    uK[2 * NM + 3] = K.E;
    uS[2 * NM + 3] = Ks.E;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        uK[2 * NM + 4 + q] = K.V[q];
        uS[2 * NM + 4 + q] = Ks.V[q];
    }
    if (kind == 0) {
#pragma unroll
        for (int n = NM; n < NC; ++n) F[n] = f[n];
#pragma unroll
        for (int c = 0; c < 3; ++c) F[NC + 1 + c] = K.V[d * 3 + c];
    } else if (kind == 1) {
#pragma unroll
        for (int n = NM; n < NC; ++n) F[n] = f[n] + SK * (uS[n] - uK[n]);
#pragma unroll
        for (int c = 0; c < 3; ++c) F[NC + 1 + c] = Ks.V[d * 3 + c];
    } else {
        // getStarStarState (hllc.cpp:87-129) with getSigmaStar (hllc.cpp:75-83)
        Rt sigS[3];
        const Rt aL = Ls.rho * (Ss - SLT), aR = Rs.rho * (Ss - SRT);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            sigS[i] = tdiv(Ls.rho * (Ss - SLT) * Rs.rho * (Ss - SRT) * (L.u[i] - R.u[i]) + Ls.rho * (Ss - SLT) * R.sig[i] -
                               Rs.rho * (Ss - SRT) * L.sig[i],
                           aL - aR);
        Rt uss[NC];
#pragma unroll
        for (int n = 0; n < NC; ++n) uss[n] = uS[n];
        Rt ussu[3] = {Ks.u[0], Ks.u[1], Ks.u[2]};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            if (r == d) continue;
            uss[2 * NM + r] += tdiv(sigS[r] - K.sig[r], Ss - SKT);
            ussu[r] = tdiv(uss[2 * NM + r], Ks.rho);
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            if (r == d) continue;
            uss[2 * NM + 3] += tdiv(ussu[r] * sigS[r] - K.u[r] * K.sig[r], Ss - SKT);
#pragma unroll
            for (int c = 0; c < 3; ++c)
                uss[2 * NM + 4 + r * 3 + c] += tdiv(uss[2 * NM + 4 + d * 3 + c] * (ussu[r] - K.u[r]), Ss - SKT);
        }
#pragma unroll
        for (int n = NM; n < NC; ++n) F[n] = f[n] + SK * (uS[n] - uK[n]) + SKT * (uss[n] - uS[n]);
#pragma unroll
        for (int c = 0; c < 3; ++c) F[NC + 1 + c] = uss[2 * NM + 4 + d * 3 + c];
    }
    if (tisnan(SL) || tisnan(SR) || tisnan(Ss) || tisnan(SLT) || tisnan(SRT)) st |= 2u;
}

// ------------------------------------------------------------------ plastic
// boxPlasticUpdate for one cell (plastic.cpp:83-225); V is overwritten with
// the Cayley update even without yield.  Returns status bits.
template <int NM, bool RW, typename R>
__device__ inline unsigned plastic_cell(const Level& lv, const R* al, const R* ar, const R* rk, R* e, R* are, R* V) {
    R ts = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m)
        if (mat<R>(lv, m).solid) ts += al[m];
    if (ts <= 0.5) {
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            e[m] = R(0);
            are[m] = R(0);
        }
        return 0u;
    }
    const double s3 = __builtin_sqrt(3.0), s23 = __builtin_sqrt(2.0 / 3.0);
    R dH[9];
    const R hj2 = hencky(V, dH);
    R JT = R(0), JB = R(0), norm = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatT<R>& M = mat<R>(lv, m);
        R J, Jn;
        if (M.solid) {
            J = tsqrt(hj2);
            if (shearG<RW>(M, rk[m]) * J * s3 > M.yield) Jn = tdiv(M.yield, s3 * shearG<RW>(M, rk[m]));
            else Jn = J;
            e[m] += s23 * (J - Jn);
            are[m] = e[m] * ar[m];
        } else {
            J = R(0);
            Jn = R(0);
            e[m] = R(0);
            are[m] = R(0);
        }
        JB += tdiv(al[m], M.Gg) * J;
        JT += tdiv(al[m], M.Gg) * Jn;
        norm += tdiv(al[m], M.Gg);
    }
    JB *= tdiv(R(1), norm);
    JT *= tdiv(R(1), norm);
    if (JB == 0.0 && JT == 0.0) {
        JB = R(1);
        JT = R(1);
    }
    unsigned st = 0;
    R t1[9], t2[9], t2i[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const double dl = (q / 3 == q % 3) ? 1.0 : 0.0;
        if (dH[q] != 0.0 && JB == 0.0 && JT != 0.0) st |= 16u;
        else dH[q] *= tdiv(JT, JB);
        t1[q] = dl + 0.5 * dH[q];
        t2[q] = dl - 0.5 * dH[q];
    }
    inv3(t2, t2i);
    mm3(t2i, t1, V);
    return st;
}

// cleanUpV of one cell (boxaccesscellarray.cpp:476-528).  R = double in the
// FP64 form; the mixed form evaluates it in FP32 (the square root's Jacobi
// sweeps are its cost; the stored V stays FP64).
__device__ __forceinline__ double trcbrt(double x) { return pm_rcbrt(x); }
__device__ __forceinline__ float trcbrt(float x) { return __builtin_amdgcn_rcpf(cbrtf(x)); }
template <int NM, typename R = double, typename S>
__device__ inline void cleanup_v(const Level& lv, const S* al, S* V) {
    R ta = R(0), tsd = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        ta += (R)al[m];
        if (lv.m[m].solid) tsd += (R)al[m];
    }
    R Vr[9], G[9], Sq[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) Vr[q] = (R)V[q];
    mmT3(Vr, Vr, G);
    sqrtm_spd(G, Sq);
    const R norm = trcbrt(det3(Sq));
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const R dl = (q / 3 == q % 3) ? R(1) : R(0);
        V[q] = (S)tdiv((Sq[q] * norm * tsd) + dl * (ta - tsd), ta);
    }
}

}  // inline namespace TJ_VIEW_NS
}  // namespace tj
