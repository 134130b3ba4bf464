// Elastoplastic (solid) per-cell physics: 3x3 algebra, Hencky strain and
// stress, the Us-up Mie-Gruneisen solid EOS, primitive/conservative maps,
// sound and transverse speeds, the 5-wave HLLC-type flux, the perfect-plastic
// radial return and the V clean-up.  Operation order follows the reference
// files cited beside each function.
#pragma once
#include "tj_fluid.h"

namespace tj {

// ------------------------------------------------------------------ 3x3 algebra
// tensor.cpp:81-115 (products: k = 0,1,2 summed left to right), 140-150
// (trace), 187-233 (adjugate inverse), 296-333 (cofactor determinant).
__device__ __forceinline__ double det2(double a, double b, double c, double d) { return a * d - b * c; }
__device__ __forceinline__ double det3(const double* M) {
    double det = 0.0;
    det += M[0] * det2(M[4], M[5], M[7], M[8]);
    det += (-M[1]) * det2(M[3], M[5], M[6], M[8]);
    det += M[2] * det2(M[3], M[4], M[6], M[7]);
    return det;
}
__device__ __forceinline__ void inv3(const double* M, double* Mi) {
    const double D = det3(M);
    double cof[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int k0 = i == 0 ? 1 : 0, k1 = i == 2 ? 1 : 2;
            const int m0 = j == 0 ? 1 : 0, m1 = j == 2 ? 1 : 2;
            const double sign = ((i + j) % 2) ? -1.0 : 1.0;
            cof[i * 3 + j] = sign * det2(M[k0 * 3 + m0], M[k0 * 3 + m1], M[k1 * 3 + m0], M[k1 * 3 + m1]);
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Mi[i * 3 + j] = cof[j * 3 + i] / D;
}
__device__ __forceinline__ void mm3(const double* A, const double* B, double* C) {  // C = A B
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            C[r * 3 + c] = (A[r * 3] * B[c] + A[r * 3 + 1] * B[3 + c]) + A[r * 3 + 2] * B[6 + c];
}
__device__ __forceinline__ void mmT3(const double* A, const double* B, double* C) {  // C = A B^T
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            C[r * 3 + c] = (A[r * 3] * B[c * 3] + A[r * 3 + 1] * B[c * 3 + 1]) + A[r * 3 + 2] * B[c * 3 + 2];
}

// getDeviatoricHenckyStrain + getHenckyJ2 (boxaccesscellarray.cpp:326-380):
// devH = (G - G^-1)/4 with G = V V^T, HJ2 = tr(devH devH^T).
__device__ __forceinline__ double hencky(const double* V, double* dH) {
    double G[9], Gi[9];
    mmT3(V, V, G);
    inv3(G, Gi);
#pragma unroll
    for (int q = 0; q < 9; ++q) dH[q] = 0.5 * 0.5 * (G[q] - Gi[q]);
    double t = 0.0;
#pragma unroll
    for (int r = 0; r < 3; ++r) t += (dH[r * 3] * dH[r * 3] + dH[r * 3 + 1] * dH[r * 3 + 1]) + dH[r * 3 + 2] * dH[r * 3 + 2];
    return t;
}

// Principal square root of an SPD 3x3 matrix by cyclic Jacobi rotations
// (Golub & Van Loan, Matrix Computations, 8.5), six sweeps; the oracle
// evaluates the same sequence.
__device__ inline void sqrtm_spd(const double* Ain, double* S) {
    double a[9], q[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        a[i] = Ain[i];
        q[i] = (i / 3 == i % 3) ? 1.0 : 0.0;
    }
    for (int sweep = 0; sweep < 6; ++sweep) {
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) {
            const int p = rr == 2 ? 1 : 0, qq = rr == 0 ? 1 : 2;
            const double apq = a[p * 3 + qq];
            if (apq == 0.0) continue;
            const double app = a[p * 3 + p], aqq = a[qq * 3 + qq];
            const double theta = (aqq - app) / (2.0 * apq);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) / (__builtin_fabs(theta) + __builtin_sqrt(theta * theta + 1.0));
            const double c = 1.0 / __builtin_sqrt(t * t + 1.0);
            const double s = t * c;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double akp = a[k * 3 + p], akq = a[k * 3 + qq];
                a[k * 3 + p] = c * akp - s * akq;
                a[k * 3 + qq] = s * akp + c * akq;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double apk = a[p * 3 + k], aqk = a[qq * 3 + k];
                a[p * 3 + k] = c * apk - s * aqk;
                a[qq * 3 + k] = s * apk + c * aqk;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double qkp = q[k * 3 + p], qkq = q[k * 3 + qq];
                q[k * 3 + p] = c * qkp - s * qkq;
                q[k * 3 + qq] = s * qkp + c * qkq;
            }
        }
    }
    const double sl0 = __builtin_sqrt(a[0]), sl1 = __builtin_sqrt(a[4]), sl2 = __builtin_sqrt(a[8]);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            S[r * 3 + c] = (q[r * 3] * sl0 * q[c * 3] + q[r * 3 + 1] * sl1 * q[c * 3 + 1]) + q[r * 3 + 2] * sl2 * q[c * 3 + 2];
}

// ------------------------------------------------------------------ Us-up solid EOS
// Hugoniot reference curve p_H = rho0 c0^2 eta/(1 - s eta)^2, e_H = p_H eta/(2 rho0),
// eta = 1 - rho0/rho, with the Mie-Gruneisen form p = p_H + Gamma rho (e - e_H).
__device__ __forceinline__ double pH(const MatC& M, double rho) {
    const double eta = 1.0 - M.rho0 / rho;
    const double den = 1.0 - M.s * eta;
    return M.rho0 * M.c0 * M.c0 * eta / (den * den);
}
__device__ __forceinline__ double eH(const MatC& M, double rho) {
    const double eta = 1.0 - M.rho0 / rho;
    return pH(M, rho) * eta / (2.0 * M.rho0);
}
__device__ __forceinline__ double dpH(const MatC& M, double rho) {
    const double eta = 1.0 - M.rho0 / rho;
    const double den = 1.0 - M.s * eta;
    const double detadrho = M.rho0 / (rho * rho);
    return M.rho0 * M.c0 * M.c0 * (1.0 + M.s * eta) / (den * den * den) * detadrho;
}
__device__ __forceinline__ double deH(const MatC& M, double rho) {
    const double eta = 1.0 - M.rho0 / rho;
    const double detadrho = M.rho0 / (rho * rho);
    return (dpH(M, rho) * eta + pH(M, rho) * detadrho) / (2.0 * M.rho0);
}
__device__ __forceinline__ double shearG(const MatC& M) { return M.eos == 1 ? M.G0 : 0.0; }

// effective sums of boxaccesscellarray.cpp:218-252 with the solid terms
template <int NM>
__device__ __forceinline__ void eos_sums_s(const Level& lv, const double* a, const double* ar, const double* rk, double hj2,
                                           double& G, double& Ent, double& Pnt) {
    G = 0.0;
    Ent = 0.0;
    Pnt = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) G += a[m] / lv.m[m].Gg;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatC& M = lv.m[m];
        if (M.eos == 1) Ent += (ar[m] * eH(M, rk[m]) + a[m] * (shearG(M) * hj2));
        else Ent += (M.eref * ar[m] + 0.0 * a[m]);
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatC& M = lv.m[m];
        if (M.eos == 1) Pnt += (a[m] / M.Gg) * pH(M, rk[m]) + (a[m] / M.Gg) * ((rk[m] * 0.0 - shearG(M)) * hj2);
        else Pnt += M.pref * a[m] / M.Gg + 0.0 * a[m];
    }
}
// getSoundSpeedContribution: stiffened gas (equationofstate.cpp:61-64) or the
// Us-up solid with c^2 = dp/drho|_s + 4G/(3 rho)
__device__ __forceinline__ double c2_s(const MatC& M, double p, double rk, double hj2) {
    if (M.eos == 1)
        return (p * M.gp1 - pH(M, rk) + M.gp1 * M.G0 * hj2) / rk + dpH(M, rk) - M.Gg * rk * deH(M, rk) + (4.0 / 3.0) * M.G0 / rk;
    return p * M.gp1 / rk - M.pref / rk;
}
template <int NM>
__device__ __forceinline__ double sound_s(const Level& lv, const double* al, const double* ar, const double* rk, double rho,
                                          double p, double hj2) {
    double a = 0.0, xt = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatC& M = lv.m[m];
        a += dmax(0.0, M.xi * c2_s(M, p, rk[m], hj2) * ar[m] / rho);
        xt += M.xi * al[m];
    }
    if (a <= 0.0) a = 1E-6;
    return __builtin_sqrt(a / xt);
}
// stressTensor column d (boxaccesscellarray.cpp:176-216)
template <int NM>
__device__ __forceinline__ void stress_col(const Level& lv, const double* al, double p, const double* dH, int d, double* sig) {
    double Gt = 0.0, Gi = 0.0;
    int solid = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m)
        if (lv.m[m].solid) {
            solid = 1;
            Gt += shearG(lv.m[m]) * (al[m] / lv.m[m].Gg);
        }
#pragma unroll
    for (int m = 0; m < NM; ++m) Gi += al[m] / lv.m[m].Gg;
    Gt = Gt / Gi;
#pragma unroll
    for (int r = 0; r < 3; ++r)
        sig[r] = solid ? -p * (r == d ? 1.0 : 0.0) + 2.0 * Gt * dH[r * 3 + d] : -p * (r == d ? 1.0 : 0.0);
}

// ------------------------------------------------------------------ states
template <int NM>
struct SPrim {
    double a[NM], r[NM], u[3], p, V[9], e[NM];
};
template <int NM>
struct SSt {  // face state along one direction d
    double al[NM], ar[NM], rho, u[3], E, p, a, sig[3], V[9], are[NM];
};

template <int NM>
__device__ __forceinline__ void load_sprim(const Level& lv, const MF& U, int i, int j, int k, SPrim<NM>& P) {
    bool f[3];
    map3(lv, i, j, k, f);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        P.a[m] = U.at(lv.cs[ALPHA] + m, i, j, k);
        P.r[m] = U.at(lv.cs[RHO_K] + m, i, j, k);
        P.e[m] = lv.plastic ? U.at(lv.cs[EPSILON] + m, i, j, k) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const double v = U.at(lv.cs[VELOCITY] + r, i, j, k);
        P.u[r] = f[r] ? -v : v;
    }
    P.p = U.at(lv.cs[PRESSURE], i, j, k);
#pragma unroll
    for (int q = 0; q < 9; ++q) P.V[q] = U.at(lv.cs[V_TENSOR] + q, i, j, k);
}

template <int NM>
__device__ __forceinline__ void muscl_s(const SPrim<NM>& m_, const SPrim<NM>& c, const SPrim<NM>& p_, SPrim<NM>& L, SPrim<NM>& R) {
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
template <int NM>
__device__ __forceinline__ void p2c_sound_s(const Level& lv, const SPrim<NM>& P, int d, SSt<NM>& S) {
    double rho = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        S.al[m] = P.a[m];
        S.ar[m] = P.a[m] * P.r[m];
        rho += S.ar[m];
        S.are[m] = lv.m[m].plastic ? P.e[m] * S.ar[m] : 0.0;
    }
    double dH[9];
    const double hj2 = hencky(P.V, dH);
    double ke = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        S.u[q] = P.u[q];
        ke += 0.5 * rho * P.u[q] * P.u[q];
    }
    double G, Ent, Pnt;
    eos_sums_s<NM>(lv, S.al, S.ar, P.r, hj2, G, Ent, Pnt);
    S.E = P.p * G + Ent - Pnt + ke;
    S.p = P.p;
    S.rho = rho;
    stress_col<NM>(lv, S.al, P.p, dH, d, S.sig);
#pragma unroll
    for (int q = 0; q < 9; ++q) S.V[q] = P.V[q];
    S.a = sound_s<NM>(lv, S.al, S.ar, P.r, rho, P.p, hj2);
}

// conservativeToPrimitive (boxaccesscellarray.cpp:67-122) of a solid cell;
// returns HJ2 and the deviatoric Hencky strain in dH.
template <int NM>
__device__ __forceinline__ double c2p_s(const Level& lv, const double* al, const double* ar, const double* are, const double* ru,
                                        double E, const double* V, double* rk, double* e, double& rho, double* u, double& p,
                                        double* dH) {
    rho = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        rk[m] = ar[m] / al[m];
        rho += ar[m];
        e[m] = lv.m[m].plastic ? are[m] / ar[m] : 0.0;
    }
    const double hj2 = hencky(V, dH);
    double ke = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        u[q] = ru[q] / rho;
        ke += 0.5 * rho * u[q] * u[q];
    }
    double G, Ent, Pnt;
    eos_sums_s<NM>(lv, al, ar, rk, hj2, G, Ent, Pnt);
    p = (E - ke - Ent + Pnt) / G;
    return hj2;
}

// transverseWaveSpeed of a star state (boxaccesscellarray.cpp:287-324)
template <int NM>
__device__ __forceinline__ double transverse(const Level& lv, const double* al, const double* rk) {
    double b = 0.0, g = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m)
        if (lv.m[m].solid) b += (al[m] / lv.m[m].Gg) * shearG(lv.m[m]) / rk[m];
    if (__builtin_isnan(b) || b < 0.0) b = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) g += al[m] / lv.m[m].Gg;
    double t = __builtin_sqrt(b / g);
    if (__builtin_isnan(t)) t = 0.0;
    return t;
}

// ------------------------------------------------------------------ 5-wave flux
// Star state of hllc.cpp:12-73 (solid form).
template <int NM>
struct SStar {
    double al[NM], ar[NM], rk[NM], rho, ru[3], u[3], E, V[9], are[NM];
};
template <int NM>
__device__ __forceinline__ void star_s(const Level& lv, const SSt<NM>& K, const double* ruK, double SK, double Ss, int d,
                                       SStar<NM>& T) {
    const double mult = (__builtin_fabs((Ss - K.u[d]) / Ss) < 1E-20) ? 1.0 : (SK - K.u[d]) / (SK - Ss);
    T.rho = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        T.al[m] = K.al[m];
        T.ar[m] = mult * K.ar[m];
        T.rk[m] = T.ar[m] / T.al[m];
        T.are[m] = lv.m[m].plastic ? mult * K.are[m] : 0.0;
        T.rho += T.ar[m];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        T.ru[r] = (r == d) ? mult * K.rho * Ss : mult * ruK[r];
        T.u[r] = T.ru[r] / T.rho;
#pragma unroll
        for (int c = 0; c < 3; ++c) T.V[r * 3 + c] = (r == d) ? K.V[r * 3 + c] : mult * K.V[r * 3 + c];
    }
    T.E = mult * K.rho * (K.E / K.rho + (Ss - K.u[d]) * (Ss - (K.sig[d]) / (K.rho * (SK - K.u[d]))));
}

// conservative-variable vector of a state (order of the flux layout):
// [0,NM) alpha, [NM,2NM) alphaRho, +3 rhoU, E, +9 V, +NM alphaRhoEpsilon
template <int NM>
struct SCons {
    static constexpr int N = 3 * NM + 13;
};

// Physical flux (flux.cpp:5-47) of the conserved variables of state K.
template <int NM>
__device__ __forceinline__ void pflux_s(const SSt<NM>& K, const double* ruK, int d, double* f) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        f[m] = K.al[m] * K.u[d];
        f[NM + m] = K.ar[m] * K.u[d];
        f[2 * NM + 13 + m] = K.are[m] * K.u[d];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) f[2 * NM + q] = ruK[q] * K.u[d] - K.sig[q];
    double vds = 0.0;
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
template <int NM>
__device__ inline void wave5(const Level& lv, const SSt<NM>& L, const SSt<NM>& R, int d, double* F, unsigned& st) {
    constexpr int NC = 3 * NM + 13;
    double ruL[3], ruR[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        ruL[q] = L.u[q] * L.rho;
        ruR[q] = R.u[q] * R.rho;
    }
    const double SR = dmax(__builtin_fabs(L.u[d]) + L.a, __builtin_fabs(R.u[d]) + R.a);
    const double SL = -SR;
    const double Ss = (-R.sig[d] + L.sig[d] + ruL[d] * (SL - L.u[d]) - ruR[d] * (SR - R.u[d])) /
                      (L.rho * (SL - L.u[d]) - R.rho * (SR - R.u[d]));
    if (__builtin_isnan(SL) || __builtin_isnan(SR) || __builtin_isnan(Ss)) st |= 2u;
    F[NC] = Ss;

    SStar<NM> Ls, Rs;
    double SLT = 0.0, SRT = 0.0;
    // which: 0 flux of K only, 1 K + K*, 2 K + K* + K**; left/right K
    int kind;
    bool left;
    if (SL >= 0.0) {
        star_s<NM>(lv, L, ruL, SL, Ss, d, Ls);
        kind = 0;
        left = true;
    } else if (Ss >= 0.0) {
        star_s<NM>(lv, L, ruL, SL, Ss, d, Ls);
        SLT = Ss - transverse<NM>(lv, Ls.al, Ls.rk);
        left = true;
        if (SLT >= 0.0) {
            kind = 1;
        } else {
            star_s<NM>(lv, R, ruR, SR, Ss, d, Rs);
            SRT = Ss + transverse<NM>(lv, Rs.al, Rs.rk);
            kind = 2;
        }
    } else if (SR >= 0.0) {
        star_s<NM>(lv, R, ruR, SR, Ss, d, Rs);
        SRT = Ss + transverse<NM>(lv, Rs.al, Rs.rk);
        left = false;
        if (SRT >= 0.0) {
            star_s<NM>(lv, L, ruL, SL, Ss, d, Ls);
            SLT = Ss - transverse<NM>(lv, Ls.al, Ls.rk);
            kind = 2;
        } else {
            kind = 1;
        }
    } else {
        star_s<NM>(lv, R, ruR, SR, Ss, d, Rs);
        kind = 0;
        left = false;
    }
    const SSt<NM>& K = left ? L : R;
    const double* ruK = left ? ruL : ruR;
    const SStar<NM>& Ks = left ? Ls : Rs;
    const double SK = left ? SL : SR;
    const double SKT = left ? SLT : SRT;

    double f[NC];
    pflux_s<NM>(K, ruK, d, f);
    // alpha: flux of the star state
#pragma unroll
    for (int m = 0; m < NM; ++m) F[m] = Ks.al[m] * Ks.u[d];
    // conserved vectors of K and K*
    double uK[NC], uS[NC];
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
        double sigS[3];
        const double aL = Ls.rho * (Ss - SLT), aR = Rs.rho * (Ss - SRT);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            sigS[i] = (Ls.rho * (Ss - SLT) * Rs.rho * (Ss - SRT) * (L.u[i] - R.u[i]) + Ls.rho * (Ss - SLT) * R.sig[i] -
                       Rs.rho * (Ss - SRT) * L.sig[i]) /
                      (aL - aR);
        double uss[NC];
#pragma unroll
        for (int n = 0; n < NC; ++n) uss[n] = uS[n];
        double ussu[3] = {Ks.u[0], Ks.u[1], Ks.u[2]};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            if (r == d) continue;
            uss[2 * NM + r] += (sigS[r] - K.sig[r]) / (Ss - SKT);
            ussu[r] = uss[2 * NM + r] / Ks.rho;
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            if (r == d) continue;
            uss[2 * NM + 3] += (ussu[r] * sigS[r] - K.u[r] * K.sig[r]) / (Ss - SKT);
#pragma unroll
            for (int c = 0; c < 3; ++c)
                uss[2 * NM + 4 + r * 3 + c] += (uss[2 * NM + 4 + d * 3 + c] * (ussu[r] - K.u[r])) / (Ss - SKT);
        }
#pragma unroll
        for (int n = NM; n < NC; ++n) F[n] = f[n] + SK * (uS[n] - uK[n]) + SKT * (uss[n] - uS[n]);
#pragma unroll
        for (int c = 0; c < 3; ++c) F[NC + 1 + c] = uss[2 * NM + 4 + d * 3 + c];
    }
    if (__builtin_isnan(SL) || __builtin_isnan(SR) || __builtin_isnan(Ss) || __builtin_isnan(SLT) || __builtin_isnan(SRT)) st |= 2u;
}

// ------------------------------------------------------------------ plastic
// boxPlasticUpdate for one cell (plastic.cpp:83-225); V is overwritten with
// the Cayley update even without yield.  Returns status bits.
template <int NM>
__device__ inline unsigned plastic_cell(const Level& lv, const double* al, const double* ar, double* e, double* are, double* V) {
    double ts = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m)
        if (lv.m[m].solid) ts += al[m];
    if (ts <= 0.5) {
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            e[m] = 0.0;
            are[m] = 0.0;
        }
        return 0u;
    }
    const double s3 = __builtin_sqrt(3.0), s23 = __builtin_sqrt(2.0 / 3.0);
    double dH[9];
    const double hj2 = hencky(V, dH);
    double JT = 0.0, JB = 0.0, norm = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatC& M = lv.m[m];
        double J, Jn;
        if (M.solid) {
            J = __builtin_sqrt(hj2);
            if (shearG(M) * J * s3 > M.yield) Jn = M.yield / (s3 * shearG(M));
            else Jn = J;
            e[m] += s23 * (J - Jn);
            are[m] = e[m] * ar[m];
        } else {
            J = 0.0;
            Jn = 0.0;
            e[m] = 0.0;
            are[m] = 0.0;
        }
        JB += (al[m] / M.Gg) * J;
        JT += (al[m] / M.Gg) * Jn;
        norm += (al[m] / M.Gg);
    }
    JB *= 1.0 / norm;
    JT *= 1.0 / norm;
    if (JB == 0.0 && JT == 0.0) {
        JB = 1.0;
        JT = 1.0;
    }
    unsigned st = 0;
    double t1[9], t2[9], t2i[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const double dl = (q / 3 == q % 3) ? 1.0 : 0.0;
        if (dH[q] != 0.0 && JB == 0.0 && JT != 0.0) st |= 16u;
        else dH[q] *= JT / JB;
        t1[q] = dl + 0.5 * dH[q];
        t2[q] = dl - 0.5 * dH[q];
    }
    inv3(t2, t2i);
    mm3(t2i, t1, V);
    return st;
}

// cleanUpV of one cell (boxaccesscellarray.cpp:476-528)
template <int NM>
__device__ inline void cleanup_v(const Level& lv, const double* al, double* V) {
    double ta = 0.0, tsd = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        ta += al[m];
        if (lv.m[m].solid) tsd += al[m];
    }
    double G[9], S[9];
    mmT3(V, V, G);
    sqrtm_spd(G, S);
    const double norm = pm_rcbrt(det3(S));
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const double dl = (q / 3 == q % 3) ? 1.0 : 0.0;
        V[q] = ((S[q] * norm * tsd) + dl * (ta - tsd)) / ta;
    }
}

}  // namespace tj
