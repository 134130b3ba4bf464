// Fluid (stiffened-gas / constant-Gamma Mie-Gruneisen) reconstruction, EOS
// and HLLC flux, and the directional flux kernels.
#pragma once
#include "tj_device.h"

namespace tj {

// ------------------------------------------------------------------ states
template <int NM>
struct Prim {  // primitive variables reconstructed by MUSCL (accesspattern.cpp:44-79)
    double a[NM], r[NM], u[3], p;
};

template <int NM>
struct FSt {  // face state after primitiveToConservative + getSoundSpeed
    double al[NM], ar[NM], rho, u[3], E, p, a;
};

// α-reconstruction of one cell for THINC/BVD (thinc.cpp:67-186)
template <int NM>
struct ARec {
    double ula[NM], ura[NM];  // MUSCL face alphas
    double thl[NM], thr[NM];  // THINC face alphas (valid where mixed)
    int mixed;                // bit m: mixedCell for material m
};

// ------------------------------------------------------------------ loads
__device__ __forceinline__ void map3(const Level& lv, int& i, int& j, int& k, bool f[3]) {
    i = bcmap(i, lv.n[0], lv.bc_lo[0], lv.bc_hi[0], f[0]);
    if (lv.dim > 1) j = bcmap(j, lv.n[1], lv.bc_lo[1], lv.bc_hi[1], f[1]); else f[1] = false;
    if (lv.dim > 2) k = bcmap(k, lv.n[2], lv.bc_lo[2], lv.bc_hi[2], f[2]); else f[2] = false;
}

template <int NM>
__device__ __forceinline__ void load_prim(const Level& lv, const MF& U, int i, int j, int k, Prim<NM>& P) {
    bool f[3];
    map3(lv, i, j, k, f);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        P.a[m] = U.at(lv.cs[ALPHA] + m, i, j, k);
        P.r[m] = U.at(lv.cs[RHO_K] + m, i, j, k);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double v = U.at(lv.cs[VELOCITY] + r, i, j, k);
        P.u[r] = f[r] ? -v : v;
    }
    P.p = U.at(lv.cs[PRESSURE], i, j, k);
}

__device__ __forceinline__ double load_alpha(const Level& lv, const MF& U, int m, int i, int j, int k) {
    bool f[3];
    map3(lv, i, j, k, f);
    return U.at(lv.cs[ALPHA] + m, i, j, k);
}

// ------------------------------------------------------------------ MUSCL
// hllc.cpp:527-575.  min(2/(1+r), 2r/(1+r)) is evaluated with one division:
// both quotients share the divisor, and 2r < 2 exactly when r < 1, so the
// smaller quotient is the one with the smaller numerator (ties are equal).
__device__ __forceinline__ double vanleer(double r) {
    if (r <= 0.0) return 0.0;
    double num = (r < 1.0) ? 2.0 * r : 2.0;
    return num / (1.0 + r);
}
__device__ __forceinline__ void muscl1(double um, double uc, double up, double& L, double& R) {
    double r = (uc - um) / (up - uc);
    if (__builtin_isinf(r) || __builtin_isnan(r)) r = 0.0;
    double g = vanleer(r) * 0.5 * (up - um);
    L = uc - 0.5 * g;
    R = uc + 0.5 * g;
}
template <int NM>
__device__ __forceinline__ void muscl(const Prim<NM>& m_, const Prim<NM>& c, const Prim<NM>& p_, Prim<NM>& L, Prim<NM>& R) {
#pragma unroll
    for (int m = 0; m < NM; ++m) muscl1(m_.a[m], c.a[m], p_.a[m], L.a[m], R.a[m]);
#pragma unroll
    for (int m = 0; m < NM; ++m) muscl1(m_.r[m], c.r[m], p_.r[m], L.r[m], R.r[m]);
#pragma unroll
    for (int q = 0; q < 3; ++q) muscl1(m_.u[q], c.u[q], p_.u[q], L.u[q], R.u[q]);
    muscl1(m_.p, c.p, p_.p, L.p, R.p);
}

// ------------------------------------------------------------------ EOS (kind 0)
// equationofstate.cpp:37-64 and boxaccesscellarray.cpp:218-252.
template <int NM>
__device__ __forceinline__ void eos_sums(const Level& lv, const double* a, const double* ar, double& G, double& Ent, double& Pnt) {
    G = 0.0;
    Ent = 0.0;
    Pnt = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) G += a[m] / lv.m[m].Gg;
#pragma unroll
    for (int m = 0; m < NM; ++m) Ent += (lv.m[m].eref * ar[m] + 0.0 * a[m]);
#pragma unroll
    for (int m = 0; m < NM; ++m) Pnt += lv.m[m].pref * a[m] / lv.m[m].Gg + 0.0 * a[m];
}

// primitiveToConservative + getSoundSpeed of one reconstructed state
// (boxaccesscellarray.cpp:124-174, 254-285).
template <int NM>
__device__ __forceinline__ void p2c_sound(const Level& lv, const Prim<NM>& P, FSt<NM>& S) {
    double rho = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        S.al[m] = P.a[m];
        S.ar[m] = P.a[m] * P.r[m];
        rho += S.ar[m];
    }
    double ke = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        S.u[q] = P.u[q];
        ke += 0.5 * rho * P.u[q] * P.u[q];
    }
    double G, Ent, Pnt;
    eos_sums<NM>(lv, S.al, S.ar, G, Ent, Pnt);
    S.E = P.p * G + Ent - Pnt + ke;
    S.p = P.p;
    S.rho = rho;
    double a = 0.0, xt = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatC& M = lv.m[m];
        double c2 = P.p * M.gp1 / P.r[m] - M.pref / P.r[m];
        a += dmax(0.0, M.xi * c2 * S.ar[m] / rho);
        xt += M.xi * S.al[m];
    }
    if (a <= 0.0) a = 1E-6;
    S.a = __builtin_sqrt(a / xt);
}

// conservativeToPrimitive (boxaccesscellarray.cpp:67-122), fluid part.
template <int NM>
__device__ __forceinline__ void c2p(const Level& lv, const double* al, const double* ar, const double* ru, double E,
                                    double* rk, double& rho, double* u, double& p) {
    rho = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        rk[m] = ar[m] / al[m];
        rho += ar[m];
    }
    double ke = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        u[q] = ru[q] / rho;
        ke += 0.5 * rho * u[q] * u[q];
    }
    double G, Ent, Pnt;
    eos_sums<NM>(lv, al, ar, G, Ent, Pnt);
    p = (E - ke - Ent + Pnt) / G;
}

// getSoundSpeed of a cell state (boxaccesscellarray.cpp:254-285).
template <int NM>
__device__ __forceinline__ double sound(const Level& lv, const double* al, const double* ar, const double* rk, double rho, double p) {
    double a = 0.0, xt = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatC& M = lv.m[m];
        double c2 = p * M.gp1 / rk[m] - M.pref / rk[m];
        a += dmax(0.0, M.xi * c2 * ar[m] / rho);
        xt += M.xi * al[m];
    }
    if (a <= 0.0) a = 1E-6;
    return __builtin_sqrt(a / xt);
}

// ------------------------------------------------------------------ HLLC
// hllc.cpp:5-73 (getSstar, getStarState), 384-487 (calc_fluxes), flux.cpp:5-47.
// F layout: [0,NM) alpha, [NM,2NM) alphaRho, [2NM,2NM+3) rhoU, 2NM+3 E, 2NM+4 S*.
template <int NM>
__device__ __forceinline__ void hllc(const FSt<NM>& L, const FSt<NM>& R, int d, double* F, double* sa) {
    const double sigL = -L.p * 1.0, sigR = -R.p * 1.0;  // sigma_dd = -p delta_dd
    double ruL[3], ruR[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        ruL[q] = L.u[q] * L.rho;
        ruR[q] = R.u[q] * R.rho;
    }
    const double SR = dmax(__builtin_fabs(L.u[d]) + L.a, __builtin_fabs(R.u[d]) + R.a);
    const double SL = -SR;
    const double Ss = (-sigR + sigL + ruL[d] * (SL - L.u[d]) - ruR[d] * (SR - R.u[d])) /
                      (L.rho * (SL - L.u[d]) - R.rho * (SR - R.u[d]));
    F[2 * NM + 4] = Ss;
    // branch selection of calc_fluxes; K is the state the star state is built from
    const bool useL = (SL >= 0.0) || (Ss >= 0.0);
    const bool addStar = !(SL >= 0.0) && ((Ss >= 0.0) || (SR >= 0.0));
    const FSt<NM>& K = useL ? L : R;
    const double* ruK = useL ? ruL : ruR;
    const double SK = useL ? SL : SR;
    const double sigK = useL ? sigL : sigR;
    // star state
    const double mult = (__builtin_fabs((Ss - K.u[d]) / Ss) < 1E-20) ? 1.0 : (SK - K.u[d]) / (SK - Ss);
    double arS[NM], rhoS = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        arS[m] = mult * K.ar[m];
        rhoS += arS[m];
    }
    double ruS[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) ruS[q] = (q == d) ? mult * K.rho * Ss : mult * ruK[q];
    const double uSd = ruS[d] / rhoS;
    const double ES = mult * K.rho * (K.E / K.rho + (Ss - K.u[d]) * (Ss - (sigK) / (K.rho * (SK - K.u[d]))));
    // physical flux of K (flux.cpp:33-47) with sigma = -p I
    double vds = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) vds += K.u[q] * (-K.p * (q == d ? 1.0 : 0.0));
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        F[m] = K.al[m] * uSd;
        sa[m] = K.al[m];
        double f = K.ar[m] * K.u[d];
        F[NM + m] = addStar ? f + SK * (arS[m] - K.ar[m]) : f;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        double f = ruK[q] * K.u[d] - (-K.p * (q == d ? 1.0 : 0.0));
        F[2 * NM + q] = addStar ? f + SK * (ruS[q] - ruK[q]) : f;
    }
    {
        double f = K.E * K.u[d] - vds;
        F[2 * NM + 3] = addStar ? f + SK * (ES - K.E) : f;
    }
}

template <int NM>
__device__ __forceinline__ void flip_state(FSt<NM>& S, int d) { S.u[d] = -S.u[d]; }

// ------------------------------------------------------------------ THINC
// thinc.cpp:54-88 (mixedCell, Youngs normal in 2D), 130-186 (pass 1).
// cell (i,j,k) along direction d; alphas are read from U with the boundary map.
template <int NM>
__device__ inline void arec(const Level& lv, const MF& U, int i, int j, int k, int d, double ynx, double yny,
                            ARec<NM>& A) {
    int di = d == 0, dj = d == 1, dk = d == 2;
    A.mixed = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const double aL = load_alpha(lv, U, m, i - di, j - dj, k - dk);
        const double aC = load_alpha(lv, U, m, i, j, k);
        const double aR = load_alpha(lv, U, m, i + di, j + dj, k + dk);
        muscl1(aL, aC, aR, A.ula[m], A.ura[m]);
        const double mn = dmin(aL, aR);
        const double mx = dmax(aL, aR) - mn;
        const double theta = sgnd(aR - aL);
        const double eps = 1E-20;
        const double C = (mx < eps) ? 0.0 : (aC - mn + eps) / (mx + eps);
        const double dl = 1E-4;
        const bool mixed = (C > dl) && (C < (0.999999 - dl)) && ((aR - aC) * (aC - aL) >= 0.0);
        A.thl[m] = 0.0;
        A.thr[m] = 0.0;
        if (mixed) {
            A.mixed |= 1 << m;
            double nvc = 1.0;
            if (lv.dim == 2) {
                auto a = [&](int ddi, int ddj) { return load_alpha(lv, U, m, i + ddi, j + ddj, k); };
                const double nx = ynx * (a(1, 1) + 2.0 * a(1, 0) + a(1, -1) - a(-1, 1) - 2.0 * a(-1, 0) - a(-1, -1));
                const double ny = yny * (a(1, 1) + 2.0 * a(0, 1) + a(-1, 1) - a(-1, -1) - 2.0 * a(0, -1) - a(1, -1));
                const double norm = __builtin_sqrt(nx * nx + ny * ny);
                if (norm < 1E-20) nvc = 0;
                else nvc = (d == 0) ? __builtin_fabs(nx / norm) : __builtin_fabs(ny / norm);
            }
            const double beta = nvc * lv.beta0 + 0.01;
            const double coshB = pm_cosh(beta), tanhB = pm_tanh(beta);
            const double B = pm_exp(theta * beta * ((2.0 * C) - 1.0));
            const double Ac = ((B / coshB) - 1.0) / tanhB;
            A.thr[m] = mn + (mx / 2.0) * (1.0 + theta * ((tanhB + Ac) / (1.0 + Ac * tanhB)));
            A.thl[m] = mn + (mx / 2.0) * (1.0 + theta * Ac);
        }
    }
}

// TBV of thinc.cpp:57-63 for one material.
__device__ __forceinline__ double tbv(double URl, double TRl, double ULr, double TLr, double PL, double PR) {
    return dmin(dmin(__builtin_fabs(URl - PL) + __builtin_fabs(PR - ULr), __builtin_fabs(TRl - PL) + __builtin_fabs(PR - TLr)),
                dmin(__builtin_fabs(URl - PL) + __builtin_fabs(PR - TLr), __builtin_fabs(TRl - PL) + __builtin_fabs(PR - ULr)));
}

// cleanUpAlpha of one cell (boxaccesscellarray.cpp:530-599), the Nan-index
// quirk of :553-556 included.  Returns TJ status bits.
template <int NM>
__device__ __forceinline__ unsigned cleanup_alpha(double* a) {
    double tot = 0.0;
    int nan = 0;
    unsigned st = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        if (__builtin_isnan(a[m])) {
            a[m] = 0.0;
            if (nan == 0) nan = m;
            else st |= 4u;
        } else if (a[m] < 0.0) {
            a[m] = 1E-6;
        } else if (a[m] > 1.0) {
            a[m] = 1.0;
        }
        tot += a[m];
    }
    if (nan > 0) {
#pragma unroll
        for (int m = 0; m < NM; ++m)
            if (m == nan) {
                a[m] = 1.0 - tot;
                tot += a[m];
            }
    }
    if (tot <= 0.0) st |= 8u;
#pragma unroll
    for (int m = 0; m < NM; ++m) a[m] *= 1.0 / tot;
    return st;
}

// BVD selection for one cell (thinc.cpp:191-238) given the neighbour data.
//   nb*: neighbour values with the reference's box-edge semantics already applied
//   own thl/thr: THINC values of this cell where mixed, else the stale history
template <int NM>
__device__ __forceinline__ void bvd(int mixedFlag, const double* URl, const double* TRl, const double* ULr, const double* TLr,
                                    const double* thl, const double* thr, double* ula, double* ura) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const double tm = tbv(URl[m], TRl[m], ULr[m], TLr[m], ula[m], ura[m]);
        const double tt = tbv(URl[m], TRl[m], ULr[m], TLr[m], thl[m], thr[m]);
        if (mixedFlag && tt < tm) {
            ula[m] = thl[m];
            ura[m] = thr[m];
        }
    }
}

}  // namespace tj
