// Fluid (stiffened-gas / constant-Gamma Mie-Gruneisen) reconstruction, EOS
// and HLLC flux, and the directional flux kernels.
#pragma once
#include "tj_device.h"

namespace tj {

// ------------------------------------------------------------------ states
// R: the real type of reconstruction and flux work -- double in the FP64
// form, float in the mixed-precision form.
template <int NM, typename R = double>
struct Prim {  // primitive variables reconstructed by MUSCL (accesspattern.cpp:44-79)
    R a[NM], r[NM], u[3], p;
};

template <int NM, typename R = double>
struct FSt {  // face state after primitiveToConservative + getSoundSpeed
    R al[NM], ar[NM], rho, u[3], E, p, a;
};

// α-reconstruction of one cell for THINC/BVD (thinc.cpp:67-186)
template <int NM, typename R = double>
struct ARec {
    R ula[NM], ura[NM];  // MUSCL face alphas
    R thl[NM], thr[NM];  // THINC face alphas (valid where mixed)
    int mixed;           // bit m: mixedCell for material m
};

// ------------------------------------------------------------------ loads
__device__ __forceinline__ void map3(const Level& lv, int& i, int& j, int& k, bool f[3]) {
    i = bcmap(i, lv.n[0], lv.bc_lo[0], lv.bc_hi[0], f[0]);
    if (lv.dim > 1) j = bcmap(j, lv.n[1], lv.bc_lo[1], lv.bc_hi[1], f[1]); else f[1] = false;
    if (lv.dim > 2) k = bcmap(k, lv.n[2], lv.bc_lo[2], lv.bc_hi[2], f[2]); else f[2] = false;
}

template <int NM, typename R>
__device__ __forceinline__ void load_prim(const Level& lv, const MF& U, int i, int j, int k, Prim<NM, R>& P) {
    bool f[3];
    map3(lv, i, j, k, f);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        P.a[m] = (R)U.at(lv.cs[ALPHA] + m, i, j, k);
        P.r[m] = (R)U.at(lv.cs[RHO_K] + m, i, j, k);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        R v = (R)U.at(lv.cs[VELOCITY] + r, i, j, k);
        P.u[r] = f[r] ? -v : v;
    }
    P.p = (R)U.at(lv.cs[PRESSURE], i, j, k);
}

template <typename R = double>
__device__ __forceinline__ R load_alpha(const Level& lv, const MF& U, int m, int i, int j, int k) {
    bool f[3];
    map3(lv, i, j, k, f);
    return (R)U.at(lv.cs[ALPHA] + m, i, j, k);
}

// ------------------------------------------------------------------ MUSCL
// hllc.cpp:527-575.  min(2/(1+r), 2r/(1+r)) is evaluated with one division:
// both quotients share the divisor, and 2r < 2 exactly when r < 1, so the
// smaller quotient is the one with the smaller numerator (ties are equal).
template <typename R>
__device__ __forceinline__ R vanleer(R r) {
    if (r <= R(0)) return R(0);
    R num = (r < R(1)) ? R(2) * r : R(2);
    return tdiv(num, R(1) + r);
}
template <typename R>
__device__ __forceinline__ void muscl1(R um, R uc, R up, R& L, R& Rt) {
    R r = tdiv(uc - um, up - uc);
    if (tisinf(r) || tisnan(r)) r = R(0);
    R g = vanleer(r) * R(0.5) * (up - um);
    L = uc - R(0.5) * g;
    Rt = uc + R(0.5) * g;
}
template <int NM, typename Rt>
__device__ __forceinline__ void muscl(const Prim<NM, Rt>& m_, const Prim<NM, Rt>& c, const Prim<NM, Rt>& p_,
                                      Prim<NM, Rt>& L, Prim<NM, Rt>& R) {
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
template <int NM, typename R>
__device__ __forceinline__ void eos_sums(const Level& lv, const R* a, const R* ar, R& G, R& Ent, R& Pnt) {
    G = R(0);
    Ent = R(0);
    Pnt = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) G += tdiv(a[m], mat<R>(lv, m).Gg);
#pragma unroll
    for (int m = 0; m < NM; ++m) Ent += (mat<R>(lv, m).eref * ar[m] + R(0) * a[m]);
#pragma unroll
    for (int m = 0; m < NM; ++m) Pnt += tdiv(mat<R>(lv, m).pref * a[m], mat<R>(lv, m).Gg) + R(0) * a[m];
}

// primitiveToConservative + getSoundSpeed of one reconstructed state
// (boxaccesscellarray.cpp:124-174, 254-285).
template <int NM, typename R>
__device__ __forceinline__ void p2c_sound(const Level& lv, const Prim<NM, R>& P, FSt<NM, R>& S) {
    R rho = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        S.al[m] = P.a[m];
        S.ar[m] = P.a[m] * P.r[m];
        rho += S.ar[m];
    }
    R ke = R(0);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        S.u[q] = P.u[q];
        ke += R(0.5) * rho * P.u[q] * P.u[q];
    }
    R G, Ent, Pnt;
    eos_sums<NM, R>(lv, S.al, S.ar, G, Ent, Pnt);
    S.E = P.p * G + Ent - Pnt + ke;
    S.p = P.p;
    S.rho = rho;
    R a = R(0), xt = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatT<R>& M = mat<R>(lv, m);
        R c2 = tdiv(P.p * M.gp1, P.r[m]) - tdiv(M.pref, P.r[m]);
        a += tmax(R(0), tdiv(M.xi * c2 * S.ar[m], rho));
        xt += M.xi * S.al[m];
    }
    if (a <= R(0)) a = R(1E-6);
    S.a = tsqrt(tdiv(a, xt));
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
    eos_sums<NM, double>(lv, al, ar, G, Ent, Pnt);
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
template <int NM, typename R>
__device__ __forceinline__ void hllc(const FSt<NM, R>& L, const FSt<NM, R>& Rs, int d, R* F, R* sa) {
    const FSt<NM, R>& Rr = Rs;
    const R sigL = -L.p * R(1), sigR = -Rr.p * R(1);  // sigma_dd = -p delta_dd
    R ruL[3], ruR[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        ruL[q] = L.u[q] * L.rho;
        ruR[q] = Rr.u[q] * Rr.rho;
    }
    const R SR = tmax(tabs(L.u[d]) + L.a, tabs(Rr.u[d]) + Rr.a);
    const R SL = -SR;
    const R Ss = tdiv(-sigR + sigL + ruL[d] * (SL - L.u[d]) - ruR[d] * (SR - Rr.u[d]),
                      L.rho * (SL - L.u[d]) - Rr.rho * (SR - Rr.u[d]));
    F[2 * NM + 4] = Ss;
    // branch selection of calc_fluxes; K is the state the star state is built from
    const bool useL = (SL >= R(0)) || (Ss >= R(0));
    const bool addStar = !(SL >= R(0)) && ((Ss >= R(0)) || (SR >= R(0)));
    const FSt<NM, R>& K = useL ? L : Rr;
    const R* ruK = useL ? ruL : ruR;
    const R SK = useL ? SL : SR;
    const R sigK = useL ? sigL : sigR;
    // star state
    const R mult = (tabs(tdiv(Ss - K.u[d], Ss)) < R(1E-20)) ? R(1) : tdiv(SK - K.u[d], SK - Ss);
    R arS[NM], rhoS = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        arS[m] = mult * K.ar[m];
        rhoS += arS[m];
    }
    R ruS[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) ruS[q] = (q == d) ? mult * K.rho * Ss : mult * ruK[q];
    const R uSd = tdiv(ruS[d], rhoS);
    const R ES = mult * K.rho * (tdiv(K.E, K.rho) + (Ss - K.u[d]) * (Ss - tdiv(sigK, K.rho * (SK - K.u[d]))));
    // physical flux of K (flux.cpp:33-47) with sigma = -p I
    R vds = R(0);
#pragma unroll
    for (int q = 0; q < 3; ++q) vds += K.u[q] * (-K.p * (q == d ? R(1) : R(0)));
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        F[m] = K.al[m] * uSd;
        sa[m] = K.al[m];
        R f = K.ar[m] * K.u[d];
        F[NM + m] = addStar ? f + SK * (arS[m] - K.ar[m]) : f;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        R f = ruK[q] * K.u[d] - (-K.p * (q == d ? R(1) : R(0)));
        F[2 * NM + q] = addStar ? f + SK * (ruS[q] - ruK[q]) : f;
    }
    {
        R f = K.E * K.u[d] - vds;
        F[2 * NM + 3] = addStar ? f + SK * (ES - K.E) : f;
    }
}

template <int NM, typename R>
__device__ __forceinline__ void flip_state(FSt<NM, R>& S, int d) { S.u[d] = -S.u[d]; }

// ------------------------------------------------------------------ THINC
// thinc.cpp:54-88 (mixedCell, Youngs normal in 2D), 130-186 (pass 1).
// cell (i,j,k) along direction d; alphas are read from U with the boundary map.
template <int NM, typename R>
__device__ inline void arec(const Level& lv, const MF& U, int i, int j, int k, int d, R ynx, R yny, ARec<NM, R>& A) {
    int di = d == 0, dj = d == 1, dk = d == 2;
    A.mixed = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const R aL = load_alpha<R>(lv, U, m, i - di, j - dj, k - dk);
        const R aC = load_alpha<R>(lv, U, m, i, j, k);
        const R aR = load_alpha<R>(lv, U, m, i + di, j + dj, k + dk);
        muscl1(aL, aC, aR, A.ula[m], A.ura[m]);
        const R mn = tmin(aL, aR);
        const R mx = tmax(aL, aR) - mn;
        const R theta = tsgn(aR - aL);
        const R eps = R(1E-20);
        const R C = (mx < eps) ? R(0) : tdiv(aC - mn + eps, mx + eps);
        const R dl = R(1E-4);
        const bool mixed = (C > dl) && (C < (R(0.999999) - dl)) && ((aR - aC) * (aC - aL) >= R(0));
        A.thl[m] = R(0);
        A.thr[m] = R(0);
        if (mixed) {
            A.mixed |= 1 << m;
            R nvc = R(1);
            if (lv.dim == 2) {
                auto a = [&](int ddi, int ddj) { return load_alpha<R>(lv, U, m, i + ddi, j + ddj, k); };
                const R nx = ynx * (a(1, 1) + R(2) * a(1, 0) + a(1, -1) - a(-1, 1) - R(2) * a(-1, 0) - a(-1, -1));
                const R ny = yny * (a(1, 1) + R(2) * a(0, 1) + a(-1, 1) - a(-1, -1) - R(2) * a(0, -1) - a(1, -1));
                const R norm = tsqrt(nx * nx + ny * ny);
                if (norm < R(1E-20)) nvc = R(0);
                else nvc = (d == 0) ? tabs(tdiv(nx, norm)) : tabs(tdiv(ny, norm));
            }
            const R beta = nvc * R(lv.beta0) + R(0.01);
            const R coshB = tcosh(beta), tanhB = ttanh(beta);
            const R B = texp(theta * beta * ((R(2) * C) - R(1)));
            const R Ac = tdiv(tdiv(B, coshB) - R(1), tanhB);
            A.thr[m] = mn + tdiv(mx, R(2)) * (R(1) + theta * tdiv(tanhB + Ac, R(1) + Ac * tanhB));
            A.thl[m] = mn + tdiv(mx, R(2)) * (R(1) + theta * Ac);
        }
    }
}

// TBV of thinc.cpp:57-63 for one material.
template <typename R>
__device__ __forceinline__ R tbv(R URl, R TRl, R ULr, R TLr, R PL, R PR) {
    return tmin(tmin(tabs(URl - PL) + tabs(PR - ULr), tabs(TRl - PL) + tabs(PR - TLr)),
                tmin(tabs(URl - PL) + tabs(PR - TLr), tabs(TRl - PL) + tabs(PR - ULr)));
}

// cleanUpAlpha of one cell (boxaccesscellarray.cpp:530-599), the Nan-index
// quirk of :553-556 included.  Returns TJ status bits.
template <int NM, typename R>
__device__ __forceinline__ unsigned cleanup_alpha(R* a) {
    R tot = R(0);
    int nan = 0;
    unsigned st = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        if (tisnan(a[m])) {
            a[m] = R(0);
            if (nan == 0) nan = m;
            else st |= 4u;
        } else if (a[m] < R(0)) {
            a[m] = R(1E-6);
        } else if (a[m] > R(1)) {
            a[m] = R(1);
        }
        tot += a[m];
    }
    if (nan > 0) {
#pragma unroll
        for (int m = 0; m < NM; ++m)
            if (m == nan) {
                a[m] = R(1) - tot;
                tot += a[m];
            }
    }
    if (tot <= R(0)) st |= 8u;
#pragma unroll
    for (int m = 0; m < NM; ++m) a[m] *= tdiv(R(1), tot);
    return st;
}

// BVD selection for one cell (thinc.cpp:191-238) given the neighbour data.
//   nb*: neighbour values with the reference's box-edge semantics already applied
//   own thl/thr: THINC values of this cell where mixed, else the stale history
template <int NM, typename R>
__device__ __forceinline__ void bvd(int mixedFlag, const R* URl, const R* TRl, const R* ULr, const R* TLr, const R* thl,
                                    const R* thr, R* ula, R* ura) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const R tm = tbv(URl[m], TRl[m], ULr[m], TLr[m], ula[m], ura[m]);
        const R tt = tbv(URl[m], TRl[m], ULr[m], TLr[m], thl[m], thr[m]);
        if (mixedFlag && tt < tm) {
            ula[m] = thl[m];
            ura[m] = thr[m];
        }
    }
}

}  // namespace tj
