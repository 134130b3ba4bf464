// Fluid (stiffened-gas / constant-Gamma Mie-Gruneisen) reconstruction, EOS
// and HLLC flux, and the directional flux kernels.
#pragma once
#include "tj_device.h"

namespace tj {
inline namespace TJ_VIEW_NS {

// ------------------------------------------------------------------ states
// R: the real type of reconstruction and flux work -- double in the FP64
// form, float in the mixed-precision form.
// MX: the level has a mixture material (MixtureEOS, lv.mixm), whose mass
// fraction lambda is a primitive and alpha rho lambda a conserved variable.
template <int NM, typename R = double>
struct Prim {  // primitive variables reconstructed by MUSCL (accesspattern.cpp:44-79)
    R a[NM], r[NM], u[3], p;
    R lam;  // mixture only
};

template <int NM, typename R = double>
struct FSt {  // face state after primitiveToConservative + getSoundSpeed
    R al[NM], ar[NM], rho, u[3], E, p, a;
    R arl;  // mixture only
};

// The mixture material's variables a closure needs: rho_k, lambda and the
// phase densities RHO_MIX (reactants a, products b).
template <typename R>
struct MixV {
    R rk, lam, ra, rb;
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

template <int NM, typename R, bool MX = false>
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
    if (MX) P.lam = (R)U.at(lv.cs[LAMBDA], i, j, k);
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
template <int NM, typename Rt, bool MX = false>
__device__ __forceinline__ void muscl(const Prim<NM, Rt>& m_, const Prim<NM, Rt>& c, const Prim<NM, Rt>& p_,
                                      Prim<NM, Rt>& L, Prim<NM, Rt>& R) {
#pragma unroll
    for (int m = 0; m < NM; ++m) muscl1(m_.a[m], c.a[m], p_.a[m], L.a[m], R.a[m]);
#pragma unroll
    for (int m = 0; m < NM; ++m) muscl1(m_.r[m], c.r[m], p_.r[m], L.r[m], R.r[m]);
#pragma unroll
    for (int q = 0; q < 3; ++q) muscl1(m_.u[q], c.u[q], p_.u[q], L.u[q], R.u[q]);
    muscl1(m_.p, c.p, p_.p, L.p, R.p);
    if (MX) muscl1(m_.lam, c.lam, p_.lam, L.lam, R.lam);
}

// ------------------------------------------------------------------ EOS (kind 0)
// MixtureEOS pieces (equationofstate.cpp:158-236, 239-247, 364-370): the
// Gruneisen coefficient from the CV-weighted adiabatic index, the products'
// density from the volume balance, and the phase densities of a primitive
// state in thermal equilibrium.
template <typename R>
__device__ __forceinline__ R mix_gamma(const MatT<R>& M, R lam) {
    return tdiv(lam * M.ga * M.CV + (R(1) - lam) * M.g2 * M.CV2, lam * M.CV + (R(1) - lam) * M.CV2) - R(1);
}
template <typename R>
__device__ __forceinline__ R mix_rhob(R lam, R rk, R ra) {
    return tdiv(R(1) - lam, tdiv(R(1), rk) - tdiv(lam, ra));
}
template <typename R>
__device__ __forceinline__ void mix_define(const MatT<R>& M, R p, MixV<R>& x) {
    x.ra = x.rk * (x.lam + (R(1) - x.lam) * tdiv(p - M.pref, p - M.pref2) * tdiv(M.CV2 * M.Gg2, M.CV * M.Gg));
    x.rb = mix_rhob(x.lam, x.rk, x.ra);
}
// mixtureSoundSpeed (equationofstate.cpp:212-226)
template <typename R>
__device__ __forceinline__ R mix_c2(const MatT<R>& M, R p, const MixV<R>& x) {
    R dedrho = (x.lam) * tdiv(M.pref - p, M.Gg * x.rk * x.rk);
    dedrho += (R(1) - x.lam) * tdiv(M.pref2 - p, M.Gg2 * x.rk * x.rk);
    return tdiv(tdiv(p, x.rk * x.rk) - dedrho, tdiv(x.lam, x.ra * M.Gg) + tdiv(R(1) - x.lam, x.rb * M.Gg2));
}

// equationofstate.cpp:37-64 and boxaccesscellarray.cpp:218-252; the mixture
// material (MX) by equationofstate.cpp:158-175 (its shear terms are 0.0).
template <int NM, typename R, bool MX = false>
__device__ __forceinline__ void eos_sums(const Level& lv, const R* a, const R* ar, R& G, R& Ent, R& Pnt,
                                         const MixV<R>& x = MixV<R>{}) {
    G = R(0);
    Ent = R(0);
    Pnt = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatT<R>& M = mat<R>(lv, m);
        if (MX && m == lv.mixm) G += tdiv(a[m], mix_gamma(M, x.lam));
        else G += tdiv(a[m], M.Gg);
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatT<R>& M = mat<R>(lv, m);
        if (MX && m == lv.mixm) Ent += (x.lam * (M.eref * ar[m]) + (R(1) - x.lam) * (M.eref2 * ar[m]) + R(0));
        else Ent += (M.eref * ar[m] + R(0) * a[m]);
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatT<R>& M = mat<R>(lv, m);
        if (MX && m == lv.mixm) {
            const R pa = tdiv(M.pref * a[m], M.Gg), pb = tdiv(M.pref2 * a[m], M.Gg2);
            Pnt += tdiv(pa * x.lam * x.rk, x.ra) + tdiv(pb * (R(1) - x.lam) * x.rk, x.rb) + R(0);
        } else {
            Pnt += tdiv(M.pref * a[m], M.Gg) + R(0) * a[m];
        }
    }
}

// primitiveToConservative + getSoundSpeed of one reconstructed state
// (boxaccesscellarray.cpp:124-174, 254-285).
template <int NM, typename R, bool MX = false>
__device__ __forceinline__ void p2c_sound(const Level& lv, const Prim<NM, R>& P, FSt<NM, R>& S, MixV<R>* xo = nullptr) {
    R rho = R(0);
    MixV<R> x{};
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        S.al[m] = P.a[m];
        S.ar[m] = P.a[m] * P.r[m];
        rho += S.ar[m];
        if (MX && m == lv.mixm) {
            S.arl = P.lam * P.a[m] * P.r[m];
            x.rk = P.r[m];
            x.lam = P.lam;
            mix_define(mat<R>(lv, m), P.p, x);
        }
    }
    R ke = R(0);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        S.u[q] = P.u[q];
        ke += R(0.5) * rho * P.u[q] * P.u[q];
    }
    R G, Ent, Pnt;
    eos_sums<NM, R, MX>(lv, S.al, S.ar, G, Ent, Pnt, x);
    S.E = P.p * G + Ent - Pnt + ke;
    S.p = P.p;
    S.rho = rho;
    R a = R(0), xt = R(0);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatT<R>& M = mat<R>(lv, m);
        const bool mx = MX && m == lv.mixm;
        const R c2 = mx ? mix_c2(M, P.p, x) : tdiv(P.p * M.gp1, P.r[m]) - tdiv(M.pref, P.r[m]);
        const R xi = mx ? tdiv(R(1), mix_gamma(M, x.lam)) : M.xi;
        a += tmax(R(0), tdiv(xi * c2 * S.ar[m], rho));
        xt += xi * S.al[m];
    }
    if (a <= R(0)) a = R(1E-6);
    S.a = tsqrt(tdiv(a, xt));
    if (MX && xo) *xo = x;
}

// MixtureEOS::rootFind (equationofstate.cpp:275-354): the phase densities
// of the mixture material of a cell in thermal equilibrium, by bisection on
// the reactants' density; bisectionFunction (:264-273) evaluates the cell's
// pressure at each trial.  Every term of the effective sums that does not
// depend on the trial density is formed once, with the reference's operands,
// and the sums are taken in the reference's order, so each trial rounds
// exactly as a full re-evaluation would.  The sign of f at the lower end is
// carried over instead of re-evaluated each pass (the reference evaluates it
// again; f has no state, so the result is the same).  The reference's loop
// has no bound; this one stops after 2000 passes (a finite bracket converges
// in under 100).  Returns TJ status bits (32: no sign change, where the
// reference exits).
template <int NM>
__device__ __forceinline__ unsigned root_find(const Level& lv, const double* al, const double* ar, double E, double ke, MixV<double>& x) {
    const MatC& M = lv.m[lv.mixm];
    const double lam = x.lam;
    // trial-independent parts of getEffectiveInverseGruneisen /
    // NonThermalInternalEnergy / NonThermalPressure (boxaccesscellarray.cpp:218-252)
    double G = 0.0, Ent = 0.0, pconst[NM], pal = 0.0, pbl = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatC& Mm = lv.m[m];
        if (m == lv.mixm) G += al[m] / (mix_gamma(Mm, lam));
        else G += al[m] / Mm.Gg;
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatC& Mm = lv.m[m];
        if (m == lv.mixm) Ent += (lam * (Mm.eref * ar[m]) + (1.0 - lam) * (Mm.eref2 * ar[m]) + 0.0);
        else Ent += (Mm.eref * ar[m] + 0.0 * al[m]);
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatC& Mm = lv.m[m];
        if (m == lv.mixm) {
            pal = (Mm.pref * al[m] / Mm.Gg) * lam * x.rk;
            pbl = (Mm.pref2 * al[m] / Mm.Gg2) * (1.0 - lam) * x.rk;
            pconst[m] = 0.0;
        } else {
            pconst[m] = Mm.pref * al[m] / Mm.Gg + 0.0 * al[m];
        }
    }
    const double cE = E - ke - Ent;
    const double inv_rk = 1.0 / x.rk;
    auto rhob = [&](double ra) { return (1.0 - lam) / (inv_rk - lam / ra); };
    auto fn = [&](double ra) {
        const double rb = rhob(ra);
        double Pnt = 0.0;
#pragma unroll
        for (int m = 0; m < NM; ++m) Pnt += (m == lv.mixm) ? (pal / (ra) + pbl / (rb) + 0.0) : pconst[m];
        const double pp = (cE + Pnt) / (G);
        return ((pp - M.pref) / (ra * M.Gg * M.CV)) - ((pp - M.pref2) / (rb * M.Gg2 * M.CV2));
    };
    if (M.pref == 0.0 && M.pref2 == 0.0) {
        x.ra = x.rk * (lam + (1.0 - lam) * (M.Gg2 * M.CV2) / (M.Gg * M.CV));
        x.rb = mix_rhob(lam, x.rk, x.ra);
        return 0u;
    }
    // lambda at either end: both phases at the mixture's density.  One branch
    // with one store order: two branches storing ra and rb in opposite orders
    // were merged into a store through a select of the two fields' addresses,
    // which kept x in scratch memory in every mixture kernel.
    if (lam > 1.0 - 1E-3 || lam < 1E-3) {
        x.ra = x.rk;
        x.rb = x.rk;
        return 0u;
    }
    double a = lam * x.rk + 1E-10;
    double b = 1000000.0;
    double mid = 0.5 * (a + b);
    const double fa = fn(a);
    if (__builtin_fabs(fa) <= 1E-4) {
        x.ra = a;
        x.rb = rhob(a);
        return 0u;
    }
    const double fb = fn(b);
    if (__builtin_fabs(fb) <= 1E-4) {
        x.ra = b;
        x.rb = rhob(b);
        return 0u;
    }
    double sa = sgnd(fa);
    if (sa == sgnd(fb)) {  // the trial densities of b stay, as in the oracle
        x.ra = b;
        x.rb = rhob(b);
        return 32u;
    }
    for (int n = 0;; n++) {
        const double smid = sgnd(fn(mid));
        if (smid == sa) a = mid;  // f(a) keeps its sign: a moves only to a point of that sign
        else b = mid;
        mid = 0.5 * (a + b);
        if (__builtin_fabs(a - b) / (a + b) < 1E-7 || n >= 2000) {
            x.ra = mid;
            x.rb = rhob(mid);
            return n >= 2000 ? 64u : 0u;
        }
    }
}

// conservativeToPrimitive (boxaccesscellarray.cpp:67-122), fluid part; with
// a mixture (MX) lambda = alpha rho lambda / alpha rho and rootFind come
// after the kinetic energy (:104-112).  Returns TJ status bits.
template <int NM, bool MX = false>
__device__ __forceinline__ unsigned c2p(const Level& lv, const double* al, const double* ar, const double* ru, double E,
                                        double* rk, double& rho, double* u, double& p, double arl = 0.0,
                                        MixV<double>* x = nullptr) {
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
    unsigned st = 0;
    MixV<double> xm{};
    if (MX) {
#pragma unroll
        for (int m = 0; m < NM; ++m)
            if (m == lv.mixm) {
                xm.rk = rk[m];
                xm.lam = arl / ar[m];
            }
        st = root_find<NM>(lv, al, ar, E, ke, xm);
        *x = xm;
    }
    double G, Ent, Pnt;
    eos_sums<NM, double, MX>(lv, al, ar, G, Ent, Pnt, xm);
    p = (E - ke - Ent + Pnt) / G;
    return st;
}

// getSoundSpeed of a cell state (boxaccesscellarray.cpp:254-285); a mixture
// reads its stored lambda and phase densities.
template <int NM, bool MX = false>
__device__ __forceinline__ double sound(const Level& lv, const double* al, const double* ar, const double* rk, double rho, double p,
                                        const MixV<double>& x = MixV<double>{}) {
    double a = 0.0, xt = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatC& M = lv.m[m];
        const bool mx = MX && m == lv.mixm;
        const double c2 = mx ? mix_c2(M, p, x) : p * M.gp1 / rk[m] - M.pref / rk[m];
        const double xi = mx ? 1.0 / (mix_gamma(M, x.lam)) : M.xi;
        a += dmax(0.0, xi * c2 * ar[m] / rho);
        xt += xi * al[m];
    }
    if (a <= 0.0) a = 1E-6;
    return __builtin_sqrt(a / xt);
}

// ------------------------------------------------------------------ HLLC
// hllc.cpp:5-73 (getSstar, getStarState), 384-487 (calc_fluxes), flux.cpp:5-47.
// F layout: [0,NM) alpha, [NM,2NM) alphaRho, [2NM,2NM+3) rhoU, 2NM+3 E, 2NM+4 S*.
// MX: F[2NM+5] is the alpha rho lambda flux (flux.cpp:39, hllc.cpp:26-29).
template <int NM, typename R, bool MX = false>
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
    const R mult = (tabs(tdiv(Ss - K.u[d], Ss)) < R(1E-20)) ? R(1) : tdiv_cr(SK - K.u[d], SK - Ss);
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
    if (MX) {
        const R arlS = mult * K.arl;
        const R f = K.arl * K.u[d];
        F[2 * NM + 5] = addStar ? f + SK * (arlS - K.arl) : f;
    }
}

template <int NM, typename R>
__device__ __forceinline__ void flip_state(FSt<NM, R>& S, int d) { S.u[d] = -S.u[d]; }

// ------------------------------------------------------------------ THINC
// thinc.cpp:54-88 (mixedCell, Youngs normal in 2D), 130-186 (pass 1).
// cell (i,j,k) along direction d; alphas are read from U with the boundary map
// (V: the state view, MF, or the tangent-linear form's value + derivative view).
template <int NM, typename R, typename V = MF>
__device__ inline void arec(const Level& lv, const V& U, int i, int j, int k, int d, R ynx, R yny, ARec<NM, R>& A) {
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

}  // inline namespace TJ_VIEW_NS
}  // namespace tj
