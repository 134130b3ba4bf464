// Tangent-linear (forward-mode differentiated) form: the FP64 step of the
// solid kernels, and of the MUSCL(+THINC/BVD) + HLLC fluid kernels, on Dual<K>
// numbers, carrying K directional derivatives of the whole state with respect
// to per-material EOS / plasticity parameters (Gamma0, s, yield stress; for
// fluids the Gruneisen coefficient adiabaticIndex - 1 and pref).  Each kernel mirrors
// its FP64 counterpart (k_init_p2c_s, k_sound_dt_s, k_flux_x_s /
// k_flux_lmarch_s, k_update_s, k_clock_dt, k_clock_advance) operation for
// operation, so the value parts equal the FP64 trajectory bit for bit.
//
// Layout: a state's values are the caller's CellArray; its derivative parts
// are K further blocks of the same components, compact (i fastest).  Face
// planes hold the NF flux components of the values, then NF per direction.
// The time step is differentiated through dt = CFL min(dx / (a + |u|)) along
// the cell that attains the minimum (the lowest cell index on ties).
#pragma once
#include "tj_dual.h"
#include "tj_solid_kernels.h"

namespace tj {

template <int K>
struct DMF {
    MF u;           // values
    MF dd;          // derivative block 0 (compact strides)
    long long bs;   // stride between derivative blocks
    __device__ __forceinline__ Dual<K> get(int c, int i, int j, int k) const {
        Dual<K> x;
        x.v = u.at(c, i, j, k);
        const long long o = (long long)c * dd.ns + (long long)k * dd.ks + (long long)j * dd.js + i;
#pragma unroll
        for (int t = 0; t < K; ++t) x.d[t] = dd.p[t * bs + o];
        return x;
    }
    // NT: the cache hints of the FP64 update kernels (get_base / put_state):
    // the update's input and base reads and its stage-2 store (512^2 plate
    // gradient: 937 -> 916 ms)
    template <bool NT>
    __device__ __forceinline__ Dual<K> getx(int c, int i, int j, int k) const {
        if constexpr (!NT) return get(c, i, j, k);
        Dual<K> x;
        x.v = get_base(u.at(c, i, j, k));
        const long long o = (long long)c * dd.ns + (long long)k * dd.ks + (long long)j * dd.js + i;
#pragma unroll
        for (int t = 0; t < K; ++t) x.d[t] = get_base(dd.p[t * bs + o]);
        return x;
    }
    template <bool NT>
    __device__ __forceinline__ void putx(int c, int i, int j, int k, const Dual<K>& x) const {
        if constexpr (!NT) { put(c, i, j, k, x); return; }
        put_state<true>(u.at(c, i, j, k), x.v);
        const long long o = (long long)c * dd.ns + (long long)k * dd.ks + (long long)j * dd.js + i;
#pragma unroll
        for (int t = 0; t < K; ++t) put_state<true>(dd.p[t * bs + o], x.d[t]);
    }
    __device__ __forceinline__ void put(int c, int i, int j, int k, const Dual<K>& x) const {
        u.at(c, i, j, k) = x.v;
        const long long o = (long long)c * dd.ns + (long long)k * dd.ks + (long long)j * dd.js + i;
#pragma unroll
        for (int t = 0; t < K; ++t) dd.p[t * bs + o] = x.d[t];
    }
};

template <int NM, int K>
__device__ __forceinline__ void load_tprim(const Level& lv, const DMF<K>& U, int i, int j, int k,
                                           SPrim<NM, Dual<K>>& P) {
    bool f[3];
    map3(lv, i, j, k, f);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        P.a[m] = U.get(lv.cs[ALPHA] + m, i, j, k);
        P.r[m] = U.get(lv.cs[RHO_K] + m, i, j, k);
        P.e[m] = lv.plastic ? U.get(lv.cs[EPSILON] + m, i, j, k) : Dual<K>(0.0);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const Dual<K> v = U.get(lv.cs[VELOCITY] + r, i, j, k);
        P.u[r] = f[r] ? -v : v;
    }
    P.p = U.get(lv.cs[PRESSURE], i, j, k);
#pragma unroll
    for (int q = 0; q < 9; ++q) P.V[q] = U.get(lv.cs[V_TENSOR] + q, i, j, k);
}

// Faces along D with one reconstruction per cell.  Each wave covers a tile
// of cells: 64 along x (D = 0), or 8 along x by 8 along D (D = 1, 2; lanes
// x fastest, so a load reads 64-byte row pieces).  A lane reconstructs one
// cell (MUSCL, then p2c + sound speed of both face states), and face f takes
// its left state UR(f-1) from the lane one cell below along D by shuffle, so
// a tile column of L cells gives L - 1 faces (L = 64 or 8).  At the domain
// faces the left state of face 0 is the image of UR(0) and the right state
// of face n the image of UL(n-1) (normal velocity flipped when reflective),
// as in the FP64 passes.  The right state is shuffled before the left one is
// built, so two face states are live at a time.
template <int NM, int K, int D>
__global__ __launch_bounds__(256) void k_tflux_ts(Level lv, DMF<K> U, Faces F, Clock* clk) {
    if (clk->done) return;
    constexpr int LN = D == 0 ? 64 : 8;  // lanes along D
    constexpr int LT = 64 / LN;           // lanes along x (D = 1, 2)
    constexpr int PER = LN - 1;           // faces per tile column
    constexpr int NF = 3 * NM + 17;
    using T = Dual<K>;
    const int lane = threadIdx.x & 63;
    const int ld = lane / LT, lt = lane % LT;
    const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + wave_in_block();
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2], nd = lv.n[D];
    const int nseg = (nd + 1 + PER - 1) / PER;
    const int nxt = D == 0 ? 1 : (nx + LT - 1) / LT;
    const int no = D == 0 ? ny * nz : (D == 1 ? nz : ny);
    if (wave >= (long long)nxt * nseg * no) return;  // uniform per wave
    const int xt = (int)(wave % nxt);
    const long long rest = wave / nxt;
    const int seg = (int)(rest % nseg);
    const int o = (int)(rest / nseg);
    const int c = seg * PER - 1 + ld;  // the lane's cell along D
    bool fl;
    const int s = bcmap(c, nd, lv.bc_lo[D], lv.bc_hi[D], fl);
    const int i0 = xt * LT + lt;
    const bool act = D == 0 || i0 < nx;
    const int i = D == 0 ? s : (act ? i0 : nx - 1);
    const int j = D == 0 ? o % ny : (D == 1 ? s : o);
    const int k = D == 0 ? o / ny : (D == 1 ? o : s);
    auto at = [&](int r, SPrim<NM, T>& P) {
        load_tprim<NM, K>(lv, U, D == 0 ? r : i, D == 1 ? r : j, D == 2 ? r : k, P);
    };
    SPrim<NM, T> PL;
    SSt<NM, T> Lf;
    {
        SPrim<NM, T> Pm, Pc, Pp, PR;
        at(s - 1, Pm);
        at(s, Pc);
        at(s + 1, Pp);
        muscl_s<NM, T>(Pm, Pc, Pp, PL, PR);
        SSt<NM, T> SRs;
        p2c_sound_s<NM, false, T>(lv, PR, D, SRs);
        Lf = shfl_state<NM, T>(SRs, (c == 0 || ld == 0) ? lane : lane - LT);
    }
    if (c == 0 && lv.bc_lo[D]) flip_s<NM, T>(Lf, D);
    SSt<NM, T> Rf;
    p2c_sound_s<NM, false, T>(lv, PL, D, Rf);
    if (c == nd && lv.bc_hi[D]) flip_s<NM, T>(Rf, D);
    unsigned st = 0;
    if (act && ld >= 1 && c >= 0 && c <= nd) {
        T Fv[NF];
        wave5<NM, false, T>(lv, Lf, Rf, D, Fv, st);
        const int fi = D == 0 ? c : i, fj = D == 1 ? c : j, fk = D == 2 ? c : k;
        const long long id = ((long long)fk * F.ny + fj) * F.nx + fi;
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            F.p[q * F.cs + id] = Fv[q].v;
#pragma unroll
            for (int t = 0; t < K; ++t) F.p[(NF * (t + 1) + q) * F.cs + id] = Fv[q].d[t];
        }
    }
    if (st) atomicOr(&clk->status, st);
}

template <int NM, int K>
struct TCell {
    Dual<K> al[NM], ar[NM], rk[NM], e[NM], are[NM], rho, ru[3], u[3], E, p, V[9], hj2;
};
template <int NM, int K, bool NT = false>
__device__ __forceinline__ void load_tcell(const Level& lv, const DMF<K>& U, int i, int j, int k, TCell<NM, K>& C) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        C.al[m] = U.template getx<NT>(lv.cs[ALPHA] + m, i, j, k);
        C.ar[m] = U.template getx<NT>(lv.cs[ALPHARHO] + m, i, j, k);
        C.rk[m] = U.template getx<NT>(lv.cs[RHO_K] + m, i, j, k);
        C.e[m] = lv.plastic ? U.template getx<NT>(lv.cs[EPSILON] + m, i, j, k) : Dual<K>(0.0);
        C.are[m] = lv.plastic ? U.template getx<NT>(lv.cs[ALPHARHOEPSILON] + m, i, j, k) : Dual<K>(0.0);
    }
    C.rho = U.template getx<NT>(lv.cs[RHO], i, j, k);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        C.ru[q] = U.template getx<NT>(lv.cs[RHOU] + q, i, j, k);
        C.u[q] = U.template getx<NT>(lv.cs[VELOCITY] + q, i, j, k);
    }
    C.E = U.template getx<NT>(lv.cs[TOTAL_E], i, j, k);
    C.p = U.template getx<NT>(lv.cs[PRESSURE], i, j, k);
#pragma unroll
    for (int q = 0; q < 9; ++q) C.V[q] = U.template getx<NT>(lv.cs[V_TENSOR] + q, i, j, k);
    C.hj2 = U.template getx<NT>(lv.cs[HJ2], i, j, k);
}
template <int NM, int K, bool NT = false>
__device__ __forceinline__ void store_tcell(const Level& lv, const DMF<K>& U, int i, int j, int k, const TCell<NM, K>& C) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        U.template putx<NT>(lv.cs[ALPHA] + m, i, j, k, C.al[m]);
        U.template putx<NT>(lv.cs[ALPHARHO] + m, i, j, k, C.ar[m]);
        U.template putx<NT>(lv.cs[RHO_K] + m, i, j, k, C.rk[m]);
        if (lv.plastic) {
            U.template putx<NT>(lv.cs[EPSILON] + m, i, j, k, C.e[m]);
            U.template putx<NT>(lv.cs[ALPHARHOEPSILON] + m, i, j, k, C.are[m]);
        }
    }
    U.template putx<NT>(lv.cs[RHO], i, j, k, C.rho);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        U.template putx<NT>(lv.cs[RHOU] + q, i, j, k, C.ru[q]);
        U.template putx<NT>(lv.cs[VELOCITY] + q, i, j, k, C.u[q]);
    }
    U.template putx<NT>(lv.cs[TOTAL_E], i, j, k, C.E);
    U.template putx<NT>(lv.cs[PRESSURE], i, j, k, C.p);
#pragma unroll
    for (int q = 0; q < 9; ++q) U.template putx<NT>(lv.cs[V_TENSOR] + q, i, j, k, C.V[q]);
    U.template putx<NT>(lv.cs[HJ2], i, j, k, C.hj2);
}

// getSoundSpeed + the CFL candidate of one cell; the candidate's value and
// derivative parts go to cand (K + 1 planes of ncell).
template <int NM, int K>
__device__ __forceinline__ double tsound_cand(const Level& lv, const DMF<K>& U, const TCell<NM, K>& C, int i, int j,
                                              int k, long long id, double* cand) {
    using T = Dual<K>;
    const T a = sound_s<NM, false, T>(lv, C.al, C.ar, C.rk, C.rho, C.p, C.hj2);
    U.put(lv.cs[SOUNDSPEED], i, j, k, a);
    T c = T(__builtin_huge_val());
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (d < lv.dim) c = tmin(c, tdiv(lv.dx[d], a + tabs(C.u[d])));
    const long long nc = lv.ncells;
    cand[id] = c.v;
#pragma unroll
    for (int t = 0; t < K; ++t) cand[(t + 1) * nc + id] = c.d[t];
    return c.v;
}

// update + conservativeToPrimitive; stage 2 the RK average and the post-step
// chain (plastic + c2p, cleanUpV, cleanUpAlpha, NaN scan, CFL candidate).
// dt: value clk->dt, derivatives tdt[1..K].
template <int NM, int STAGE, int K>
__global__ __launch_bounds__(256) void k_tupdate_s(Level lv, DMF<K> Uin, DMF<K> Ub, DMF<K> Uo, Faces F0, Faces F1,
                                                   Faces F2, Clock* clk, const double* tdt, unsigned ops, double* cand) {
    if (clk->done) return;
    using T = Dual<K>;
    constexpr int NC = 3 * NM + 13, NF = NC + 4;
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncell = (long long)nx * ny * nz;
    double cv = __builtin_huge_val();
    unsigned st = 0;
    if (id < ncell) {
        const int i = (int)(id % nx);
        const int j = (int)((id / nx) % ny);
        const int k = (int)(id / ((long long)nx * ny));
        T dt;
        dt.v = clk->dt;
#pragma unroll
        for (int t = 0; t < K; ++t) dt.d[t] = tdt[1 + t];
        TCell<NM, K> C;
        load_tcell<NM, K, true>(lv, Uin, i, j, k, C);
        T a0[NM], V0[9], u0[3];
#pragma unroll
        for (int m = 0; m < NM; ++m) a0[m] = C.al[m];
#pragma unroll
        for (int q = 0; q < 9; ++q) V0[q] = C.V[q];
#pragma unroll
        for (int q = 0; q < 3; ++q) u0[q] = C.u[q];
        for (int d = 0; d < lv.dim; ++d) {
            const Faces& Fd = d == 0 ? F0 : (d == 1 ? F1 : F2);
            const long long lo = ((long long)k * Fd.ny + j) * Fd.nx + i;
            const long long hi = ((long long)(k + (d == 2)) * Fd.ny + (j + (d == 1))) * Fd.nx + (i + (d == 0));
            auto fv = [&](int q, long long o) {
                T x;
                x.v = Fd.p[q * Fd.cs + o];
#pragma unroll
                for (int t = 0; t < K; ++t) x.d[t] = Fd.p[(NF * (t + 1) + q) * Fd.cs + o];
                return x;
            };
            const T dtdx = tdiv(dt, lv.dx[d]);
            const T s0 = fv(NC, lo), s1 = fv(NC, hi);
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                C.al[m] += dtdx * (fv(m, lo) - fv(m, hi) - a0[m] * (s0 - s1));
                C.ar[m] += dtdx * (fv(NM + m, lo) - fv(NM + m, hi));
                if (lv.plastic) C.are[m] += dtdx * (fv(2 * NM + 13 + m, lo) - fv(2 * NM + 13 + m, hi));
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) C.ru[q] += dtdx * (fv(2 * NM + q, lo) - fv(2 * NM + q, hi));
            C.E += dtdx * (fv(2 * NM + 3, lo) - fv(2 * NM + 3, hi));
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int n = 2 * NM + 4 + r * 3 + c;
                    const T vs0 = fv(NC + 1 + c, lo), vs1 = fv(NC + 1 + c, hi);
                    C.V[r * 3 + c] += dtdx * (fv(n, lo) - fv(n, hi) - (2.0 / 3.0) * V0[r * 3 + c] * (s0 - s1) +
                                              u0[r] * (vs0 - vs1));
                }
        }
        T dH[9];
        C.hj2 = c2p_s<NM, false, T>(lv, C.al, C.ar, C.are, C.ru, C.E, C.V, C.rk, C.e, C.rho, C.u, C.p, dH);
        if (STAGE == 2) {
            TCell<NM, K> B;
            load_tcell<NM, K, true>(lv, Ub, i, j, k, B);
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                C.al[m] = B.al[m] * (1.0 / 2.0) + C.al[m] * (1.0 / 2.0);
                C.ar[m] = B.ar[m] * (1.0 / 2.0) + C.ar[m] * (1.0 / 2.0);
                C.rk[m] = B.rk[m] * (1.0 / 2.0) + C.rk[m] * (1.0 / 2.0);
                C.e[m] = B.e[m] * (1.0 / 2.0) + C.e[m] * (1.0 / 2.0);
                C.are[m] = B.are[m] * (1.0 / 2.0) + C.are[m] * (1.0 / 2.0);
            }
            C.rho = B.rho * (1.0 / 2.0) + C.rho * (1.0 / 2.0);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                C.ru[q] = B.ru[q] * (1.0 / 2.0) + C.ru[q] * (1.0 / 2.0);
                C.u[q] = B.u[q] * (1.0 / 2.0) + C.u[q] * (1.0 / 2.0);
            }
            C.E = B.E * (1.0 / 2.0) + C.E * (1.0 / 2.0);
            C.p = B.p * (1.0 / 2.0) + C.p * (1.0 / 2.0);
#pragma unroll
            for (int q = 0; q < 9; ++q) C.V[q] = B.V[q] * (1.0 / 2.0) + C.V[q] * (1.0 / 2.0);
            C.hj2 = B.hj2 * (1.0 / 2.0) + C.hj2 * (1.0 / 2.0);
            if (ops & 1u) {
                st |= plastic_cell<NM, false, T>(lv, C.al, C.ar, C.rk, C.e, C.are, C.V);
                C.hj2 = c2p_s<NM, false, T>(lv, C.al, C.ar, C.are, C.ru, C.E, C.V, C.rk, C.e, C.rho, C.u, C.p, dH);
            }
            if (ops & 2u) cleanup_v<NM, T>(lv, C.al, C.V);
            if (ops & 4u) st |= cleanup_alpha<NM, T>(C.al);
            if (ops & 16u) {
                bool bad = false, tbad = false;
#pragma unroll
                for (int m = 0; m < NM; ++m) {
                    bad |= tisnan(C.al[m]) || tisnan(C.ar[m]) || tisnan(C.rk[m]) ||
                           (lv.plastic && (tisnan(C.e[m]) || tisnan(C.are[m])));
                    tbad |= tan_nan(C.al[m]) || tan_nan(C.ar[m]) || tan_nan(C.rk[m]) || tan_nan(C.e[m]) ||
                            tan_nan(C.are[m]);
                }
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    bad |= tisnan(C.ru[q]) || tisnan(C.u[q]);
                    tbad |= tan_nan(C.ru[q]) || tan_nan(C.u[q]);
                }
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    bad |= tisnan(C.V[q]);
                    tbad |= tan_nan(C.V[q]);
                }
                bad |= tisnan(C.E) || tisnan(C.p);
                tbad |= tan_nan(C.E) || tan_nan(C.p);
                if (bad) st |= 1u;
                if (tbad) st |= 128u;
            }
            if (ops & 8u) cv = tsound_cand<NM, K>(lv, Uo, C, i, j, k, id, cand);
        }
        store_tcell<NM, K, STAGE == 2>(lv, Uo, i, j, k, C);
    }
    if (STAGE == 2 && (ops & 8u)) block_min_to(cv, &clk->dtmin_bits);
    if (st) atomicOr(&clk->status, st);
}

// getTimeStep of a state (k_sound_dt_s) with the candidates' derivatives.
template <int NM, int K>
__global__ __launch_bounds__(256) void k_tsound_dt_s(Level lv, DMF<K> U, Clock* clk, double* cand) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double cv = __builtin_huge_val();
    if (id < (long long)nx * ny * nz) {
        const int i = (int)(id % nx);
        const int j = (int)((id / nx) % ny);
        const int k = (int)(id / ((long long)nx * ny));
        TCell<NM, K> C;
        load_tcell<NM, K>(lv, U, i, j, k, C);
        cv = tsound_cand<NM, K>(lv, U, C, i, j, k, id, cand);
    }
    block_min_to(cv, &clk->dtmin_bits);
}

// setInitialConditions of a solid level (k_init_p2c_s) with derivatives: the
// thermal -> total pressure shift and primitiveToConservative depend on the
// EOS parameters through the materials' derivative parts.
template <int NM, int K>
__global__ __launch_bounds__(256) void k_tinit_p2c_s(Level lv, DMF<K> U, int shift) {
    using T = Dual<K>;
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)nx * ny * nz) return;
    const int i = (int)(id % nx);
    const int j = (int)((id / nx) % ny);
    const int k = (int)(id / ((long long)nx * ny));
    SPrim<NM, T> P;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        P.a[m] = U.get(lv.cs[ALPHA] + m, i, j, k);
        P.r[m] = U.get(lv.cs[RHO_K] + m, i, j, k);
        P.e[m] = lv.plastic ? U.get(lv.cs[EPSILON] + m, i, j, k) : T(0.0);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) P.u[q] = U.get(lv.cs[VELOCITY] + q, i, j, k);
    P.p = U.get(lv.cs[PRESSURE], i, j, k);
#pragma unroll
    for (int q = 0; q < 9; ++q) P.V[q] = U.get(lv.cs[V_TENSOR] + q, i, j, k);
    if (shift) {
        T ar[NM], G, Ent, Pnt;
#pragma unroll
        for (int m = 0; m < NM; ++m) ar[m] = P.a[m] * P.r[m];
        eos_sums_s<NM, false, T>(lv, P.a, ar, P.r, U.get(lv.cs[HJ2], i, j, k), G, Ent, Pnt);
        P.p += tdiv(Pnt, G);
        U.put(lv.cs[PRESSURE], i, j, k, P.p);
    }
    SSt<NM, T> S;
    p2c_sound_s<NM, false, T>(lv, P, 0, S);
    T dH[9];
    const T hj2 = hencky(P.V, dH);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        U.put(lv.cs[ALPHARHO] + m, i, j, k, S.ar[m]);
        if (lv.plastic) U.put(lv.cs[ALPHARHOEPSILON] + m, i, j, k, S.are[m]);
    }
    U.put(lv.cs[RHO], i, j, k, S.rho);
#pragma unroll
    for (int q = 0; q < 3; ++q) U.put(lv.cs[RHOU] + q, i, j, k, S.u[q] * S.rho);
    U.put(lv.cs[TOTAL_E], i, j, k, S.E);
    U.put(lv.cs[HJ2], i, j, k, hj2);
#pragma unroll
    for (int q = 0; q < 9; ++q) U.put(lv.cs[DEVH] + q, i, j, k, dH[q]);
}

// The lowest cell index whose candidate attains the level minimum.
__global__ __launch_bounds__(256) void k_targmin(const Clock* clk, const double* cand, long long ncell,
                                                 unsigned long long* arg) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < ncell && __builtin_bit_cast(unsigned long long, cand[id]) == clk->dtmin_bits)
        atomicMin(arg, (unsigned long long)id);
}

// k_clock_dt (driver form) with dt's derivatives: CFL times the minimising
// cell's candidate derivatives; minus the derivatives of t where dt is
// clipped to the final time; times 0.2 over the first five steps.
template <int K>
__global__ void k_tclock_dt(Clock* clk, double cfl, const double* cand, long long ncell, unsigned long long* arg,
                            double* tdt, const double* tt) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (clk->done) return;
    if (!(clk->t < clk->final_time)) {
        clk->done = 1;
        return;
    }
    const double mn = __builtin_bit_cast(double, clk->dtmin_bits);
    double dt = cfl * mn;
    const unsigned long long a = *arg;
    double g[K];
#pragma unroll
    for (int t = 0; t < K; ++t) g[t] = a < (unsigned long long)ncell ? cfl * cand[(t + 1) * ncell + (long long)a] : 0.0;
    if (clk->t + dt > clk->final_time) {
        dt = clk->final_time - clk->t;
#pragma unroll
        for (int t = 0; t < K; ++t) g[t] = -tt[t];
    }
    if (clk->n < 5) {
        dt *= 0.2;
#pragma unroll
        for (int t = 0; t < K; ++t) g[t] *= 0.2;
    }
    clk->dt = dt;
    tdt[0] = dt;
#pragma unroll
    for (int t = 0; t < K; ++t) tdt[1 + t] = g[t];
    clk->dtmin_bits = 0x7FEFFFFFFFFFFFFFull;
    *arg = ~0ull;
}
// start of a run (k_clock_final of tj_begin_run): the final time, no
// minimising cell yet, t's derivatives zero
template <int K>
__global__ void k_tclock_begin(Clock* clk, double final_time, unsigned long long* arg, double* tt) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    clk->final_time = final_time;
    *arg = ~0ull;
#pragma unroll
    for (int t = 0; t < K; ++t) tt[t] = 0.0;
}
// k_clock_advance with t's derivatives.
template <int K>
__global__ void k_tclock_advance(Clock* clk, const double* tdt, double* tt) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (clk->done) return;
    if (clk->status & ST_STOP) {
        clk->done = 1;
        return;
    }
    clk->t += clk->dt;
    clk->n += 1;
#pragma unroll
    for (int t = 0; t < K; ++t) tt[t] += tdt[1 + t];
}
// alpha of a cell with the boundary map (load_alpha of tj_fluid.h), for arec
template <typename R, int K>
__device__ __forceinline__ R load_alpha(const Level& lv, const DMF<K>& U, int m, int i, int j, int k) {
    bool f[3];
    map3(lv, i, j, k, f);
    return U.get(lv.cs[ALPHA] + m, i, j, k);
}

// THINC/BVD history planes (Hist) with derivative parts: plane blocks of
// values, then K blocks of derivatives, each `*blk` doubles apart.
struct THist {
    double* urstar;
    const double* star_in;
    double* star_out;
    const double* edge_in;
    double* edge_out;
    long long pstride, estride;  // plane strides (cells, faces of the direction)
    long long pblk, eblk;        // derivative block strides (NM planes)
};
template <int K>
__device__ __forceinline__ Dual<K> hget(const double* p, long long blk, long long idx) {
    Dual<K> x;
    x.v = p[idx];
#pragma unroll
    for (int t = 0; t < K; ++t) x.d[t] = p[(t + 1) * blk + idx];
    return x;
}
template <int K>
__device__ __forceinline__ void hput(double* p, long long blk, long long idx, const Dual<K>& x) {
    p[idx] = x.v;
#pragma unroll
    for (int t = 0; t < K; ++t) p[(t + 1) * blk + idx] = x.d[t];
}

// ------------------------------------------------------------------ fluid levels
// The MUSCL(+THINC/BVD) + HLLC path of fluid levels (no mixture, no geometric
// source), mirroring k_flux_x / k_flux_march, update_body with finish_cell,
// init_p2c_body and sound_dt_body.
template <int NM, int K>
__device__ __forceinline__ void load_tprim_f(const Level& lv, const DMF<K>& U, int i, int j, int k,
                                             Prim<NM, Dual<K>>& P) {
    bool f[3];
    map3(lv, i, j, k, f);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        P.a[m] = U.get(lv.cs[ALPHA] + m, i, j, k);
        P.r[m] = U.get(lv.cs[RHO_K] + m, i, j, k);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const Dual<K> v = U.get(lv.cs[VELOCITY] + r, i, j, k);
        P.u[r] = f[r] ? -v : v;
    }
    P.p = U.get(lv.cs[PRESSURE], i, j, k);
}

// lane src's fluid face state within the wave (no mixture part)
template <int NM, int K>
__device__ __forceinline__ FSt<NM, Dual<K>> shfl_tfst(const FSt<NM, Dual<K>>& S, int src) {
    FSt<NM, Dual<K>> T;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        T.al[m] = shfl_any(S.al[m], src);
        T.ar[m] = shfl_any(S.ar[m], src);
    }
    T.rho = shfl_any(S.rho, src);
#pragma unroll
    for (int q = 0; q < 3; ++q) T.u[q] = shfl_any(S.u[q], src);
    T.E = shfl_any(S.E, src);
    T.p = shfl_any(S.p, src);
    T.a = shfl_any(S.a, src);
    return T;
}

// Faces along D with one reconstruction per cell, in the tiles of k_tflux_ts
// (64 x 1 cells in x, 8 x 8 in y and z; face f takes UR(f-1) from the lane
// one cell below by shuffle).  A lane reconstructs its cell: MUSCL, and with
// THINC (TH) the BVD selection of the alpha faces of flux_x_body /
// flux_march_body (the cell's and its neighbours' THINC reconstructions, the
// box-edge semantics and the stale history planes, read with their
// derivative parts) and cleanUpAlpha; then p2c + sound speed of both faces.
// The lane of face f then writes the history of this pass: urstar of cell f
// (where mixed), the star alpha of cell f and, at box-high faces, the edge
// plane.  Within a pass urstar is read only where a cell is not mixed and
// written only where it is, so the writes race with no read.
template <int NM, int K, int D, bool TH>
__global__ __launch_bounds__(256) void k_tflux_tf(Level lv, DMF<K> U, Faces F, THist h, Clock* clk) {
    if (clk->done) return;
    constexpr int LN = D == 0 ? 64 : 8;  // lanes along D
    constexpr int LT = 64 / LN;           // lanes along x (D = 1, 2)
    constexpr int PER = LN - 1;           // faces per tile column
    constexpr int NF = 2 * NM + 5;
    using T = Dual<K>;
    const int lane = threadIdx.x & 63;
    const int ld = lane / LT, lt = lane % LT;
    const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + wave_in_block();
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2], nd = lv.n[D];
    const int nseg = (nd + 1 + PER - 1) / PER;
    const int nxt = D == 0 ? 1 : (nx + LT - 1) / LT;
    const int no = D == 0 ? ny * nz : (D == 1 ? nz : ny);
    if (wave >= (long long)nxt * nseg * no) return;  // uniform per wave
    const int xt = (int)(wave % nxt);
    const long long rest = wave / nxt;
    const int seg = (int)(rest % nseg);
    const int o = (int)(rest / nseg);
    const int c = seg * PER - 1 + ld;  // the lane's cell along D
    bool fl;
    const int s = bcmap(c, nd, lv.bc_lo[D], lv.bc_hi[D], fl);
    const int i0 = xt * LT + lt;
    const bool act = D == 0 || i0 < nx;
    const int i = D == 0 ? s : (act ? i0 : nx - 1);
    const int j = D == 0 ? o % ny : (D == 1 ? s : o);
    const int k = D == 0 ? o / ny : (D == 1 ? o : s);
    auto ci = [&](int r) { return D == 0 ? r : i; };
    auto cj = [&](int r) { return D == 1 ? r : j; };
    auto ck = [&](int r) { return D == 2 ? r : k; };
    auto cellidx = [&](int r) { return ((long long)ck(r) * ny + cj(r)) * nx + ci(r); };
    const long long cstride = D == 0 ? 1 : (D == 1 ? (long long)nx : (long long)nx * ny);
    // face index of face g along D in the face grid of this direction
    auto faceidx = [&](int g) {
        return D == 0 ? ((long long)k * ny + j) * (nx + 1) + g
                      : (D == 1 ? ((long long)k * (ny + 1) + g) * nx + i : ((long long)g * ny + j) * nx + i);
    };
    const bool face = act && ld >= 1 && c >= 0 && c <= nd;  // this lane computes face c
    unsigned st = 0;
    Prim<NM, T> PL, PR;
    {
        Prim<NM, T> Pm, Pc, Pp;
        load_tprim_f<NM, K>(lv, U, ci(s - 1), cj(s - 1), ck(s - 1), Pm);
        load_tprim_f<NM, K>(lv, U, ci(s), cj(s), ck(s), Pc);
        load_tprim_f<NM, K>(lv, U, ci(s + 1), cj(s + 1), ck(s + 1), Pp);
        muscl<NM, T>(Pm, Pc, Pp, PL, PR);
    }
    if (TH) {
        const uint8_t ef = lv.edge[D][s];
        const bool blo = ef & 1, bhi = ef & 2;
        const T ynx = T(lv.yn[0]), yny = T(lv.yn[1]);
        ARec<NM, T> Acur;
        arec<NM, T>(lv, U, ci(s), cj(s), ck(s), D, ynx, yny, Acur);
        const int mixedFlag = (Acur.mixed >> (NM - 1)) & 1;
        const long long cc = cellidx(s);
        if (mixedFlag) {
            ARec<NM, T> Aprev, Anext;
            if (!blo) arec<NM, T>(lv, U, ci(s - 1), cj(s - 1), ck(s - 1), D, ynx, yny, Aprev);
            if (!bhi) arec<NM, T>(lv, U, ci(s + 1), cj(s + 1), ck(s + 1), D, ynx, yny, Anext);
            T URl[NM], TRl[NM], ULr[NM], TLr[NM], thl[NM], thr[NM];
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                if (blo) {
                    URl[m] = load_alpha<T>(lv, U, m, ci(s - 1), cj(s - 1), ck(s - 1));
                    TRl[m] = T(0.0);
                } else {
                    URl[m] = Aprev.ura[m];
                    TRl[m] = ((Aprev.mixed >> m) & 1) ? Aprev.thr[m]
                                                      : hget<K>(h.urstar, h.pblk, m * h.pstride + cc - cstride);
                }
                if (bhi) {
                    ULr[m] = load_alpha<T>(lv, U, m, ci(s + 1), cj(s + 1), ck(s + 1));
                    TLr[m] = hget<K>(h.edge_in, h.eblk, m * h.estride + faceidx(s + 1));
                } else {
                    ULr[m] = Anext.ula[m];
                    TLr[m] = ((Anext.mixed >> m) & 1) ? Anext.thl[m]
                                                      : hget<K>(h.star_in, h.pblk, m * h.pstride + cc + cstride);
                }
                if ((Acur.mixed >> m) & 1) {
                    thl[m] = Acur.thl[m];
                    thr[m] = Acur.thr[m];
                } else {
                    thl[m] = hget<K>(h.star_in, h.pblk, m * h.pstride + cc);
                    thr[m] = hget<K>(h.urstar, h.pblk, m * h.pstride + cc);
                }
            }
            bvd<NM, T>(1, URl, TRl, ULr, TLr, thl, thr, PL.a, PR.a);
        }
        st |= cleanup_alpha<NM, T>(PL.a);
        st |= cleanup_alpha<NM, T>(PR.a);
        if (face && c < nd) {  // urstar of cell c, the upper cell of face c (c == s here)
#pragma unroll
            for (int m = 0; m < NM; ++m)
                if ((Acur.mixed >> m) & 1) hput<K>(h.urstar, h.pblk, m * h.pstride + cc, Acur.thr[m]);
        }
    }
    FSt<NM, T> Lf;
    {
        FSt<NM, T> SRs;
        p2c_sound<NM, T>(lv, PR, SRs);
        Lf = shfl_tfst<NM, K>(SRs, (c == 0 || ld == 0) ? lane : lane - LT);
    }
    if (c == 0 && lv.bc_lo[D]) flip_state<NM, T>(Lf, D);
    FSt<NM, T> Rf;
    p2c_sound<NM, T>(lv, PL, Rf);
    if (c == nd && lv.bc_hi[D]) flip_state<NM, T>(Rf, D);
    if (face) {
        T Fv[NF], sa[NM];
        hllc<NM, T>(Lf, Rf, D, Fv, sa);
        const long long fi = faceidx(c);
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            F.p[q * F.cs + fi] = Fv[q].v;
#pragma unroll
            for (int t = 0; t < K; ++t) F.p[(NF * (t + 1) + q) * F.cs + fi] = Fv[q].d[t];
        }
        if (TH) {
            if (c < nd) {
#pragma unroll
                for (int m = 0; m < NM; ++m) hput<K>(h.star_out, h.pblk, m * h.pstride + cellidx(c), sa[m]);
            }
            if (c == nd || (c > 0 && (lv.edge[D][c - 1] & 2))) {
#pragma unroll
                for (int m = 0; m < NM; ++m) hput<K>(h.edge_out, h.eblk, m * h.estride + fi, sa[m]);
            }
        }
    }
    if (st) atomicOr(&clk->status, st);
}

// conservativeToPrimitive (c2p of tj_fluid.h, no mixture) on dual numbers
template <int NM, int K>
__device__ __forceinline__ void tc2p_f(const Level& lv, const Dual<K>* al, const Dual<K>* ar, const Dual<K>* ru,
                                       Dual<K> E, Dual<K>* rk, Dual<K>& rho, Dual<K>* u, Dual<K>& p) {
    using T = Dual<K>;
    rho = T(0.0);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        rk[m] = tdiv(ar[m], al[m]);
        rho += ar[m];
    }
    T ke = T(0.0);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        u[q] = tdiv(ru[q], rho);
        ke += 0.5 * rho * u[q] * u[q];
    }
    T G, Ent, Pnt;
    eos_sums<NM, T>(lv, al, ar, G, Ent, Pnt);
    p = tdiv(E - ke - Ent + Pnt, G);
}
// getSoundSpeed (sound of tj_fluid.h, no mixture) on dual numbers
template <int NM, int K>
__device__ __forceinline__ Dual<K> tsound_f(const Level& lv, const Dual<K>* al, const Dual<K>* ar, const Dual<K>* rk,
                                            Dual<K> rho, Dual<K> p) {
    using T = Dual<K>;
    T a = T(0.0), xt = T(0.0);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const MatT<T>& M = mat<T>(lv, m);
        const T c2 = tdiv(p * M.gp1, rk[m]) - tdiv(M.pref, rk[m]);
        a += tmax(T(0.0), tdiv(M.xi * c2 * ar[m], rho));
        xt += M.xi * al[m];
    }
    if (a <= 0.0) a = T(1E-6);
    return tsqrt(tdiv(a, xt));
}
template <int NM, int K>
__device__ __forceinline__ double tcand_f(const Level& lv, const DMF<K>& U, Dual<K> a, const Dual<K>* u, int i, int j,
                                          int k, long long id, double* cand) {
    using T = Dual<K>;
    U.put(lv.cs[SOUNDSPEED], i, j, k, a);
    T c = T(__builtin_huge_val());
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (d < lv.dim) c = tmin(c, tdiv(lv.dx[d], a + tabs(u[d])));
    const long long nc = lv.ncells;
    cand[id] = c.v;
#pragma unroll
    for (int t = 0; t < K; ++t) cand[(t + 1) * nc + id] = c.d[t];
    return c.v;
}

template <int NM, int STAGE, int K>
__global__ __launch_bounds__(256) void k_tupdate_f(Level lv, DMF<K> Uin, DMF<K> Ub, DMF<K> Uo, Faces F0, Faces F1,
                                                   Faces F2, Clock* clk, const double* tdt, unsigned ops, double* cand) {
    if (clk->done) return;
    using T = Dual<K>;
    constexpr int NF = 2 * NM + 5;
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncell = (long long)nx * ny * nz;
    double cv = __builtin_huge_val();
    unsigned st = 0;
    if (id < ncell) {
        const int i = (int)(id % nx);
        const int j = (int)((id / nx) % ny);
        const int k = (int)(id / ((long long)nx * ny));
        T dt;
        dt.v = clk->dt;
#pragma unroll
        for (int t = 0; t < K; ++t) dt.d[t] = tdt[1 + t];
        T al[NM], ar[NM], ru[3], E, a0[NM];
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            al[m] = Uin.get(lv.cs[ALPHA] + m, i, j, k);
            ar[m] = Uin.get(lv.cs[ALPHARHO] + m, i, j, k);
            a0[m] = al[m];
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) ru[q] = Uin.get(lv.cs[RHOU] + q, i, j, k);
        E = Uin.get(lv.cs[TOTAL_E], i, j, k);
        for (int d = 0; d < lv.dim; ++d) {
            const Faces& Fd = d == 0 ? F0 : (d == 1 ? F1 : F2);
            const long long lo = ((long long)k * Fd.ny + j) * Fd.nx + i;
            const long long hi = ((long long)(k + (d == 2)) * Fd.ny + (j + (d == 1))) * Fd.nx + (i + (d == 0));
            auto fv = [&](int q, long long o) {
                T x;
                x.v = Fd.p[q * Fd.cs + o];
#pragma unroll
                for (int t = 0; t < K; ++t) x.d[t] = Fd.p[(NF * (t + 1) + q) * Fd.cs + o];
                return x;
            };
            const T dtdx = tdiv(dt, lv.dx[d]);
            const T s0 = fv(2 * NM + 4, lo), s1 = fv(2 * NM + 4, hi);
#pragma unroll
            for (int m = 0; m < NM; ++m) al[m] += dtdx * (fv(m, lo) - fv(m, hi) - a0[m] * (s0 - s1));
#pragma unroll
            for (int m = 0; m < NM; ++m) ar[m] += dtdx * (fv(NM + m, lo) - fv(NM + m, hi));
#pragma unroll
            for (int q = 0; q < 3; ++q) ru[q] += dtdx * (fv(2 * NM + q, lo) - fv(2 * NM + q, hi));
            E += dtdx * (fv(2 * NM + 3, lo) - fv(2 * NM + 3, hi));
        }
        T rk[NM], rho, u[3], p;
        tc2p_f<NM, K>(lv, al, ar, ru, E, rk, rho, u, p);
        if (STAGE == 2) {
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                al[m] = Ub.get(lv.cs[ALPHA] + m, i, j, k) * (1.0 / 2.0) + al[m] * (1.0 / 2.0);
                ar[m] = Ub.get(lv.cs[ALPHARHO] + m, i, j, k) * (1.0 / 2.0) + ar[m] * (1.0 / 2.0);
                rk[m] = Ub.get(lv.cs[RHO_K] + m, i, j, k) * (1.0 / 2.0) + rk[m] * (1.0 / 2.0);
            }
            rho = Ub.get(lv.cs[RHO], i, j, k) * (1.0 / 2.0) + rho * (1.0 / 2.0);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                ru[q] = Ub.get(lv.cs[RHOU] + q, i, j, k) * (1.0 / 2.0) + ru[q] * (1.0 / 2.0);
                u[q] = Ub.get(lv.cs[VELOCITY] + q, i, j, k) * (1.0 / 2.0) + u[q] * (1.0 / 2.0);
            }
            E = Ub.get(lv.cs[TOTAL_E], i, j, k) * (1.0 / 2.0) + E * (1.0 / 2.0);
            p = Ub.get(lv.cs[PRESSURE], i, j, k) * (1.0 / 2.0) + p * (1.0 / 2.0);
            if (ops & 4u) st |= cleanup_alpha<NM, T>(al);
            if (ops & 16u) {
                bool bad = false, tbad = false;
#pragma unroll
                for (int m = 0; m < NM; ++m) {
                    bad |= tisnan(al[m]) || tisnan(ar[m]) || tisnan(rk[m]);
                    tbad |= tan_nan(al[m]) || tan_nan(ar[m]) || tan_nan(rk[m]);
                }
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    bad |= tisnan(ru[q]) || tisnan(u[q]);
                    tbad |= tan_nan(ru[q]) || tan_nan(u[q]);
                }
                bad |= tisnan(E) || tisnan(p);
                tbad |= tan_nan(E) || tan_nan(p);
                if (bad) st |= 1u;
                if (tbad) st |= 128u;
            }
            if (ops & 8u) cv = tcand_f<NM, K>(lv, Uo, tsound_f<NM, K>(lv, al, ar, rk, rho, p), u, i, j, k, id, cand);
        }
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            Uo.put(lv.cs[ALPHA] + m, i, j, k, al[m]);
            Uo.put(lv.cs[ALPHARHO] + m, i, j, k, ar[m]);
            Uo.put(lv.cs[RHO_K] + m, i, j, k, rk[m]);
        }
        Uo.put(lv.cs[RHO], i, j, k, rho);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            Uo.put(lv.cs[RHOU] + q, i, j, k, ru[q]);
            Uo.put(lv.cs[VELOCITY] + q, i, j, k, u[q]);
        }
        Uo.put(lv.cs[TOTAL_E], i, j, k, E);
        Uo.put(lv.cs[PRESSURE], i, j, k, p);
    }
    if (STAGE == 2 && (ops & 8u)) block_min_to(cv, &clk->dtmin_bits);
    if (st) atomicOr(&clk->status, st);
}

template <int NM, int K>
__global__ __launch_bounds__(256) void k_tsound_dt_f(Level lv, DMF<K> U, Clock* clk, double* cand) {
    using T = Dual<K>;
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double cv = __builtin_huge_val();
    if (id < (long long)nx * ny * nz) {
        const int i = (int)(id % nx);
        const int j = (int)((id / nx) % ny);
        const int k = (int)(id / ((long long)nx * ny));
        T al[NM], ar[NM], rk[NM], u[3];
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            al[m] = U.get(lv.cs[ALPHA] + m, i, j, k);
            ar[m] = U.get(lv.cs[ALPHARHO] + m, i, j, k);
            rk[m] = U.get(lv.cs[RHO_K] + m, i, j, k);
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) u[q] = U.get(lv.cs[VELOCITY] + q, i, j, k);
        const T a = tsound_f<NM, K>(lv, al, ar, rk, U.get(lv.cs[RHO], i, j, k), U.get(lv.cs[PRESSURE], i, j, k));
        cv = tcand_f<NM, K>(lv, U, a, u, i, j, k, id, cand);
    }
    block_min_to(cv, &clk->dtmin_bits);
}

template <int NM, int K>
__global__ __launch_bounds__(256) void k_tinit_p2c_f(Level lv, DMF<K> U, int shift) {
    using T = Dual<K>;
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)nx * ny * nz) return;
    const int i = (int)(id % nx);
    const int j = (int)((id / nx) % ny);
    const int k = (int)(id / ((long long)nx * ny));
    Prim<NM, T> P;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        P.a[m] = U.get(lv.cs[ALPHA] + m, i, j, k);
        P.r[m] = U.get(lv.cs[RHO_K] + m, i, j, k);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) P.u[q] = U.get(lv.cs[VELOCITY] + q, i, j, k);
    P.p = U.get(lv.cs[PRESSURE], i, j, k);
    if (shift) {
        T ar[NM], G, Ent, Pnt;
#pragma unroll
        for (int m = 0; m < NM; ++m) ar[m] = P.a[m] * P.r[m];
        eos_sums<NM, T>(lv, P.a, ar, G, Ent, Pnt);
        P.p += tdiv(Pnt, G);
        U.put(lv.cs[PRESSURE], i, j, k, P.p);
    }
    FSt<NM, T> S;
    p2c_sound<NM, T>(lv, P, S);
#pragma unroll
    for (int m = 0; m < NM; ++m) U.put(lv.cs[ALPHARHO] + m, i, j, k, S.ar[m]);
    U.put(lv.cs[RHO], i, j, k, S.rho);
#pragma unroll
    for (int q = 0; q < 3; ++q) U.put(lv.cs[RHOU] + q, i, j, k, S.u[q] * S.rho);
    U.put(lv.cs[TOTAL_E], i, j, k, S.E);
}

// Z = X over ncomp components of the valid cells (views of any strides)
__global__ __launch_bounds__(256) void k_copy_mf(MF X, MF Z, int ncomp, int nx, int ny, int nz) {
    const long long ncell = (long long)nx * ny * nz;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ncell * ncomp) return;
    const int c = (int)(id / ncell);
    const long long cell = id % ncell;
    const int i = (int)(cell % nx);
    const int j = (int)((cell / nx) % ny);
    const int k = (int)(cell / ((long long)nx * ny));
    Z.at(c, i, j, k) = X.at(c, i, j, k);
}
// value and derivatives of one component of one cell into a trace row
template <int K>
__global__ void k_tprobe(DMF<K> U, int c, int i, int j, int k, double* row) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const Dual<K> x = U.get(c, i, j, k);
    row[0] = x.v;
#pragma unroll
    for (int t = 0; t < K; ++t) row[1 + t] = x.d[t];
}

}  // namespace tj
