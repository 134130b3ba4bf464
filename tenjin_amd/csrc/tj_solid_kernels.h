// Kernels of the elastoplastic path (SOLID = 1): directional 5-wave flux
// passes and the fused update / post-step kernel.
#pragma once
#include "tj_kernels.h"
#include "tj_solid.h"

namespace tj {
inline namespace TJ_VIEW_NS {

template <int NM, typename R>
__device__ __forceinline__ SSt<NM, R> shfl_state(const SSt<NM, R>& S, int src) {
    SSt<NM, R> T;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        T.al[m] = shfl_any(S.al[m], src);
        T.ar[m] = shfl_any(S.ar[m], src);
        T.are[m] = shfl_any(S.are[m], src);
    }
    T.rho = shfl_any(S.rho, src);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        T.u[q] = shfl_any(S.u[q], src);
        T.sig[q] = shfl_any(S.sig[q], src);
    }
    T.E = shfl_any(S.E, src);
    T.p = shfl_any(S.p, src);
    T.a = shfl_any(S.a, src);
#pragma unroll
    for (int q = 0; q < 9; ++q) T.V[q] = shfl_any(S.V[q], src);
    return T;
}

// lane - 1's state (lane 0 keeps its own) by DPP row moves, as lane_dn
template <int NM, typename R>
__device__ __forceinline__ SSt<NM, R> state_dn(const SSt<NM, R>& S) {
    SSt<NM, R> T;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        T.al[m] = lane_dn(S.al[m]);
        T.ar[m] = lane_dn(S.ar[m]);
        T.are[m] = lane_dn(S.are[m]);
    }
    T.rho = lane_dn(S.rho);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        T.u[q] = lane_dn(S.u[q]);
        T.sig[q] = lane_dn(S.sig[q]);
    }
    T.E = lane_dn(S.E);
    T.p = lane_dn(S.p);
    T.a = lane_dn(S.a);
#pragma unroll
    for (int q = 0; q < 9; ++q) T.V[q] = lane_dn(S.V[q]);
    return T;
}

// The left image of a reconstructed state across a reflective face:
// reflect_odd flips the normal velocity and momentum only (boundary.cpp:3-42).
template <int NM, typename R>
__device__ __forceinline__ void flip_s(SSt<NM, R>& S, int d) { S.u[d] = -S.u[d]; }

// x pass of the solid path (same wave layout as k_flux_x, no THINC).
template <int NM, bool RW, typename R>
__global__ __launch_bounds__(256) void k_flux_x_s(Level lv, MF U, FacesT<R> F, Clock* clk) {
    if (clk->done) return;
    constexpr int PER = 63;
    constexpr int NF = 3 * NM + 17;
    const int lane = threadIdx.x & 63;
    const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + wave_in_block();
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const int nseg = (nx + 1 + PER - 1) / PER;
    if (wave >= (long long)nseg * ny * nz) return;
    const int seg = (int)(wave % nseg);
    const long long row = wave / nseg;
    const int j = (int)(row % ny), k = (int)(row / ny);
    const int c = seg * PER - 1 + lane;
    bool fl;
    const int s = bcmap(c, nx, lv.bc_lo[0], lv.bc_hi[0], fl);
    SSt<NM, R> SLs, SRs;
    {
        SPrim<NM, R> Pm, Pc, Pp, PL, PR;
        load_sprim<NM, R>(lv, U, s - 1, j, k, Pm);
        load_sprim<NM, R>(lv, U, s, j, k, Pc);
        load_sprim<NM, R>(lv, U, s + 1, j, k, Pp);
        muscl_s<NM, R>(Pm, Pc, Pp, PL, PR);
        p2c_sound_s<NM, RW, R>(lv, PL, 0, SLs);
        p2c_sound_s<NM, RW, R>(lv, PR, 0, SRs);
    }
    SSt<NM, R> Lf = state_dn<NM, R>(SRs);
    unsigned st = 0;
    if (lane >= 1 && c >= 0 && c <= nx) {
        if (c == 0) {
            Lf = SRs;
            if (lv.bc_lo[0]) flip_s<NM, R>(Lf, 0);
        }
        SSt<NM, R> Rf = SLs;
        if (c == nx && lv.bc_hi[0]) flip_s<NM, R>(Rf, 0);
        R Fv[NF];
        wave5<NM, RW, R>(lv, Lf, Rf, 0, Fv, st);
#pragma unroll
        for (int q = 0; q < NF; ++q) F.at(q, c, j, k) = Fv[q];
    }
    if (st) atomicOr(&clk->status, st);
}

// A solid face state in LDS, one slot per field, lane-minor (conflict-free).
template <int NM, typename R>
struct SStSlots {
    static constexpr int N = 5 * NM + 15;
};
template <int NM, typename R>
__device__ __forceinline__ void sst_store(R* sh, int lane, const SSt<NM, R>& S) {
    int q = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        sh[(q++) * 64 + lane] = S.al[m];
        sh[(q++) * 64 + lane] = S.ar[m];
        sh[(q++) * 64 + lane] = S.are[m];
    }
    sh[(q++) * 64 + lane] = S.rho;
    sh[(q++) * 64 + lane] = S.E;
    sh[(q++) * 64 + lane] = S.p;
    sh[(q++) * 64 + lane] = S.a;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        sh[(q++) * 64 + lane] = S.u[c];
        sh[(q++) * 64 + lane] = S.sig[c];
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) sh[(q++) * 64 + lane] = S.V[c];
}
template <int NM, typename R>
__device__ __forceinline__ void sst_load(const R* sh, int lane, SSt<NM, R>& S) {
    int q = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        S.al[m] = sh[(q++) * 64 + lane];
        S.ar[m] = sh[(q++) * 64 + lane];
        S.are[m] = sh[(q++) * 64 + lane];
    }
    S.rho = sh[(q++) * 64 + lane];
    S.E = sh[(q++) * 64 + lane];
    S.p = sh[(q++) * 64 + lane];
    S.a = sh[(q++) * 64 + lane];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        S.u[c] = sh[(q++) * 64 + lane];
        S.sig[c] = sh[(q++) * 64 + lane];
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) S.V[c] = sh[(q++) * 64 + lane];
}

// y / z pass of the solid path: lanes along x (coalesced), marching along D
// over a segment of faces.  The right state of row r-1 waits for face r in
// the wave's own LDS slice (not in registers).  The primitive rows slide:
// rows r-1 and r stay in registers and row r+1 is the one load per row.
// Reloading all three rows from cache missed the L2 (a 4 MB L2 per XCD
// against the rows of all resident waves): 565 B/cell of HBM traffic against
// 336, which made the FP32 (mixed) pass bandwidth-bound (measured: FP32 pass
// 1.6 -> 1.1 ms, FP64 3.2 -> 3.0 ms on 256^3).
template <int NM, int D, bool RW, typename R>
__device__ __forceinline__ void flux_lmarch_s_body(const Level& lv, const MF& U, const FacesT<R>& F, int seglen, Clock* clk) {
    if (clk->done) return;
    constexpr int NF = 3 * NM + 17;
    constexpr int NS = SStSlots<NM, R>::N;
    __shared__ R shm[4 * NS * 64];
    const int lane = threadIdx.x & 63;
    R* sh = shm + wave_in_block() * NS * 64;
    const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + wave_in_block();
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const int nd = lv.n[D];
    const int no = (D == 1) ? nz : ny;
    const int nstrip = (nx + 63) / 64;
    const int nseg = (nd + 1 + seglen - 1) / seglen;
    if (wave >= (long long)nstrip * nseg * no) return;  // uniform per wave; no block barrier below
    const int strip = (int)(wave % nstrip);
    const long long rest = wave / nstrip;
    const int seg = (int)(rest % nseg);
    const int o = (int)(rest / nseg);
    const int i0 = strip * 64 + lane;
    const bool act = i0 < nx;
    const int i = act ? i0 : nx - 1;
    const int f0 = seg * seglen;
    const int f1 = min(f0 + seglen, nd + 1);
    const int rb = max(f0 - 1, 0), re = min(f1 - 1, nd - 1);
    auto cj = [&](int r) { return D == 1 ? r : o; };
    auto ck = [&](int r) { return D == 2 ? r : o; };
    auto faceidx = [&](int r) {
        return D == 1 ? ((long long)ck(r) * (ny + 1) + r) * nx + i : ((long long)r * ny + cj(r)) * nx + i;
    };
    unsigned st = 0;
    // reconstruct row r (MUSCL on rows r-1, r, r+1; p2c and sound speed)
    auto recon = [&](int r, SSt<NM, R>& SL, SSt<NM, R>& SR) {
        SPrim<NM, R> Pm, Pc, Pp, PL, PR;
        load_sprim<NM, R>(lv, U, i, cj(r - 1), ck(r - 1), Pm);
        load_sprim<NM, R>(lv, U, i, cj(r), ck(r), Pc);
        load_sprim<NM, R>(lv, U, i, cj(r + 1), ck(r + 1), Pp);
        muscl_s<NM, R>(Pm, Pc, Pp, PL, PR);
        p2c_sound_s<NM, RW, R>(lv, PL, D, SL);
        p2c_sound_s<NM, RW, R>(lv, PR, D, SR);
    };
    auto emit = [&](int face, const SSt<NM, R>& A, const SSt<NM, R>& B) {
        R Fv[NF];
        wave5<NM, RW, R>(lv, A, B, D, Fv, st);
        if (!act) return;
        const long long fi = faceidx(face);
#pragma unroll
        for (int q = 0; q < NF; ++q) F.p[q * F.cs + fi] = Fv[q];
    };
    constexpr bool SLIDE = true;
    SPrim<NM, R> W0, W1;  // rows r-1 and r (sliding form)
    if (SLIDE) {
        load_sprim<NM, R>(lv, U, i, cj(rb - 1), ck(rb - 1), W0);
        load_sprim<NM, R>(lv, U, i, cj(rb), ck(rb), W1);
    }
    for (int r = rb; r <= re; ++r) {
        SSt<NM, R> SLs, SRs;
        if (SLIDE) {
            SPrim<NM, R> W2, PL, PR;
            load_sprim<NM, R>(lv, U, i, cj(r + 1), ck(r + 1), W2);
            muscl_s<NM, R>(W0, W1, W2, PL, PR);
            p2c_sound_s<NM, RW, R>(lv, PL, D, SLs);
            p2c_sound_s<NM, RW, R>(lv, PR, D, SRs);
            W0 = W1;
            W1 = W2;
        } else {
            recon(r, SLs, SRs);
        }
        if (r >= f0) {  // face r: left UR(r-1) (the image of UR(0) at the domain face), right UL(r)
            SSt<NM, R> A;
            if (r == 0) {
                A = SRs;
                if (lv.bc_lo[D]) flip_s<NM, R>(A, D);
            } else {
                sst_load<NM, R>(sh, lane, A);
            }
            emit(r, A, SLs);
        }
        sst_store<NM, R>(sh, lane, SRs);
    }
    if (f1 == nd + 1) {  // domain face nd: UR(nd-1) and the image of UL(nd-1)
        SSt<NM, R> SLs, SRs;
        recon(nd - 1, SLs, SRs);
        if (lv.bc_hi[D]) flip_s<NM, R>(SLs, D);
        emit(nd, SRs, SLs);
    }
    if (st) atomicOr(&clk->status, st);
}
template <int NM, int D, bool RW, typename R>
__global__ __launch_bounds__(256) void k_flux_lmarch_s(Level lv, MF U, FacesT<R> F, int seglen, Clock* clk) {
    flux_lmarch_s_body<NM, D, RW, R>(lv, U, F, seglen, clk);
}

// Per-cell solid state held by the update kernel.
template <int NM>
struct SCell {
    double al[NM], ar[NM], rk[NM], e[NM], are[NM], rho, ru[3], u[3], E, p, V[9], hj2;
};

template <bool NT>
__device__ __forceinline__ double ld(const double& src) {
    if constexpr (NT) return get_base(src);
    else return src;
}
// NT: non-temporal loads (get_base), for the update's input and, in stage 2,
// the step's starting state (256^3 plate: the update stages 0.3-0.8 % faster)
template <int NM, bool NT = false>
__device__ __forceinline__ void load_scell(const Level& lv, const MF& U, int i, int j, int k, SCell<NM>& C) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        C.al[m] = ld<NT>(U.at(lv.cs[ALPHA] + m, i, j, k));
        C.ar[m] = ld<NT>(U.at(lv.cs[ALPHARHO] + m, i, j, k));
        C.rk[m] = ld<NT>(U.at(lv.cs[RHO_K] + m, i, j, k));
        C.e[m] = lv.plastic ? ld<NT>(U.at(lv.cs[EPSILON] + m, i, j, k)) : 0.0;
        C.are[m] = lv.plastic ? ld<NT>(U.at(lv.cs[ALPHARHOEPSILON] + m, i, j, k)) : 0.0;
    }
    C.rho = ld<NT>(U.at(lv.cs[RHO], i, j, k));
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        C.ru[q] = ld<NT>(U.at(lv.cs[RHOU] + q, i, j, k));
        C.u[q] = ld<NT>(U.at(lv.cs[VELOCITY] + q, i, j, k));
    }
    C.E = ld<NT>(U.at(lv.cs[TOTAL_E], i, j, k));
    C.p = ld<NT>(U.at(lv.cs[PRESSURE], i, j, k));
#pragma unroll
    for (int q = 0; q < 9; ++q) C.V[q] = ld<NT>(U.at(lv.cs[V_TENSOR] + q, i, j, k));
    C.hj2 = ld<NT>(U.at(lv.cs[HJ2], i, j, k));
}
// NT: non-temporal stores (put_state), for the end-of-step state
template <int NM, bool NT = false>
__device__ __forceinline__ void store_scell(const Level& lv, const MF& U, int i, int j, int k, const SCell<NM>& C) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        put_state<NT>(U.at(lv.cs[ALPHA] + m, i, j, k), C.al[m]);
        put_state<NT>(U.at(lv.cs[ALPHARHO] + m, i, j, k), C.ar[m]);
        put_state<NT>(U.at(lv.cs[RHO_K] + m, i, j, k), C.rk[m]);
        if (lv.plastic) {
            put_state<NT>(U.at(lv.cs[EPSILON] + m, i, j, k), C.e[m]);
            put_state<NT>(U.at(lv.cs[ALPHARHOEPSILON] + m, i, j, k), C.are[m]);
        }
    }
    put_state<NT>(U.at(lv.cs[RHO], i, j, k), C.rho);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        put_state<NT>(U.at(lv.cs[RHOU] + q, i, j, k), C.ru[q]);
        put_state<NT>(U.at(lv.cs[VELOCITY] + q, i, j, k), C.u[q]);
    }
    put_state<NT>(U.at(lv.cs[TOTAL_E], i, j, k), C.E);
    put_state<NT>(U.at(lv.cs[PRESSURE], i, j, k), C.p);
#pragma unroll
    for (int q = 0; q < 9; ++q) put_state<NT>(U.at(lv.cs[V_TENSOR] + q, i, j, k), C.V[q]);
    put_state<NT>(U.at(lv.cs[HJ2], i, j, k), C.hj2);
}

// geometricSourceTerm (radial.cpp:5-31, flux.cpp:51-77) of a solid cell
// whose stored stress is that of its own state, as it is after the
// conservativeToPrimitive that ends plasticUpdate: sigma from p and the
// deviatoric Hencky strain dH (stressTensor, boxaccesscellarray.cpp:176-203).
// V rows 0 and 1 take +1/3 V u_x, row 2 -2/3 V u_x.
template <int NM, bool RW>
__device__ __forceinline__ void radial_solid(const Level& lv, double f, SCell<NM>& C, const double* dH) {
    double s0[3], s1[3], s2[3];
    stress_col<NM, RW>(lv, C.al, C.rk, C.p, dH, 0, s0);
    stress_col<NM, RW>(lv, C.al, C.rk, C.p, dH, 1, s1);
    stress_col<NM, RW>(lv, C.al, C.rk, C.p, dH, 2, s2);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        C.ar[m] += f * (C.ar[m] * C.u[0]);
        // a select, not a conditional store: the latter put C.are in scratch
        C.are[m] = lv.plastic ? C.are[m] + f * (C.are[m] * C.u[0]) : C.are[m];
    }
    radial_momentum_energy(f, C.u, s0, s1[0], s2[0], s2[2], C.ru, C.E);
#pragma unroll
    for (int q = 0; q < 6; ++q) C.V[q] += f * ((1.0 / 3.0) * C.V[q] * C.u[0]);
#pragma unroll
    for (int q = 6; q < 9; ++q) C.V[q] += f * (-(2.0 / 3.0) * C.V[q] * C.u[0]);
}
// plasticUpdate's conservativeToPrimitive leaves p and dH of the state; a
// stand-alone source recomputes dH from V (hencky does what getHenckyJ2 does).
template <int NM, bool RW>
__device__ __forceinline__ void radial_solid_c2p(const Level& lv, double f, SCell<NM>& C, double* dH, bool have_dH) {
    if (!have_dH) hencky(C.V, dH);
    radial_solid<NM, RW>(lv, f, C, dH);
    C.hj2 = c2p_s<NM, RW>(lv, C.al, C.ar, C.are, C.ru, C.E, C.V, C.rk, C.e, C.rho, C.u, C.p, dH);
}

// update (hllc.cpp:493-523, with the V rows of :511-514) in x, y, z order,
// then conservativeToPrimitive.  Stage 2 adds the RK average (hllc.cpp:705),
// plasticUpdate (main.cpp:154-157, plastic.cpp:227-244), the geometric source
// (main.cpp:159-162), cleanUpV (main.cpp:164-167), cleanUpAlpha
// (main.cpp:169-171), the NaN scan and the next getTimeStep.
template <int NM, int STAGE, bool RW, typename T>
__global__ __launch_bounds__(256) void k_update_s(Level lv, MF Uin, MF Ub, MF Uo, FacesT<T> F0, FacesT<T> F1, FacesT<T> F2,
                                                  Clock* clk, unsigned ops) {
    if (clk->done) return;
    constexpr int NC = 3 * NM + 13;
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = xcd_block(blockIdx.x, gridDim.x, lv.dim) * blockDim.x + threadIdx.x;
    const long long ncell = (long long)nx * ny * nz;
    double cand = __builtin_huge_val();
    unsigned st = 0;
    if (id < ncell) {
        const int i = (int)(id % nx);
        const int j = (int)((id / nx) % ny);
        const int k = (int)(id / ((long long)nx * ny));
        const double dt = clk->dt;
        SCell<NM> C;
        load_scell<NM, true>(lv, Uin, i, j, k, C);
        double a0[NM], V0[9], u0[3];
#pragma unroll
        for (int m = 0; m < NM; ++m) a0[m] = C.al[m];
#pragma unroll
        for (int q = 0; q < 9; ++q) V0[q] = C.V[q];
#pragma unroll
        for (int q = 0; q < 3; ++q) u0[q] = C.u[q];
        // unrolled over the three directions (a loop to lv.dim kept C.are in
        // scratch memory: the plastic-only update made its index dynamic)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d >= lv.dim) break;
            const FacesT<T>& Fd = d == 0 ? F0 : (d == 1 ? F1 : F2);
            const int i1 = i + (d == 0), j1 = j + (d == 1), k1 = k + (d == 2);
            const double dtdx = dt / lv.dx[d];
            const double s0 = (double)Fd.at(NC, i, j, k), s1 = (double)Fd.at(NC, i1, j1, k1);
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                C.al[m] += dtdx * ((double)Fd.at(m, i, j, k) - (double)Fd.at(m, i1, j1, k1) - a0[m] * (s0 - s1));
                C.ar[m] += dtdx * ((double)Fd.at(NM + m, i, j, k) - (double)Fd.at(NM + m, i1, j1, k1));
                if (lv.plastic) {
                    const double fl = (double)Fd.at(2 * NM + 13 + m, i, j, k), fh = (double)Fd.at(2 * NM + 13 + m, i1, j1, k1);
                    C.are[m] += dtdx * (fl - fh);
                }
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) C.ru[q] += dtdx * ((double)Fd.at(2 * NM + q, i, j, k) - (double)Fd.at(2 * NM + q, i1, j1, k1));
            C.E += dtdx * ((double)Fd.at(2 * NM + 3, i, j, k) - (double)Fd.at(2 * NM + 3, i1, j1, k1));
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int n = 2 * NM + 4 + r * 3 + c;
                    const double vs0 = (double)Fd.at(NC + 1 + c, i, j, k), vs1 = (double)Fd.at(NC + 1 + c, i1, j1, k1);
                    C.V[r * 3 + c] += dtdx * ((double)Fd.at(n, i, j, k) - (double)Fd.at(n, i1, j1, k1) - (2.0 / 3.0) * V0[r * 3 + c] * (s0 - s1) +
                                              u0[r] * (vs0 - vs1));
                }
        }
        double dH[9];
        // Stage 2 with the plastic relaxation: the conservativeToPrimitive that
        // ends it rewrites rho, u, p, HJ2 and dH, so this one forms only what
        // the relaxation reads (rho_k and epsilon, as c2p_s does).  The values
        // are those of the full sequence.
        const bool relax = STAGE == 2 && (ops & 1u) != 0;
        SCell<NM> B;
        if (STAGE == 2) load_scell<NM, true>(lv, Ub, i, j, k, B);  // in flight across the branch
        if (relax) {
            c2p_rk_e<NM>(lv, C.al, C.ar, C.are, C.rk, C.e);
            C.rho = C.p = C.hj2 = 0.0;  // rewritten after the relaxation (no loads of Uin's values)
        } else C.hj2 = c2p_s<NM, RW>(lv, C.al, C.ar, C.are, C.ru, C.E, C.V, C.rk, C.e, C.rho, C.u, C.p, dH);
        if (STAGE == 2) {
            // U1 = (U * 1/2) + (U2 * 1/2) over every component
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
            if (relax) {  // plastic relaxation, then conservativeToPrimitive
                st |= plastic_cell<NM, RW>(lv, C.al, C.ar, C.rk, C.e, C.are, C.V);
                C.hj2 = c2p_s<NM, RW>(lv, C.al, C.ar, C.are, C.ru, C.E, C.V, C.rk, C.e, C.rho, C.u, C.p, dH);
            }
            if (ops & 32u) radial_solid_c2p<NM, RW>(lv, radial_factor(lv, i, dt), C, dH, (ops & 1u) != 0);
            if (ops & 2u) cleanup_v<NM, T>(lv, C.al, C.V);  // FP32 in the mixed form
            if (ops & 4u) st |= cleanup_alpha<NM>(C.al);
            if (ops & 16u) {
                bool bad = false;
#pragma unroll
                for (int m = 0; m < NM; ++m)
                    bad |= __builtin_isnan(C.al[m]) || __builtin_isnan(C.ar[m]) || __builtin_isnan(C.rk[m]) ||
                           (lv.plastic && (__builtin_isnan(C.e[m]) || __builtin_isnan(C.are[m])));
#pragma unroll
                for (int q = 0; q < 3; ++q) bad |= __builtin_isnan(C.ru[q]) || __builtin_isnan(C.u[q]);
#pragma unroll
                for (int q = 0; q < 9; ++q) bad |= __builtin_isnan(C.V[q]);
                bad |= __builtin_isnan(C.E) || __builtin_isnan(C.p);
                if (bad) st |= 1u;
            }
            if (ops & 8u) {
                const double a = sound_s<NM, RW>(lv, C.al, C.ar, C.rk, C.rho, C.p, C.hj2);
                put_state<STAGE == 2>(Uo.at(lv.cs[SOUNDSPEED], i, j, k), a);
                #pragma unroll
                for (int d = 0; d < 3; ++d)  // unrolled: constant indices keep the state in registers
                    if (d < lv.dim) cand = dmin(cand, lv.dx[d] / (a + __builtin_fabs(C.u[d])));
            }
        }
        store_scell<NM, STAGE == 2>(lv, Uo, i, j, k, C);
    }
    if (STAGE == 2 && (ops & 8u)) {
        block_min_to(cand, dt_slot(lv, clk));
    }
    if (st) atomicOr(&clk->status, st);
}

template <int NM, bool RW>
__global__ __launch_bounds__(256) void k_sound_dt_s(Level lv, MF U, Clock* clk) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncell = (long long)nx * ny * nz;
    double cand = __builtin_huge_val();
    if (id < ncell) {
        const int i = (int)(id % nx);
        const int j = (int)((id / nx) % ny);
        const int k = (int)(id / ((long long)nx * ny));
        SCell<NM> C;
        load_scell<NM>(lv, U, i, j, k, C);
        const double a = sound_s<NM, RW>(lv, C.al, C.ar, C.rk, C.rho, C.p, C.hj2);
        U.at(lv.cs[SOUNDSPEED], i, j, k) = a;
        #pragma unroll
        for (int d = 0; d < 3; ++d)  // unrolled: constant indices keep the state in registers
            if (d < lv.dim) cand = dmin(cand, lv.dx[d] / (a + __builtin_fabs(C.u[d])));
    }
    block_min_to(cand, dt_slot(lv, clk));
}

// setInitialConditions for solid levels: thermal -> total pressure shift
// (initialise.cpp:104, with the HJ2 component as given) then
// primitiveToConservative (which computes HJ2 and devH).
template <int NM, bool RW>
__global__ __launch_bounds__(256) void k_init_p2c_s(Level lv, MF U, int shift) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)nx * ny * nz) return;
    const int i = (int)(id % nx);
    const int j = (int)((id / nx) % ny);
    const int k = (int)(id / ((long long)nx * ny));
    SPrim<NM> P;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        P.a[m] = U.at(lv.cs[ALPHA] + m, i, j, k);
        P.r[m] = U.at(lv.cs[RHO_K] + m, i, j, k);
        P.e[m] = lv.plastic ? U.at(lv.cs[EPSILON] + m, i, j, k) : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) P.u[q] = U.at(lv.cs[VELOCITY] + q, i, j, k);
    P.p = U.at(lv.cs[PRESSURE], i, j, k);
#pragma unroll
    for (int q = 0; q < 9; ++q) P.V[q] = U.at(lv.cs[V_TENSOR] + q, i, j, k);
    if (shift) {
        double ar[NM], G, Ent, Pnt;
#pragma unroll
        for (int m = 0; m < NM; ++m) ar[m] = P.a[m] * P.r[m];
        eos_sums_s<NM, RW>(lv, P.a, ar, P.r, U.at(lv.cs[HJ2], i, j, k), G, Ent, Pnt);
        P.p += Pnt / G;
        U.at(lv.cs[PRESSURE], i, j, k) = P.p;
    }
    SSt<NM> S;
    p2c_sound_s<NM, RW>(lv, P, 0, S);
    double dH[9];
    const double hj2 = hencky(P.V, dH);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        U.at(lv.cs[ALPHARHO] + m, i, j, k) = S.ar[m];
        if (lv.plastic) U.at(lv.cs[ALPHARHOEPSILON] + m, i, j, k) = S.are[m];
    }
    U.at(lv.cs[RHO], i, j, k) = S.rho;
#pragma unroll
    for (int q = 0; q < 3; ++q) U.at(lv.cs[RHOU] + q, i, j, k) = S.u[q] * S.rho;
    U.at(lv.cs[TOTAL_E], i, j, k) = S.E;
    U.at(lv.cs[HJ2], i, j, k) = hj2;
#pragma unroll
    for (int q = 0; q < 9; ++q) U.at(lv.cs[DEVH] + q, i, j, k) = dH[q];
}

// derived components of a solid state: devH, HJ2 from V and sigma
// (boxaccesscellarray.cpp:176-216, 326-380)
template <int NM, bool RW>
__global__ __launch_bounds__(256) void k_derive_s(Level lv, MF U) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)nx * ny * nz) return;
    const int i = (int)(id % nx);
    const int j = (int)((id / nx) % ny);
    const int k = (int)(id / ((long long)nx * ny));
    double V[9], dH[9], al[NM], rk[NM];
#pragma unroll
    for (int q = 0; q < 9; ++q) V[q] = U.at(lv.cs[V_TENSOR] + q, i, j, k);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        al[m] = U.at(lv.cs[ALPHA] + m, i, j, k);
        rk[m] = U.at(lv.cs[RHO_K] + m, i, j, k);
    }
    const double hj2 = hencky(V, dH);
    const double p = U.at(lv.cs[PRESSURE], i, j, k);
    U.at(lv.cs[HJ2], i, j, k) = hj2;
#pragma unroll
    for (int q = 0; q < 9; ++q) U.at(lv.cs[DEVH] + q, i, j, k) = dH[q];
    for (int d = 0; d < 3; ++d) {
        double sig[3];
        stress_col<NM, RW>(lv, al, rk, p, dH, d, sig);
        for (int r = 0; r < 3; ++r) U.at(lv.cs[SIGMA] + 3 * r + d, i, j, k) = sig[r];
    }
}

}  // inline namespace TJ_VIEW_NS
}  // namespace tj

namespace tj {
inline namespace TJ_VIEW_NS {

// conservativeToPrimitive over the valid cells of a solid level.
template <int NM, bool RW>
__global__ __launch_bounds__(256) void k_c2p_s(Level lv, MF U) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)nx * ny * nz) return;
    const int i = (int)(id % nx);
    const int j = (int)((id / nx) % ny);
    const int k = (int)(id / ((long long)nx * ny));
    SCell<NM> C;
    load_scell<NM>(lv, U, i, j, k, C);
    double dH[9];
    C.hj2 = c2p_s<NM, RW>(lv, C.al, C.ar, C.are, C.ru, C.E, C.V, C.rk, C.e, C.rho, C.u, C.p, dH);
    store_scell<NM>(lv, U, i, j, k, C);
}

// The post-step chain of main.cpp:154-171 on a state, stand-alone
// (tj_post_step): plastic + c2p, the geometric source + c2p (with the clock's
// dt), cleanUpV, cleanUpAlpha, NaN scan, sound speed and the CFL minimum, each
// when its bit is set.
template <int NM, bool RW>
__global__ __launch_bounds__(256) void k_post_s(Level lv, MF U, Clock* clk, unsigned ops) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncell = (long long)nx * ny * nz;
    double cand = __builtin_huge_val();
    unsigned st = 0;
    if (id < ncell) {
        const int i = (int)(id % nx);
        const int j = (int)((id / nx) % ny);
        const int k = (int)(id / ((long long)nx * ny));
        SCell<NM> C;
        load_scell<NM>(lv, U, i, j, k, C);
        double dH[9];
        if (ops & 1u) {
            st |= plastic_cell<NM, RW>(lv, C.al, C.ar, C.rk, C.e, C.are, C.V);
            C.hj2 = c2p_s<NM, RW>(lv, C.al, C.ar, C.are, C.ru, C.E, C.V, C.rk, C.e, C.rho, C.u, C.p, dH);
        }
        if (ops & 32u) radial_solid_c2p<NM, RW>(lv, radial_factor(lv, i, clk->dt), C, dH, (ops & 1u) != 0);
        if (ops & 2u) cleanup_v<NM>(lv, C.al, C.V);
        if (ops & 4u) st |= cleanup_alpha<NM>(C.al);
        if (ops & 16u) {
            bool bad = false;
#pragma unroll
            for (int m = 0; m < NM; ++m)
                bad |= __builtin_isnan(C.al[m]) || __builtin_isnan(C.ar[m]) || __builtin_isnan(C.rk[m]) ||
                       (lv.plastic && (__builtin_isnan(C.e[m]) || __builtin_isnan(C.are[m])));
#pragma unroll
            for (int q = 0; q < 3; ++q) bad |= __builtin_isnan(C.ru[q]) || __builtin_isnan(C.u[q]);
#pragma unroll
            for (int q = 0; q < 9; ++q) bad |= __builtin_isnan(C.V[q]);
            bad |= __builtin_isnan(C.E) || __builtin_isnan(C.p);
            if (bad) st |= 1u;
        }
        if (ops & 8u) {
            const double a = sound_s<NM, RW>(lv, C.al, C.ar, C.rk, C.rho, C.p, C.hj2);
            U.at(lv.cs[SOUNDSPEED], i, j, k) = a;
            #pragma unroll
            for (int d = 0; d < 3; ++d)  // unrolled: constant indices keep the state in registers
                if (d < lv.dim) cand = dmin(cand, lv.dx[d] / (a + __builtin_fabs(C.u[d])));
        }
        store_scell<NM>(lv, U, i, j, k, C);
    }
    if (ops & 8u) {
        block_min_to(cand, dt_slot(lv, clk));
    }
    if (st) atomicOr(&clk->status, st);
}

}  // inline namespace TJ_VIEW_NS
}  // namespace tj
