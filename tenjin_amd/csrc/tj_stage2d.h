// Stage kernels of two-dimensional fluid levels: one RK stage of
// HLLCadvance (hllc.cpp:579-695) in two launches, with no face fluxes in HBM.
//
//   x kernel  lanes are cells along x (a wave covers one row segment).  MUSCL
//             (+ THINC/BVD) reconstruction, primitiveToConservative and sound
//             speed of both faces, HLLC at each lane's lower face; then the x
//             part of update (hllc.cpp:493-523) for each cell whose two faces
//             the wave holds, the upper face arriving by a lane shuffle.  Out:
//             the partially updated conserved state Ux (8 FP64 components).
//   y kernel  a block owns TC columns x a band of rows (16 x 12 by
//             default).  The primitives of the tile (the band with its halo
//             rows, and the alphas of the two neighbouring columns for
//             Youngs' stencil) are staged in LDS with coalesced
//             loads; each lane segment then reconstructs one column with lanes along y
//             (neighbours by shuffle, as in the x kernel) and computes the
//             y faces; the face fluxes go to LDS, and a lanes-along-x phase
//             adds the y part of update to Ux and completes the cell:
//             conserved-to-primitive, and in stage 2 the RK average and the
//             post-step chain (finish_cell), with coalesced state I/O.
//
// The y kernel reads the THINC history the x kernel of the same stage wrote
// (star alphas, URStar values), exactly as the reference's y sweep reads what
// its x sweep left in ULStar/URStar; the two launches order them.  Every
// value is computed by the same per-cell functions, in the same order, as the
// flux-plane path (k_flux_x, k_flux_march, k_update), so both forms are
// bit-identical.
#pragma once
#include "tj_kernels.h"

namespace tj {
inline namespace TJ_VIEW_NS {

// Partially updated conserved state between the two kernels of a stage:
// component-major planes of FP64 over the level's cells, in the order
// alpha (NM), alpha rho (NM), rho u (3), E.
struct UxPlanes {
    double* p;
    long long cs;  // component stride (= cells)
    __device__ __forceinline__ double& at(int q, long long cell) const { return p[(long long)q * cs + cell]; }
};

// ------------------------------------------------------------------ x kernel
// Lane l holds cell c = s0 - H + l; the wave computes faces s0 .. s0 + CPW and
// updates cells s0 .. s0 + CPW - 1 (CPW = 60 with THINC, 62 without).
template <bool TH>
struct XGeom {
    static constexpr int H = TH ? 2 : 1;
    static constexpr int CPW = TH ? 60 : 62;
};

// With FIN = 1 or 2 (one-dimensional levels) the x part is the whole stage:
// the wave completes its cells itself (finish_cell: c2p; in stage 2 the RK
// average with Ub and the post-step chain) into Uo instead of writing Ux, and
// in stage 2 puts its CFL candidate into the partial slots.
struct XFin {
    MF Ub, Uo;
    unsigned ops;
};
template <int NM, bool TH, typename R, int FIN = 0>
__device__ __forceinline__ void xstage_body(const Level& lv, const MF& U, const UxPlanes& Ux, const Hist& h, Clock* clk,
                                            long long wave, const XFin& fin = XFin{}) {
    constexpr int NF = 2 * NM + 5;
    if (clk->done) return;
    constexpr int H = XGeom<TH>::H, CPW = XGeom<TH>::CPW;
    const int lane = threadIdx.x & 63;
    const int nx = lv.n[0], ny = lv.n[1];
    const int nseg = (nx + CPW - 1) / CPW;
    if (wave >= (long long)nseg * ny) return;  // uniform per wave
    const int seg = (int)(wave % nseg);
    const int j = (int)(wave / nseg), k = 0;
    const int s0 = seg * CPW;
    const int c = s0 - H + lane;
    bool fl;
    const int s = bcmap(c, nx, lv.bc_lo[0], lv.bc_hi[0], fl);
    const int lofs = s - c;

    Prim<NM, R> Pm, Pc, Pp, PL, PR;
    load_prim<NM, R>(lv, U, s - 1, j, k, Pm);
    load_prim<NM, R>(lv, U, s, j, k, Pc);
    load_prim<NM, R>(lv, U, s + 1, j, k, Pp);
    muscl<NM, R>(Pm, Pc, Pp, PL, PR);

    unsigned st = 0;
    const long long cell_s = (long long)j * nx + s;
    // the lane owns cell c's history and update, and face c's history
    const bool own = lane >= H && c >= s0 && c < s0 + CPW && c < nx;
    if (TH) {
        ARec<NM, R> A;
        arec<NM, R>(lv, U, s, j, k, 0, (R)lv.yn[0], (R)lv.yn[1], A);
        const uint8_t eflag = lv.edge[0][s];
        const bool blo = eflag & 1, bhi = eflag & 2;
        int srcm = lane + lofs - 1, srcp = lane + lofs + 1;
        srcm = srcm < 0 ? 0 : (srcm > 63 ? 63 : srcm);
        srcp = srcp < 0 ? 0 : (srcp > 63 ? 63 : srcp);
        const int mixm = __shfl(A.mixed, srcm, 64);
        const int mixp = __shfl(A.mixed, srcp, 64);
        const int mixedFlag = (A.mixed >> (NM - 1)) & 1;
        R URl[NM], TRl[NM], ULr[NM], TLr[NM], thl[NM], thr[NM];
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            const R uraM = __shfl(A.ura[m], srcm, 64);
            const R thrM = __shfl(A.thr[m], srcm, 64);
            const R ulaP = __shfl(A.ula[m], srcp, 64);
            const R thlP = __shfl(A.thl[m], srcp, 64);
            if (mixedFlag) {
                if (blo) {
                    URl[m] = load_alpha<R>(lv, U, m, s - 1, j, k);
                    TRl[m] = R(0);
                } else {
                    URl[m] = uraM;
                    TRl[m] = ((mixm >> m) & 1) ? thrM : (R)h.urstar[m * h.pstride + cell_s - 1];
                }
                if (bhi) {
                    ULr[m] = load_alpha<R>(lv, U, m, s + 1, j, k);
                    TLr[m] = (R)h.edge_in[m * h.estride + (long long)j * (nx + 1) + s + 1];
                } else {
                    ULr[m] = ulaP;
                    TLr[m] = ((mixp >> m) & 1) ? thlP : (R)h.star_in[m * h.pstride + cell_s + 1];
                }
                if ((A.mixed >> m) & 1) {
                    thl[m] = A.thl[m];
                    thr[m] = A.thr[m];
                } else {
                    thl[m] = (R)h.star_in[m * h.pstride + cell_s];
                    thr[m] = (R)h.urstar[m * h.pstride + cell_s];
                }
            }
        }
        if (mixedFlag) bvd<NM, R>(1, URl, TRl, ULr, TLr, thl, thr, PL.a, PR.a);
        st |= cleanup_alpha<NM, R>(PL.a);
        st |= cleanup_alpha<NM, R>(PR.a);
        if (own) {
#pragma unroll
            for (int m = 0; m < NM; ++m)
                if ((A.mixed >> m) & 1) h.urstar[m * h.pstride + cell_s] = (double)A.thr[m];
        }
    }

    FSt<NM, R> SLs, SRs;
    p2c_sound<NM, R>(lv, PL, SLs);
    p2c_sound<NM, R>(lv, PR, SRs);

    // left state of face c: UR(c-1) from the neighbouring lane
    FSt<NM, R> Lf;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        Lf.al[m] = lane_dn(SRs.al[m]);
        Lf.ar[m] = lane_dn(SRs.ar[m]);
    }
    Lf.rho = lane_dn(SRs.rho);
#pragma unroll
    for (int q = 0; q < 3; ++q) Lf.u[q] = lane_dn(SRs.u[q]);
    Lf.E = lane_dn(SRs.E);
    Lf.p = lane_dn(SRs.p);
    Lf.a = lane_dn(SRs.a);

    // face c: lanes H .. H + CPW (faces s0 .. s0 + CPW), up to the domain's top face
    const bool face = lane >= H && lane <= H + CPW && c >= 0 && c <= nx;
    R Fv[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q) Fv[q] = R(0);
    if (face) {
        if (c == 0) {
            Lf = SRs;
            if (lv.bc_lo[0]) flip_state<NM, R>(Lf, 0);
        }
        FSt<NM, R> Rf = SLs;
        if (c == nx && lv.bc_hi[0]) flip_state<NM, R>(Rf, 0);
        R sa[NM];
        hllc<NM, R>(Lf, Rf, 0, Fv, sa);
        if (TH && (c < s0 + CPW || c == nx)) {
            if (c < nx) {
#pragma unroll
                for (int m = 0; m < NM; ++m) h.star_out[m * h.pstride + (long long)j * nx + c] = (double)sa[m];
            }
            if (c == nx || (c > 0 && (lv.edge[0][c - 1] & 2))) {
#pragma unroll
                for (int m = 0; m < NM; ++m) h.edge_out[m * h.estride + (long long)j * (nx + 1) + c] = (double)sa[m];
            }
        }
    }
    // upper face of cell c from lane + 1
    R Fh[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q) Fh[q] = lane_up(Fv[q]);
    double cand = __builtin_huge_val();
    if (own) {
        const long long cell = (long long)j * nx + c;
        double al[NM], ar[NM], ru[3], E, a0[NM];
        load_cons<NM>(lv, U, c, j, k, al, ar, ru, E);
#pragma unroll
        for (int m = 0; m < NM; ++m) a0[m] = al[m];
        accum_dir<NM, R>(al, ar, ru, E, a0, clk->dt / lv.dx[0], Fv, Fh);
        if constexpr (FIN == 0) {
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                Ux.at(m, cell) = al[m];
                Ux.at(NM + m, cell) = ar[m];
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) Ux.at(2 * NM + q, cell) = ru[q];
            Ux.at(2 * NM + 3, cell) = E;
        } else {
            finish_cell<NM, FIN>(lv, fin.Ub, fin.Uo, c, j, k, al, ar, ru, E, clk->dt, fin.ops, cand, st);
        }
    }
    if constexpr (FIN == 2) {
        if (fin.ops & 8u) wave_min_to(cand, dt_slot(lv, clk));
    }
    if (st) atomicOr(&clk->status, st);
}

template <int NM, bool TH, typename R = double>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TH && sizeof(R) == 8 ? 4 : 1)))
void k_xstage(Level lv, MF U, UxPlanes Ux, Hist h, Clock* clk) {
    const long long nb = gridDim.x;
    const long long b = xcd_block(blockIdx.x, nb, 2);
    xstage_body<NM, TH, R>(lv, U, Ux, h, clk, b * (blockDim.x >> 6) + wave_in_block());
}

// One RK stage of a one-dimensional fluid level in one launch (FIN = STAGE).
// With TAIL the launch also ends the step: every wave, done or not, takes a
// ticket after its work, and the wave holding the last one advances the clock
// as k_clock_advance does and re-arms the ticket (no wave waits on another).
template <int NM, bool TH, int STAGE, bool TAIL, typename R = double>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TH && sizeof(R) == 8 ? 4 : 1)))
void k_xstage1d(Level lv, MF U, MF Ub, MF Uo, Hist h, Clock* clk, unsigned ops, unsigned* ticket) {
    const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + wave_in_block();
    xstage_body<NM, TH, R, STAGE>(lv, U, UxPlanes{}, h, clk, wave, XFin{Ub, Uo, ops});
    if constexpr (TAIL) {
        if ((threadIdx.x & 63) == 0) {
            __threadfence();
            const unsigned waves = gridDim.x * (blockDim.x >> 6);
            if (atomicAdd(ticket, 1u) == waves - 1) {
                __threadfence();
                *ticket = 0u;
                if (!clk->done) {
                    if (atomicOr(&clk->status, 0u) & ST_STOP) {
                        clk->done = 1;
                    } else {
                        clk->t += clk->dt;
                        clk->n += 1;
                    }
                }
                __threadfence();
            }
        }
    }
}

// ------------------------------------------------------------------ y kernel
// A block owns columns i0 .. i0 + TC - 1 and a segment of rows.  Lanes run
// along y in segments of LPS lanes (64: one column per wave; 32 or 16: two
// or four columns per wave).  Shorter segments recompute more halo rows
// (THINC: 12 of 16 lanes complete a row, against 28 of 32) but make smaller
// blocks with less LDS, so that more blocks share a compute unit and one's
// loads after its barrier overlap the others' arithmetic: lane l of a
// segment holds row r = s0 - H + l, the segment computes faces s0 .. s0 + CPW
// and completes cells s0 .. s0 + CPW - 1.
template <int NM, bool TH, int TC, int LPS, typename R>
struct YTile {
    static constexpr int H = TH ? 2 : 1;
    static constexpr int CPW = LPS - 2 * H;    // rows completed per tile
    static constexpr int NP = 2 * NM + 4;      // staged primitives: alpha, rho_k, u (3), p
    static constexpr int ROWS = LPS + 2;       // rows s0 - H - 1 .. s0 - H + LPS
    static constexpr int NF = 2 * NM + 5;      // flux components per face
    static constexpr int FR = CPW + 1;         // faces per column
    static constexpr int NT = TC * LPS;        // threads per block
    static constexpr int NW = NT / 64;         // waves per block
    static constexpr int PRIMS = NP * TC * ROWS + (TH ? NM * 2 * ROWS : 0);
    static constexpr int FLUX = NF * TC * FR;
    static constexpr int WORDS = PRIMS > FLUX ? PRIMS : FLUX;
};

// The block minimum of the positive CFL candidates (ordered bits), one
// atomic per block.
template <int NW>
__device__ __forceinline__ void block_min_to_n(double cand, unsigned long long* dst, unsigned long long* red) {
    const unsigned long long b = wave_min_bits(cand);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long mn = red[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) mn = red[w] < mn ? red[w] : mn;
        atomicMin(dst, mn);
    }
}

template <int NM, bool TH, int TC, int LPS, int STAGE, typename R>
__device__ __forceinline__ void ystage_body(const Level& lv, const MF& U, const UxPlanes& Ux, const MF& Ub,
                                            const MF& Uo, const Hist& h, Clock* clk, unsigned ops, long long tile) {
    using T = YTile<NM, TH, TC, LPS, R>;
    constexpr int H = T::H, CPW = T::CPW, NP = T::NP, ROWS = T::ROWS, NF = T::NF, FR = T::FR, NT = T::NT;
    __shared__ R lds[T::WORDS];
    __shared__ unsigned long long red[T::NW];
    if (clk->done) return;  // uniform: the clock does not change during the launch
    const int nx = lv.n[0], ny = lv.n[1];
    const int ntc = (nx + TC - 1) / TC;
    const int ct = (int)(tile % ntc);
    const int seg = (int)(tile / ntc);
    const int i0 = ct * TC, s0 = seg * CPW;
    const int base = s0 - H - 1;  // global row of LDS row 0
    const int tid = threadIdx.x;
    // LDS views: primitive q of tile column col at LDS row t; alphas of the
    // neighbouring columns (side 0: i0 - 1, side 1: i0 + TC)
    auto P = [&](int q, int col, int t) -> R& { return lds[(q * TC + col) * ROWS + t]; };
    auto AX = [&](int m, int side, int t) -> R& { return lds[NP * TC * ROWS + (m * 2 + side) * ROWS + t]; };
    auto FL = [&](int q, int col, int f) -> R& { return lds[(q * TC + col) * FR + f]; };

    // ---- stage the tile's primitives (load_prim of every LDS row and column)
    for (int e = tid; e < ROWS * TC; e += NT) {
        const int col = e % TC, t = e / TC;
        Prim<NM, R> Pq;
        load_prim<NM, R>(lv, U, i0 + col, base + t, 0, Pq);
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            P(m, col, t) = Pq.a[m];
            P(NM + m, col, t) = Pq.r[m];
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) P(2 * NM + q, col, t) = Pq.u[q];
        P(2 * NM + 3, col, t) = Pq.p;
    }
    if (TH) {
        for (int e = tid; e < ROWS * 2; e += NT) {
            const int side = e & 1, t = e >> 1;
#pragma unroll
            for (int m = 0; m < NM; ++m) AX(m, side, t) = load_alpha<R>(lv, U, m, side ? i0 + TC : i0 - 1, base + t, 0);
        }
    }
    __syncthreads();

    // ---- one column per segment of LPS lanes, lanes along y
    const int lane = tid & 63;
    const int sl = lane % LPS;                 // lane within the segment
    const int lo = lane - sl, hi = lo + LPS - 1;  // the segment's lanes
    const int col = (tid >> 6) * (64 / LPS) + lane / LPS;
    const int i = i0 + col;
    const bool colok = i < nx;
    const int r = s0 - H + sl;
    bool fl;
    const int s = bcmap(r, ny, lv.bc_lo[1], lv.bc_hi[1], fl);
    const int lofs = s - r;
    auto lrow = [&](int g) {  // LDS row of global row g (g - 1 .. g + 1 of a lane's own row stay in range)
        const int t = g - base;
        return t < 0 ? 0 : (t > ROWS - 1 ? ROWS - 1 : t);
    };
    auto prim_at = [&](int g, Prim<NM, R>& Pq) {
        const int t = lrow(g);
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            Pq.a[m] = P(m, col, t);
            Pq.r[m] = P(NM + m, col, t);
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) Pq.u[q] = P(2 * NM + q, col, t);
        Pq.p = P(2 * NM + 3, col, t);
    };
    // load_alpha of (i + di, g): the tile's own column or a neighbouring one
    auto alpha_at = [&](int m, int di, int g) -> R {
        const int t = lrow(g);
        const int cc = col + di;
        if (cc < 0) return AX(m, 0, t);
        if (cc >= TC) return AX(m, 1, t);
        return P(m, cc, t);
    };
    // rows s-1, s, s+1 as load_prim maps them (the staged row g holds the
    // boundary image of g, so g = s +- 1 reads what load_prim(s +- 1) reads)
    Prim<NM, R> Pm, Pc, Pp, PL, PR;
    prim_at(s - 1, Pm);
    prim_at(s, Pc);
    prim_at(s + 1, Pp);
    muscl<NM, R>(Pm, Pc, Pp, PL, PR);

    unsigned st = 0;
    const long long cell_s = (long long)s * nx + i;
    const bool own = colok && sl >= H && r >= s0 && r < s0 + CPW && r < ny;
    if (TH) {
        // arec (thinc.cpp:54-88, 130-186) of cell (i, s) along y from the staged alphas
        ARec<NM, R> A;
        A.mixed = 0;
        const R ynx = (R)lv.yn[0], yny = (R)lv.yn[1];
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            const R aL = alpha_at(m, 0, s - 1);
            const R aC = alpha_at(m, 0, s);
            const R aR = alpha_at(m, 0, s + 1);
            muscl1(aL, aC, aR, A.ula[m], A.ura[m]);
            const R mn = tmin(aL, aR);
            const R mx = tmax(aL, aR) - mn;
            const R theta = tsgn(aR - aL);
            const R eps = R(1E-20);
            const R Cq = (mx < eps) ? R(0) : tdiv(aC - mn + eps, mx + eps);
            const R dl = R(1E-4);
            const bool mixed = (Cq > dl) && (Cq < (R(0.999999) - dl)) && ((aR - aC) * (aC - aL) >= R(0));
            A.thl[m] = R(0);
            A.thr[m] = R(0);
            if (mixed) {
                A.mixed |= 1 << m;
                auto a = [&](int ddi, int ddj) { return alpha_at(m, ddi, s + ddj); };
                const R nxv = ynx * (a(1, 1) + R(2) * a(1, 0) + a(1, -1) - a(-1, 1) - R(2) * a(-1, 0) - a(-1, -1));
                const R nyv = yny * (a(1, 1) + R(2) * a(0, 1) + a(-1, 1) - a(-1, -1) - R(2) * a(0, -1) - a(1, -1));
                const R norm = tsqrt(nxv * nxv + nyv * nyv);
                const R nvc = (norm < R(1E-20)) ? R(0) : tabs(tdiv(nyv, norm));
                const R beta = nvc * R(lv.beta0) + R(0.01);
                const R coshB = tcosh(beta), tanhB = ttanh(beta);
                const R B = texp(theta * beta * ((R(2) * Cq) - R(1)));
                const R Ac = tdiv(tdiv(B, coshB) - R(1), tanhB);
                A.thr[m] = mn + tdiv(mx, R(2)) * (R(1) + theta * tdiv(tanhB + Ac, R(1) + Ac * tanhB));
                A.thl[m] = mn + tdiv(mx, R(2)) * (R(1) + theta * Ac);
            }
        }
        const uint8_t eflag = lv.edge[1][s];
        const bool blo = eflag & 1, bhi = eflag & 2;
        int srcm = lane + lofs - 1, srcp = lane + lofs + 1;
        srcm = srcm < lo ? lo : (srcm > hi ? hi : srcm);
        srcp = srcp < lo ? lo : (srcp > hi ? hi : srcp);
        const int mixm = __shfl(A.mixed, srcm, 64);
        const int mixp = __shfl(A.mixed, srcp, 64);
        const int mixedFlag = (A.mixed >> (NM - 1)) & 1;
        R URl[NM], TRl[NM], ULr[NM], TLr[NM], thl[NM], thr[NM];
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            const R uraM = __shfl(A.ura[m], srcm, 64);
            const R thrM = __shfl(A.thr[m], srcm, 64);
            const R ulaP = __shfl(A.ula[m], srcp, 64);
            const R thlP = __shfl(A.thl[m], srcp, 64);
            if (mixedFlag && colok) {
                if (blo) {
                    URl[m] = alpha_at(m, 0, s - 1);
                    TRl[m] = R(0);
                } else {
                    URl[m] = uraM;
                    TRl[m] = ((mixm >> m) & 1) ? thrM : (R)h.urstar[m * h.pstride + cell_s - nx];
                }
                if (bhi) {
                    ULr[m] = alpha_at(m, 0, s + 1);
                    TLr[m] = (R)h.edge_in[m * h.estride + (long long)(s + 1) * nx + i];
                } else {
                    ULr[m] = ulaP;
                    TLr[m] = ((mixp >> m) & 1) ? thlP : (R)h.star_in[m * h.pstride + cell_s + nx];
                }
                if ((A.mixed >> m) & 1) {
                    thl[m] = A.thl[m];
                    thr[m] = A.thr[m];
                } else {
                    thl[m] = (R)h.star_in[m * h.pstride + cell_s];
                    thr[m] = (R)h.urstar[m * h.pstride + cell_s];
                }
            }
        }
        if (mixedFlag && colok) bvd<NM, R>(1, URl, TRl, ULr, TLr, thl, thr, PL.a, PR.a);
        st |= cleanup_alpha<NM, R>(PL.a);
        st |= cleanup_alpha<NM, R>(PR.a);
        if (own) {
#pragma unroll
            for (int m = 0; m < NM; ++m)
                if ((A.mixed >> m) & 1) h.urstar[m * h.pstride + cell_s] = (double)A.thr[m];
        }
    }

    FSt<NM, R> SLs, SRs;
    p2c_sound<NM, R>(lv, PL, SLs);
    p2c_sound<NM, R>(lv, PR, SRs);
    // UR of the row below from lane - 1; the segment's first lane takes the
    // previous segment's value, but it is a halo lane (sl = 0 < H) whose Lf is
    // never read
    FSt<NM, R> Lf;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        Lf.al[m] = lane_dn(SRs.al[m]);
        Lf.ar[m] = lane_dn(SRs.ar[m]);
    }
    Lf.rho = lane_dn(SRs.rho);
#pragma unroll
    for (int q = 0; q < 3; ++q) Lf.u[q] = lane_dn(SRs.u[q]);
    Lf.E = lane_dn(SRs.E);
    Lf.p = lane_dn(SRs.p);
    Lf.a = lane_dn(SRs.a);

    const bool face = colok && sl >= H && sl <= H + CPW && r >= 0 && r <= ny;
    R Fv[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q) Fv[q] = R(0);
    if (face) {
        if (r == 0) {
            Lf = SRs;
            if (lv.bc_lo[1]) flip_state<NM, R>(Lf, 1);
        }
        FSt<NM, R> Rf = SLs;
        if (r == ny && lv.bc_hi[1]) flip_state<NM, R>(Rf, 1);
        R sa[NM];
        hllc<NM, R>(Lf, Rf, 1, Fv, sa);
        if (TH && (r < s0 + CPW || r == ny)) {
            if (r < ny) {
#pragma unroll
                for (int m = 0; m < NM; ++m) h.star_out[m * h.pstride + (long long)r * nx + i] = (double)sa[m];
            }
            if (r == ny || (r > 0 && (lv.edge[1][r - 1] & 2))) {
#pragma unroll
                for (int m = 0; m < NM; ++m) h.edge_out[m * h.estride + (long long)r * nx + i] = (double)sa[m];
            }
        }
    }
    __syncthreads();  // every wave has read its primitives: the flux staging may overwrite them
    if (sl >= H && sl <= H + CPW) {
#pragma unroll
        for (int q = 0; q < NF; ++q) FL(q, col, sl - H) = Fv[q];
    }
    __syncthreads();

    // ---- y part of update and completion, lanes along x
    double cand = __builtin_huge_val();
    const double dt = clk->dt;
    for (int e = tid; e < CPW * TC; e += NT) {
        const int cc = e % TC, f = e / TC;
        const int ii = i0 + cc, rr = s0 + f;
        if (ii < nx && rr < ny) {
            const long long cell = (long long)rr * nx + ii;
            double al[NM], ar[NM], ru[3], E, a0[NM];
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                al[m] = Ux.at(m, cell);
                ar[m] = Ux.at(NM + m, cell);
                a0[m] = get_base(U.at(lv.cs[ALPHA] + m, ii, rr, 0));
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) ru[q] = Ux.at(2 * NM + q, cell);
            E = Ux.at(2 * NM + 3, cell);
            R Flo[NF], Fhi[NF];
#pragma unroll
            for (int q = 0; q < NF; ++q) {
                Flo[q] = FL(q, cc, f);
                Fhi[q] = FL(q, cc, f + 1);
            }
            accum_dir<NM, R>(al, ar, ru, E, a0, dt / lv.dx[1], Flo, Fhi);
            finish_cell<NM, STAGE>(lv, Ub, Uo, ii, rr, 0, al, ar, ru, E, dt, ops, cand, st);
        }
    }
    if (STAGE == 2 && (ops & 8u)) block_min_to_n<T::NW>(cand, dt_slot(lv, clk), red);
    if (st) atomicOr(&clk->status, st);
}

// The mixed form's y kernels take six waves per SIMD (71 VGPRs in stage 2,
// from 95, no spill): stage 2 0.211 -> 0.202 ms on the 2048 x 1024 level.
// The FP64 form at five spills about 100 bytes per lane and runs 1.6x slower.
template <int NM, bool TH, int TC, int LPS, int STAGE, typename R = double>
__global__ __launch_bounds__(TC * LPS) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 6 : 1))) void k_ystage(Level lv, MF U, UxPlanes Ux, MF Ub, MF Uo, Hist h, Clock* clk,
                                                     unsigned ops) {
    // tiles renumbered per XCD (xcd_block): neighbouring column tiles, which
    // share cache lines of rows whose pitch is not a multiple of the line,
    // and the row bands above and below, which share halo rows, run in one L2
    ystage_body<NM, TH, TC, LPS, STAGE, R>(lv, U, Ux, Ub, Uo, h, clk, ops, xcd_block(blockIdx.x, gridDim.x, 2));
}

}  // inline namespace TJ_VIEW_NS
}  // namespace tj
