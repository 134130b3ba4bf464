// Batched ensemble of independent fluid levels (1D and 2D, uneven sizes),
// each with its own time step, advanced together: every pass of a step is
// one launch over all problems of a dimension group, the problem of a wave
// (or block) found in a prefix table, and a run of steps can be captured as
// one hipGraph.  The per-problem arithmetic is the single-level kernels'
// (the same device bodies), so each problem's result is bit-identical to
// running it alone.  Included at the end of tj_lib.hip.
#pragma once

namespace tj {
inline namespace TJ_VIEW_NS {

// Everything one problem's passes need, in device memory.
struct BProb {
    Level lv;
    MF S, T;
    Faces F[2];
    FacesT<float> Ff[2];
    UxPlanes ux;  // partial state of the stage kernels (2D problems)
    int tile;     // y-tile shape of the stage kernels (BATCH_TILES)
    double* urstar;
    double* star[2];
    double* estar[2][2];
    long long estride[2];
    Clock* clk;
    int seglen;
};

// A dimension group: its problems and the prefix tables of work units.
struct BTab {
    const BProb* p;
    const long long* start;  // np + 1 entries: first unit of each problem
    int np;
};

// The problem records and prefix tables are read-only while a kernel runs.
// Reading them through the constant address space, at a wave-uniform index,
// lets the backend use scalar loads and keep or re-load the fields in SGPRs
// across the kernel's own global stores.  Through a plain pointer at an index
// derived from threadIdx.x, every field of the problem's Level was a per-lane
// vector load (measured: 188 vector loads per x-stage wave against 75 on a
// single level, whose Level is a kernel argument).
template <typename T>
__device__ __forceinline__ const T& ro_at(const T* base, long long i) {
    using CT = const __attribute__((address_space(4))) T;
    return *(const T*)((CT*)base + i);
}
// the wave's index in the grid, known to the compiler to be wave-uniform
__device__ __forceinline__ long long bwave() {
    return (long long)blockIdx.x * (blockDim.x >> 6) + wave_in_block();
}

// largest q with start[q] <= u (start[0] = 0, u < start[np])
__device__ __forceinline__ int bfind(const long long* start, int np, long long u) {
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ro_at(start, mid) <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
// the number of work units of a table (its last prefix entry)
__device__ __forceinline__ long long bunits(const BTab& t) { return ro_at(t.start, t.np); }

__device__ __forceinline__ Hist bhist(const BProb& P, int d, int q, int e) {
    Hist h{};
    if (!P.lv.thinc) return h;
    h.urstar = P.urstar;
    h.star_in = P.star[q];
    h.star_out = P.star[q ^ 1];
    h.edge_in = P.estar[d][e];
    h.edge_out = P.estar[d][e ^ 1];
    h.pstride = P.lv.ncells;
    h.estride = P.estride[d];
    return h;
}

// y-tile shapes of the stage kernels a 2D problem can take (columns x rows
// per lane segment); the autotuner picks one per size bucket
constexpr int NBATCH_TILES = 3;
constexpr int BATCH_TILE_TC[NBATCH_TILES] = {16, 8, 16};
constexpr int BATCH_TILE_LPS[NBATCH_TILES] = {32, 32, 16};

template <typename R>
__device__ __forceinline__ const FacesT<R>& bfaces(const BProb& P, int d);
template <>
__device__ __forceinline__ const FacesT<double>& bfaces<double>(const BProb& P, int d) { return P.F[d]; }
template <>
__device__ __forceinline__ const FacesT<float>& bfaces<float>(const BProb& P, int d) { return P.Ff[d]; }

// x pass over all problems of a group; which = 0 reads S, 1 reads T
template <bool TH, typename R>
__global__ __launch_bounds__(256) void k_b_flux_x(BTab t, int which, int q, int e) {
    const long long gw = bwave();
    if (gw >= bunits(t)) return;
    const int p = bfind(t.start, t.np, gw);
    const BProb& P = ro_at(t.p, p);
    flux_x_body<2, TH, R>(P.lv, which ? P.T : P.S, bfaces<R>(P, 0), bhist(P, 0, q, e), P.clk, gw - ro_at(t.start, p));
}

// y pass (marching) over all problems of a 2D group
template <bool TH, typename R>
__device__ __forceinline__ void b_flux_y(const BTab& t, int which, int q, int e) {
    const long long gw = bwave();
    if (gw >= bunits(t)) return;
    const int p = bfind(t.start, t.np, gw);
    const BProb& P = ro_at(t.p, p);
    const MF z{};
    const FacesT<R> zf{};
    flux_march_body<2, TH, 1, false, 1, R>(P.lv, which ? P.T : P.S, bfaces<R>(P, 1), bhist(P, 1, q, e), P.seglen,
                                           P.clk, z, z, zf, zf, 0u, gw - ro_at(t.start, p));
}
template <bool TH, typename R>
__global__ __launch_bounds__(256) void k_b_flux_y(BTab t, int which, int q, int e) {
    b_flux_y<TH, R>(t, which, q, e);
}
// two waves per SIMD for the FP64 THINC pass, as k_flux_march_w2
template <bool TH, typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_b_flux_y_w2(BTab t, int which, int q,
                                                                                              int e) {
    b_flux_y<TH, R>(t, which, q, e);
}

// stage 1: T = c2p(S + dt L(S)); stage 2: S = 1/2 S + 1/2 c2p(T + dt L(T)) + post-step ops
template <int STAGE, typename T>
__global__ __launch_bounds__(256) void k_b_update(BTab t, unsigned ops) {
    const long long gb = blockIdx.x;
    if (gb >= bunits(t)) return;  // uniform per block
    const int p = bfind(t.start, t.np, gb);
    const BProb& P = ro_at(t.p, p);
    const FacesT<T>& F0 = bfaces<T>(P, 0);
    const FacesT<T>& F1 = P.lv.dim > 1 ? bfaces<T>(P, 1) : F0;
    if (STAGE == 1) update_body<2, 1, T>(P.lv, P.S, P.S, P.T, F0, F1, F0, P.clk, ops, gb - ro_at(t.start, p));
    else update_body<2, 2, T>(P.lv, P.T, P.S, P.S, F0, F1, F0, P.clk, ops, gb - ro_at(t.start, p));
}

// stage kernels of the 2D problems (tj_stage2d.h): x part over all problems,
// then the y part over the problems of one tile shape
template <bool TH, typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TH && sizeof(R) == 8 ? 4 : 1)))
void k_b_xstage(BTab t, int which, int q, int e) {
    // blocks renumbered per XCD as on a single 2D level (xcd_block): each
    // XCD's L2 then holds neighbouring rows of one problem
    const long long gw = xcd_block(blockIdx.x, gridDim.x, 2) * (blockDim.x >> 6) +
                         wave_in_block();
    if (gw >= bunits(t)) return;
    const int p = bfind(t.start, t.np, gw);
    const BProb& P = ro_at(t.p, p);
    xstage_body<2, TH, R>(P.lv, which ? P.T : P.S, P.ux, bhist(P, 0, q, e), P.clk, gw - ro_at(t.start, p));
}
// one RK stage of every 1D problem in one launch (k_xstage1d of a single
// level): reconstruction, fluxes and the completion of each wave's cells
template <bool TH, int STAGE, typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TH && sizeof(R) == 8 ? 4 : 1)))
void k_b_xstage1d(BTab t, int q, int e, unsigned ops) {
    const long long gw = bwave();
    if (gw >= bunits(t)) return;
    const int p = bfind(t.start, t.np, gw);
    const BProb& P = ro_at(t.p, p);
    const long long w = gw - ro_at(t.start, p);
    if (STAGE == 1) xstage_body<2, TH, R, 1>(P.lv, P.S, UxPlanes{}, bhist(P, 0, q, e), P.clk, w, XFin{P.S, P.T, ops});
    else xstage_body<2, TH, R, 2>(P.lv, P.T, UxPlanes{}, bhist(P, 0, q, e), P.clk, w, XFin{P.S, P.S, ops});
}
template <bool TH, int TC, int LPS, int STAGE, typename R>
__global__ __launch_bounds__(TC * LPS) void k_b_ystage(BTab t, int q, int e, unsigned ops) {
    const long long gb = xcd_block(blockIdx.x, gridDim.x, 2);  // as k_ystage
    if (gb >= bunits(t)) return;
    const int p = bfind(t.start, t.np, gb);
    const BProb& P = ro_at(t.p, p);
    if (STAGE == 1)
        ystage_body<2, TH, TC, LPS, 1, R>(P.lv, P.S, P.ux, P.S, P.T, bhist(P, 1, q, e), P.clk, ops, gb - ro_at(t.start, p));
    else
        ystage_body<2, TH, TC, LPS, 2, R>(P.lv, P.T, P.ux, P.S, P.S, bhist(P, 1, q, e), P.clk, ops, gb - ro_at(t.start, p));
}

__global__ __launch_bounds__(256) void k_b_sound_dt(BTab t) {
    const long long gb = blockIdx.x;
    if (gb >= bunits(t)) return;
    const int p = bfind(t.start, t.np, gb);
    const BProb& P = ro_at(t.p, p);
    sound_dt_body<2>(P.lv, P.S, P.clk, gb - ro_at(t.start, p));
}

__global__ __launch_bounds__(256) void k_b_init_p2c(BTab t, int shift) {
    const long long gb = blockIdx.x;
    if (gb >= bunits(t)) return;
    const int p = bfind(t.start, t.np, gb);
    const BProb& P = ro_at(t.p, p);
    init_p2c_body<2>(P.lv, P.S, shift, gb - ro_at(t.start, p));
}

// the clock kernels of the single level, one 64-thread block per problem.
// Each problem has its own DT_PARTS partial slots of the CFL minimum (its
// Level::dtpart), as a single level does: the blocks of one problem's stage
// kernels spread their atomics over them (on one word, the 1365 blocks of a
// 511^2 problem's stage-2 y kernel serialised to twice the kernel's time).
// k_b_clock_dt folds and re-arms them before it takes the minimum.
__global__ __launch_bounds__(64) void k_b_clock_dt(const BProb* P, int np, int driver) {
    const int p = blockIdx.x;
    if (p >= np) return;
    Clock* clk = P[p].clk;
    unsigned long long* part = P[p].lv.dtpart;
    unsigned long long b = DT_NONE;
    for (int q = (int)threadIdx.x; q < DT_PARTS; q += 64) {
        const unsigned long long v = part[q];
        b = v < b ? v : b;
        part[q] = DT_NONE;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(b, off, 64);
        b = o < b ? o : b;
    }
    if (threadIdx.x != 0) return;
    if (b < clk->dtmin_bits) clk->dtmin_bits = b;
    if (driver) {
        if (clk->done) return;
        if (!(clk->t < clk->final_time)) {
            clk->done = 1;
            return;
        }
    }
    const double mn = __builtin_bit_cast(double, clk->dtmin_bits);
    double dt = P[p].lv.cfl * mn;
    if (driver) {
        if (clk->t + dt > clk->final_time) dt = clk->final_time - clk->t;
        if (clk->n < 5) dt *= 0.2;
    }
    clk->dt = dt;
    clk->dtmin_bits = 0x7FEFFFFFFFFFFFFFull;
}
__global__ void k_b_clock_advance(const BProb* P, int np) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    Clock* clk = P[p].clk;
    if (clk->done) return;
    if (clk->status & ST_STOP) {
        clk->done = 1;
        return;
    }
    clk->t += clk->dt;
    clk->n += 1;
}
// zeroes the THINC history planes of every problem in one launch: block
// row y is problem y, its blocks stride over that problem's planes
__global__ __launch_bounds__(256) void k_b_zero_hist(const BProb* P) {
    const BProb& Q = P[blockIdx.y];
    if (!Q.lv.thinc) return;
    const long long pl = Q.lv.ncells * Q.lv.nmat;
    const long long id0 = (long long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long long)gridDim.x * blockDim.x;
    for (long long i = id0; i < pl; i += stride) {
        Q.urstar[i] = 0.0;
        Q.star[0][i] = 0.0;
        Q.star[1][i] = 0.0;
    }
    for (int a = 0; a < Q.lv.dim; ++a) {
        const long long n = Q.estride[a] * Q.lv.nmat;
        for (long long i = id0; i < n; i += stride) {
            Q.estar[a][0][i] = 0.0;
            Q.estar[a][1][i] = 0.0;
        }
    }
}
__global__ __launch_bounds__(64) void k_b_clock_begin(const BProb* P, int np, double final_time) {
    const int p = blockIdx.x;
    if (p >= np) return;
    for (int q = (int)threadIdx.x; q < DT_PARTS; q += 64) P[p].lv.dtpart[q] = DT_NONE;
    if (threadIdx.x != 0) return;
    Clock* clk = P[p].clk;
    clk->t = 0.0;
    clk->dt = 0.0;
    clk->final_time = final_time;
    clk->n = 0;
    clk->done = 0;
    clk->status = 0;
    clk->dtmin_bits = 0x7FEFFFFFFFFFFFFFull;
}

}  // inline namespace TJ_VIEW_NS
}  // namespace tj

struct tj_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int np = 0;
    int thinc = 0;
    std::vector<tj_level_desc> descs;
    std::vector<tj::BProb> h;           // host copy of the problem records
    std::vector<int> order;             // record index -> caller's problem index
    tj::BProb* d_p = nullptr;
    tj::Clock* d_clk = nullptr;
    std::vector<void*> allocs;
    // dimension groups: records [g0, g1) are 1D, [g1, np) are 2D
    int g1 = 0;
    long long* d_start = nullptr;       // prefix tables, see build_tables
    unsigned long long* d_dtpart = nullptr;  // DT_PARTS partial CFL slots per problem (in allocs)
    // [group][kind]: x waves, y waves, cell blocks of the flux-plane passes;
    // x-stage waves and y tiles of each shape of the stage kernels
    static constexpr int NKIND = 4 + tj::NBATCH_TILES;
    long long tot[2][NKIND] = {};
    long long off[2][NKIND] = {};
    // 2D problems: the stage kernels (default) or the flux-plane passes
    bool stage2d = true;
    long long pass[2] = {0, 0}, dpass[2][2] = {{0, 0}, {0, 0}};
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    long long g_npass[2] = {0, 0}, g_ndpass[2][2] = {{0, 0}, {0, 0}};
    long long g_pass0[2] = {0, 0}, g_dpass0[2][2] = {{0, 0}, {0, 0}};
};

namespace {

#define BHIPCHK(b, call)                                                            \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            (b)->err = std::string(#call) + ": " + hipGetErrorString(e_);           \
            return TJ_EHIP;                                                         \
        }                                                                           \
    } while (0)

int balloc(tj_batch* b, void** p, size_t bytes) {
    BHIPCHK(b, hipMalloc(p, bytes));
    b->allocs.push_back(*p);
    return TJ_OK;
}

// rows a y tile completes (YTile::CPW) for a shape
int tile_rows(int thinc, int s) { return tj::BATCH_TILE_LPS[s] - 2 * (thinc ? 2 : 1); }

// Units per problem: x-pass waves, y-pass waves, 256-cell blocks; x-stage
// waves, y tiles of shape 0 .. NBATCH_TILES - 1 (only the problem's own shape).
void units(const tj::BProb& P, int thinc, bool stage2d, long long u[tj_batch::NKIND]) {
    const tj::Level& lv = P.lv;
    for (int k = 0; k < tj_batch::NKIND; ++k) u[k] = 0;
    const int per = thinc ? 61 : 63;
    u[0] = (long long)((lv.n[0] + 1 + per - 1) / per) * lv.n[1] * lv.n[2];
    if (stage2d) {  // x-stage waves (2D: x part of the stage kernels; 1D: the whole stage)
        const int cpw = thinc ? tj::XGeom<true>::CPW : tj::XGeom<false>::CPW;
        u[3] = (long long)((lv.n[0] + cpw - 1) / cpw) * lv.n[1];
    }
    if (lv.dim > 1) {
        const long long nstrip = (lv.n[0] + 63) / 64;
        const long long nseg = (lv.n[1] + 1 + P.seglen - 1) / P.seglen;
        u[1] = nstrip * nseg * lv.n[2];
        if (stage2d) {
            const int tc = tj::BATCH_TILE_TC[P.tile], rows = tile_rows(thinc, P.tile);
            u[4 + P.tile] = (long long)((lv.n[0] + tc - 1) / tc) * ((lv.n[1] + rows - 1) / rows);
        }
    }
    u[2] = (lv.ncells + 255) / 256;
}

// Prefix tables [group][kind] -> (np_g + 1) entries in one device array.
int build_tables(tj_batch* b) {
    std::vector<long long> all;
    for (int g = 0; g < 2; ++g) {
        const int lo = g == 0 ? 0 : b->g1, hi = g == 0 ? b->g1 : b->np;
        for (int kind = 0; kind < tj_batch::NKIND; ++kind) {
            b->off[g][kind] = (long long)all.size();
            long long acc = 0;
            all.push_back(0);
            for (int r = lo; r < hi; ++r) {
                long long u[tj_batch::NKIND];
                units(b->h[r], b->thinc, b->stage2d, u);
                acc += u[kind];
                all.push_back(acc);
            }
            b->tot[g][kind] = acc;
        }
    }
    if (b->d_start) {
        (void)hipFree(b->d_start);
        b->d_start = nullptr;
    }
    BHIPCHK(b, hipMalloc(&b->d_start, sizeof(long long) * all.size()));
    BHIPCHK(b, hipMemcpy(b->d_start, all.data(), sizeof(long long) * all.size(), hipMemcpyHostToDevice));
    return TJ_OK;
}

tj::BTab tab(const tj_batch* b, int g, int kind) {
    tj::BTab t;
    t.p = b->d_p + (g == 0 ? 0 : b->g1);
    t.np = g == 0 ? b->g1 : b->np - b->g1;
    t.start = b->d_start + b->off[g][kind];
    return t;
}

void batch_clear_graph(tj_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

// The stage kernels of group g (2D): x stage, then y stage per tile shape.
template <bool TH, typename R>
void launch_b_stage2d(tj_batch* b, int g, int stage, unsigned ops) {
    using namespace tj;
    const int which = stage == 1 ? 0 : 1;
    {
        const int q = (int)(b->pass[g] & 1), e = (int)(b->dpass[g][0] & 1);
        hipLaunchKernelGGL((k_b_xstage<TH, R>), dim3(grid1(b->tot[g][3], 4)), dim3(256), 0, b->stream, tab(b, g, 3), which,
                           q, e);
        if (b->thinc) {
            b->pass[g] += 1;
            b->dpass[g][0] += 1;
        }
    }
    const int q = (int)(b->pass[g] & 1), e = (int)(b->dpass[g][1] & 1);
    const unsigned o = stage == 1 ? 0u : ops;
    for (int sh = 0; sh < NBATCH_TILES; ++sh) {
        const long long n = b->tot[g][4 + sh];
        if (!n) continue;
        const BTab t = tab(b, g, 4 + sh);
#define TJ_B_YSTAGE(TC, LPS)                                                                                       \
    do {                                                                                                           \
        if (stage == 1)                                                                                            \
            hipLaunchKernelGGL((k_b_ystage<TH, TC, LPS, 1, R>), dim3((unsigned)n), dim3(TC * LPS), 0, b->stream, t, q, \
                               e, o);                                                                              \
        else                                                                                                       \
            hipLaunchKernelGGL((k_b_ystage<TH, TC, LPS, 2, R>), dim3((unsigned)n), dim3(TC * LPS), 0, b->stream, t, q, \
                               e, o);                                                                              \
    } while (0)
        static_assert(BATCH_TILE_TC[0] == 16 && BATCH_TILE_LPS[0] == 32 && BATCH_TILE_TC[1] == 8 &&
                          BATCH_TILE_LPS[1] == 32 && BATCH_TILE_TC[2] == 16 && BATCH_TILE_LPS[2] == 16,
                      "tile shapes of the launch switch");
        if (sh == 0) TJ_B_YSTAGE(16, 32);
        else if (sh == 1) TJ_B_YSTAGE(8, 32);
        else TJ_B_YSTAGE(16, 16);
#undef TJ_B_YSTAGE
    }
    if (b->thinc) {
        b->pass[g] += 1;
        b->dpass[g][1] += 1;
    }
}

// One loop body for every problem: driver dt, two RK stages (the post-step
// chain fused into stage 2), clock advance.
int batch_step(tj_batch* b, int prec) {
    using namespace tj;
    const bool mixed = prec == TJ_PREC_MIXED;
    const unsigned ops = TJ_OP_CLEAN_ALPHA | TJ_OP_DT | TJ_OP_NANCHECK;
    const unsigned gp = grid1(b->np, 64);
    hipLaunchKernelGGL(k_b_clock_dt, dim3((unsigned)b->np), dim3(64), 0, b->stream, b->d_p, b->np, 1);
    for (int stage = 1; stage <= 2; ++stage) {
        const int which = stage == 1 ? 0 : 1;
        for (int g = 0; g < 2; ++g) {
            if (b->tot[g][2] == 0) continue;
            if (g == 1 && b->stage2d) {
                if (b->thinc && mixed) launch_b_stage2d<true, float>(b, g, stage, ops);
                else if (b->thinc) launch_b_stage2d<true, double>(b, g, stage, ops);
                else if (mixed) launch_b_stage2d<false, float>(b, g, stage, ops);
                else launch_b_stage2d<false, double>(b, g, stage, ops);
                continue;
            }
            if (g == 0 && b->stage2d) {  // 1D problems: one launch per stage, no flux planes
                const int q = (int)(b->pass[g] & 1), e = (int)(b->dpass[g][0] & 1);
                const unsigned gx = grid1(b->tot[g][3], 4);
                const BTab t = tab(b, g, 3);
                const unsigned o = stage == 1 ? 0u : ops;
#define TJ_B_X1D(TH, R)                                                                                            \
    do {                                                                                                           \
        if (stage == 1) hipLaunchKernelGGL((k_b_xstage1d<TH, 1, R>), dim3(gx), dim3(256), 0, b->stream, t, q, e, o); \
        else hipLaunchKernelGGL((k_b_xstage1d<TH, 2, R>), dim3(gx), dim3(256), 0, b->stream, t, q, e, o);           \
    } while (0)
                if (b->thinc && mixed) TJ_B_X1D(true, float);
                else if (b->thinc) TJ_B_X1D(true, double);
                else if (mixed) TJ_B_X1D(false, float);
                else TJ_B_X1D(false, double);
#undef TJ_B_X1D
                if (b->thinc) {
                    b->pass[g] += 1;
                    b->dpass[g][0] += 1;
                }
                continue;
            }
            // x pass
            {
                const int q = (int)(b->pass[g] & 1), e = (int)(b->dpass[g][0] & 1);
                const unsigned gx = grid1(b->tot[g][0], 4);
                const BTab t = tab(b, g, 0);
                if (b->thinc && mixed) hipLaunchKernelGGL((k_b_flux_x<true, float>), dim3(gx), dim3(256), 0, b->stream, t, which, q, e);
                else if (b->thinc) hipLaunchKernelGGL((k_b_flux_x<true, double>), dim3(gx), dim3(256), 0, b->stream, t, which, q, e);
                else if (mixed) hipLaunchKernelGGL((k_b_flux_x<false, float>), dim3(gx), dim3(256), 0, b->stream, t, which, q, e);
                else hipLaunchKernelGGL((k_b_flux_x<false, double>), dim3(gx), dim3(256), 0, b->stream, t, which, q, e);
                if (b->thinc) {
                    b->pass[g] += 1;
                    b->dpass[g][0] += 1;
                }
            }
            if (g == 1) {  // y pass
                const int q = (int)(b->pass[g] & 1), e = (int)(b->dpass[g][1] & 1);
                const unsigned gy = grid1(b->tot[g][1], 4);
                const BTab t = tab(b, g, 1);
                if (b->thinc && mixed) hipLaunchKernelGGL((k_b_flux_y<true, float>), dim3(gy), dim3(256), 0, b->stream, t, which, q, e);
                else if (b->thinc) hipLaunchKernelGGL((k_b_flux_y_w2<true, double>), dim3(gy), dim3(256), 0, b->stream, t, which, q, e);
                else if (mixed) hipLaunchKernelGGL((k_b_flux_y<false, float>), dim3(gy), dim3(256), 0, b->stream, t, which, q, e);
                else hipLaunchKernelGGL((k_b_flux_y<false, double>), dim3(gy), dim3(256), 0, b->stream, t, which, q, e);
                if (b->thinc) {
                    b->pass[g] += 1;
                    b->dpass[g][1] += 1;
                }
            }
            const unsigned gc = (unsigned)b->tot[g][2];
            const BTab t = tab(b, g, 2);
            if (stage == 1) {
                if (mixed) hipLaunchKernelGGL((k_b_update<1, float>), dim3(gc), dim3(256), 0, b->stream, t, 0u);
                else hipLaunchKernelGGL((k_b_update<1, double>), dim3(gc), dim3(256), 0, b->stream, t, 0u);
            } else {
                if (mixed) hipLaunchKernelGGL((k_b_update<2, float>), dim3(gc), dim3(256), 0, b->stream, t, ops);
                else hipLaunchKernelGGL((k_b_update<2, double>), dim3(gc), dim3(256), 0, b->stream, t, ops);
            }
        }
    }
    hipLaunchKernelGGL(k_b_clock_advance, dim3(gp), dim3(64), 0, b->stream, b->d_p, b->np);
    BHIPCHK(b, hipGetLastError());
    return TJ_OK;
}

}  // namespace

extern "C" {

int tj_batch_create(tj_batch** out, int device, int nprob, const tj_level_desc* descs, const tj_mf* S, const tj_mf* T) {
    using namespace tj;
    if (!out || nprob < 1 || !descs || !S || !T) return TJ_EINVAL;
    *out = nullptr;
    if (nprob > TJ_BATCH_MAX_PROBLEMS) return TJ_EINVAL;  // one launch row per problem (grid y)
    tj_batch* b = new tj_batch();
    b->device = device;
    auto fail = [&](int r) {
        tj_batch_destroy(b);
        return r;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess) return fail(TJ_EHIP);
    b->np = nprob;
    b->thinc = descs[0].THINC;
    // 1D problems first, then 2D (stable within each group)
    for (int g = 1; g <= 2; ++g)
        for (int p = 0; p < nprob; ++p)
            if (descs[p].dim == g) b->order.push_back(p);
    if ((int)b->order.size() != nprob) {
        b->err = "batch: problems must be 1D or 2D";
        *out = b;
        return TJ_EUNSUPPORTED;
    }
    b->h.resize(nprob);
    b->descs.resize(nprob);
    b->g1 = 0;
    for (int r = 0; r < nprob; ++r)
        if (descs[b->order[r]].dim == 1) b->g1 = r + 1;
    int rc = TJ_OK;
    if ((rc = balloc(b, (void**)&b->d_clk, sizeof(Clock) * nprob))) return *out = b, rc;
    if ((rc = balloc(b, (void**)&b->d_dtpart, sizeof(unsigned long long) * DT_PARTS * nprob))) return *out = b, rc;
    for (int r = 0; r < nprob; ++r) {
        const int p = b->order[r];
        const tj_level_desc& d = descs[p];
        b->descs[r] = d;
        bool mix = false;
        for (int m = 0; m < d.nmat && m < TJ_MAX_MATERIALS; ++m) mix |= d.mat[m].mixture != 0;
        if (d.SOLID || d.THINC != b->thinc || d.RADIAL || mix) {
            b->err = "batch: fluid problems with one THINC setting, no geometric source and no mixture";
            *out = b;
            return TJ_EUNSUPPORTED;
        }
        BProb& P = b->h[r];
        P = BProb{};
        std::vector<uint8_t> edges[3];
        if ((rc = make_level(b->err, &d, P.lv, edges))) return *out = b, rc;
        P.lv.dtpart = b->d_dtpart + (long long)r * DT_PARTS;
        for (int a = 0; a < 3; ++a) {
            void* e = nullptr;
            if ((rc = balloc(b, &e, edges[a].size()))) return *out = b, rc;
            if (hipMemcpy(e, edges[a].data(), edges[a].size(), hipMemcpyHostToDevice) != hipSuccess)
                return *out = b, TJ_EHIP;
            P.lv.edge[a] = (const uint8_t*)e;
        }
        int cs[NVAR];
        const int nc = build_layout(d, cs);
        if (S[p].fabs || T[p].fabs) {
            b->err = "batch: state views must be level arrays (tj_fabset views are for single levels)";
            *out = b;
            return TJ_EUNSUPPORTED;
        }
        if (S[p].ncomp < nc || T[p].ncomp < nc) {
            b->err = "batch: state view with too few components";
            *out = b;
            return TJ_EINVAL;
        }
        P.S = view(&S[p]);
        P.T = view(&T[p]);
        const int nfc = 2 * d.nmat + 5;
        for (int a = 0; a < d.dim; ++a) {
            Faces& F = P.F[a];
            F.nx = P.lv.n[0] + (a == 0);
            F.ny = P.lv.n[1] + (a == 1);
            F.nz = P.lv.n[2] + (a == 2);
            F.cs = (long long)F.nx * F.ny * F.nz;
            if ((rc = balloc(b, (void**)&F.p, sizeof(double) * F.cs * nfc))) return *out = b, rc;
            FacesT<float>& G = P.Ff[a];
            G.nx = F.nx;
            G.ny = F.ny;
            G.nz = F.nz;
            G.cs = F.cs;
            if ((rc = balloc(b, (void**)&G.p, sizeof(float) * G.cs * nfc))) return *out = b, rc;
            P.estride[a] = F.cs;
        }
        if (d.THINC) {
            const long long pl = P.lv.ncells * d.nmat;
            if ((rc = balloc(b, (void**)&P.urstar, sizeof(double) * pl))) return *out = b, rc;
            for (int q = 0; q < 2; ++q)
                if ((rc = balloc(b, (void**)&P.star[q], sizeof(double) * pl))) return *out = b, rc;
            for (int a = 0; a < d.dim; ++a)
                for (int q = 0; q < 2; ++q)
                    if ((rc = balloc(b, (void**)&P.estar[a][q], sizeof(double) * P.estride[a] * d.nmat)))
                        return *out = b, rc;
        }
        if (d.dim == 2) {  // the stage kernels' partial state: alpha, alpha rho, rho u (3), E
            P.ux.cs = P.lv.ncells;
            if ((rc = balloc(b, (void**)&P.ux.p, sizeof(double) * P.ux.cs * (2 * d.nmat + 4))))
                return *out = b, rc;
        }
        P.tile = 0;
        P.clk = b->d_clk + r;
        P.seglen = 16;
    }
    if ((rc = balloc(b, (void**)&b->d_p, sizeof(BProb) * nprob))) return *out = b, rc;
    for (int r = 0; r < nprob; ++r)  // the partial CFL slots start empty (begin_run re-arms them)
        hipLaunchKernelGGL(k_dtpart_reset, dim3(1), dim3(64), 0, b->stream, b->d_dtpart + (long long)r * DT_PARTS);
    *out = b;
    return tj_batch_reset_history(b);
}

void tj_batch_destroy(tj_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_clear_graph(b);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->d_start) (void)hipFree(b->d_start);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* tj_batch_last_error(const tj_batch* b) { return b ? b->err.c_str() : "null batch"; }

int tj_batch_reset_history(tj_batch* b) {
    using namespace tj;
    if (!b || !b->d_p) return TJ_EINVAL;
    BHIPCHK(b, hipSetDevice(b->device));
    for (int g = 0; g < 2; ++g) {
        b->pass[g] = 0;
        b->dpass[g][0] = b->dpass[g][1] = 0;
    }
    BHIPCHK(b, hipMemcpyAsync(b->d_p, b->h.data(), sizeof(BProb) * b->np, hipMemcpyHostToDevice, b->stream));
    // the history planes of all problems in one launch (the records are on the device now)
    if (b->thinc) {
        hipLaunchKernelGGL(k_b_zero_hist, dim3(64, (unsigned)b->np), dim3(256), 0, b->stream, b->d_p);
        BHIPCHK(b, hipGetLastError());
    }
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return build_tables(b);
}

int tj_batch_set_seglen(tj_batch* b, int prob, int seglen) {
    if (!b || prob < 0 || prob >= b->np || seglen < 1) return TJ_EINVAL;
    for (int r = 0; r < b->np; ++r)
        if (b->order[r] == prob) b->h[r].seglen = seglen;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    BHIPCHK(b, hipMemcpy(b->d_p, b->h.data(), sizeof(tj::BProb) * b->np, hipMemcpyHostToDevice));
    batch_clear_graph(b);
    return build_tables(b);
}

int tj_batch_init_from_prims(tj_batch* b, int shift_pressure) {
    using namespace tj;
    if (!b) return TJ_EINVAL;
    for (int g = 0; g < 2; ++g)
        if (b->tot[g][2])
            hipLaunchKernelGGL(k_b_init_p2c, dim3((unsigned)b->tot[g][2]), dim3(256), 0, b->stream, tab(b, g, 2),
                               shift_pressure);
    BHIPCHK(b, hipGetLastError());
    return TJ_OK;
}

int tj_batch_begin_run(tj_batch* b, double final_time) {
    using namespace tj;
    if (!b) return TJ_EINVAL;
    hipLaunchKernelGGL(k_b_clock_begin, dim3((unsigned)b->np), dim3(64), 0, b->stream, b->d_p, b->np, final_time);
    for (int g = 0; g < 2; ++g)
        if (b->tot[g][2])
            hipLaunchKernelGGL(k_b_sound_dt, dim3((unsigned)b->tot[g][2]), dim3(256), 0, b->stream, tab(b, g, 2));
    BHIPCHK(b, hipGetLastError());
    return TJ_OK;
}

int tj_batch_advance(tj_batch* b, int nsteps, int prec) {
    if (!b) return TJ_EINVAL;
    if (prec != TJ_PREC_FP64 && prec != TJ_PREC_MIXED) {
        b->err = "unknown precision form";
        return TJ_EINVAL;
    }
    int r;
    for (int q = 0; q < nsteps; ++q)
        if ((r = batch_step(b, prec))) return r;
    return TJ_OK;
}

int tj_batch_graph_capture(tj_batch* b, int nsteps, int prec) {
    if (!b || nsteps < 1) return TJ_EINVAL;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    batch_clear_graph(b);
    long long p0[2], d0[2][2];
    for (int g = 0; g < 2; ++g) {
        p0[g] = b->pass[g];
        d0[g][0] = b->dpass[g][0];
        d0[g][1] = b->dpass[g][1];
    }
    BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rr = tj_batch_advance(b, nsteps, prec);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    for (int q = 0; q < 2; ++q) {
        b->g_pass0[q] = p0[q];
        b->g_npass[q] = b->pass[q] - p0[q];
        b->pass[q] = p0[q];
        for (int a = 0; a < 2; ++a) {
            b->g_dpass0[q][a] = d0[q][a];
            b->g_ndpass[q][a] = b->dpass[q][a] - d0[q][a];
            b->dpass[q][a] = d0[q][a];
        }
    }
    if (rr) {
        if (e == hipSuccess && g) (void)hipGraphDestroy(g);
        return rr;
    }
    BHIPCHK(b, e);
    b->graph = g;
    BHIPCHK(b, hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    return TJ_OK;
}

int tj_batch_graph_launch(tj_batch* b) {
    if (!b) return TJ_EINVAL;
    if (!b->gexec) {
        b->err = "batch graph_launch: nothing captured";
        return TJ_EINVAL;
    }
    bool same = true;
    for (int g = 0; g < 2; ++g) {
        same = same && ((b->pass[g] - b->g_pass0[g]) & 1) == 0;
        for (int a = 0; a < 2; ++a) same = same && ((b->dpass[g][a] - b->g_dpass0[g][a]) & 1) == 0;
    }
    if (b->thinc && !same) {
        b->err = "batch graph_launch: THINC history phase differs from the captured one";
        return TJ_EINVAL;
    }
    BHIPCHK(b, hipGraphLaunch(b->gexec, b->stream));
    for (int g = 0; g < 2; ++g) {
        b->pass[g] += b->g_npass[g];
        for (int a = 0; a < 2; ++a) b->dpass[g][a] += b->g_ndpass[g][a];
    }
    return TJ_OK;
}

int tj_batch_set_tile(tj_batch* b, int prob, int tile) {
    if (!b || prob < 0 || prob >= b->np || tile < 0 || tile >= tj::NBATCH_TILES) return TJ_EINVAL;
    for (int r = 0; r < b->np; ++r)
        if (b->order[r] == prob) b->h[r].tile = tile;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    BHIPCHK(b, hipMemcpy(b->d_p, b->h.data(), sizeof(tj::BProb) * b->np, hipMemcpyHostToDevice));
    batch_clear_graph(b);
    return build_tables(b);
}

int tj_batch_set_stage2d(tj_batch* b, int on) {
    if (!b) return TJ_EINVAL;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    b->stage2d = on != 0;
    batch_clear_graph(b);
    return build_tables(b);
}

int tj_batch_clocks(tj_batch* b, double* t, int* n, uint32_t* status) {
    if (!b) return TJ_EINVAL;
    std::vector<tj::Clock> h(b->np);
    BHIPCHK(b, hipMemcpyAsync(h.data(), b->d_clk, sizeof(tj::Clock) * b->np, hipMemcpyDeviceToHost, b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    for (int r = 0; r < b->np; ++r) {
        const int p = b->order[r];
        if (t) t[p] = h[r].t;
        if (n) n[p] = h[r].n;
        if (status) status[p] = h[r].status;
    }
    return TJ_OK;
}

int tj_batch_synchronize(tj_batch* b) {
    if (!b) return TJ_EINVAL;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return TJ_OK;
}

}  // extern "C"
