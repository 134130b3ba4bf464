// Kernels of the fluid path: directional flux passes, the fused update
// (update + c2p [+ RK average + cleanUpAlpha + NaN scan + CFL min]) and the
// time-step words.
#pragma once
#include "tj_fluid.h"

namespace tj {

// Device-resident clock of the time loop (main.cpp:121-143).
struct Clock {
    double t, dt, final_time;
    int n, done;
    unsigned long long dtmin_bits;  // min over cells of dx/(a+|u|), as ordered bits of a positive double
    unsigned int status;
};
// Status bits on which the loop stops at the step that raised them (t and n
// stay): the NaN stop of main.cpp:183-187 and every condition on which the
// reference calls exit(1) (5-wave NaN speeds, cleanUpAlpha, the plastic
// divide, rootFind's bracket) or never returns (rootFind's unbounded loop).
constexpr unsigned ST_STOP = 0x7fu;

// Partial slots of the CFL minimum (Level::dtpart): kernels with many blocks
// spread their per-block atomics over DT_PARTS words (block b on word b mod
// DT_PARTS), so that they do not all contend for the clock's one word;
// k_clock_dt folds them in.  Levels without slots use the clock's word.
constexpr int DT_PARTS = 256;
constexpr unsigned long long DT_NONE = 0x7FEFFFFFFFFFFFFFull;  // std::numeric_limits<double>::max()
__device__ __forceinline__ unsigned long long* dt_slot(const Level& lv, Clock* clk) {
    return lv.dtpart ? lv.dtpart + (blockIdx.x % DT_PARTS) : &clk->dtmin_bits;
}

// Everything below takes or launches with the state view MF (see tj_device.h).
inline namespace TJ_VIEW_NS {

// ------------------------------------------------------------------ x pass
// One wave covers a row segment; lane l holds cell c = f0 - H + l.  Lanes
// reconstruct their cell (at the boundary-mapped cell s), neighbours' data
// travel by wave shuffles, and lane l computes the flux of face c (the lower
// face of cell c).  Faces per wave: 63 (MUSCL) or 61 (MUSCL + THINC/BVD).
template <int NM, bool TH, typename R, bool MX = false>
__device__ __forceinline__ void flux_x_body(const Level& lv, const MF& U, const FacesT<R>& F, const Hist& h, Clock* clk,
                                            long long wave) {
    constexpr int NF = 2 * NM + 5 + (MX ? 1 : 0);  // flux components per face
    if (clk->done) return;
    constexpr int H = TH ? 2 : 1;
    constexpr int PER = TH ? 61 : 63;
    const int lane = threadIdx.x & 63;
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const int nseg = (nx + 1 + PER - 1) / PER;
    if (wave >= (long long)nseg * ny * nz) return;  // uniform per wave
    const int seg = (int)(wave % nseg);
    const long long row = wave / nseg;
    const int j = (int)(row % ny), k = (int)(row / ny);
    const int c = seg * PER - H + lane;
    bool fl;
    const int s = bcmap(c, nx, lv.bc_lo[0], lv.bc_hi[0], fl);
    const int lofs = s - c;  // lane offset of the reconstructed cell

    Prim<NM, R> Pm, Pc, Pp, PL, PR;
    load_prim<NM, R, MX>(lv, U, s - 1, j, k, Pm);
    load_prim<NM, R, MX>(lv, U, s, j, k, Pc);
    load_prim<NM, R, MX>(lv, U, s + 1, j, k, Pp);
    muscl<NM, R, MX>(Pm, Pc, Pp, PL, PR);

    unsigned st = 0;
    const long long cell_s = ((long long)k * ny + j) * nx + s;
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
                    TLr[m] = (R)h.edge_in[m * h.estride + ((long long)k * ny + j) * (nx + 1) + s + 1];
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
        // URStar history: the owner lane of a valid cell records its THINC values
        if (lane >= H && lane < H + PER && c >= 0 && c < nx) {
#pragma unroll
            for (int m = 0; m < NM; ++m)
                if ((A.mixed >> m) & 1) h.urstar[m * h.pstride + cell_s] = (double)A.thr[m];
        }
    }

    FSt<NM, R> SLs, SRs;
    p2c_sound<NM, R, MX>(lv, PL, SLs);
    p2c_sound<NM, R, MX>(lv, PR, SRs);

    // left state of face c: UR(c-1) from the neighbouring lane (or the
    // boundary image of this lane's UR at the domain's low face)
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
    if (MX) Lf.arl = lane_dn(SRs.arl);

    if (lane >= H && lane < H + PER && c >= 0 && c <= nx) {
        if (c == 0) {
            Lf = SRs;
            if (lv.bc_lo[0]) flip_state<NM, R>(Lf, 0);
        }
        FSt<NM, R> Rf = SLs;
        if (c == nx && lv.bc_hi[0]) flip_state<NM, R>(Rf, 0);
        R Fv[NF], sa[NM];
        hllc<NM, R, MX>(Lf, Rf, 0, Fv, sa);
#pragma unroll
        for (int q = 0; q < NF; ++q) F.at(q, c, j, k) = Fv[q];
        if (TH) {
            if (c < nx) {
#pragma unroll
                for (int m = 0; m < NM; ++m) h.star_out[m * h.pstride + ((long long)k * ny + j) * nx + c] = (double)sa[m];
            }
            if (c == nx || (c > 0 && (lv.edge[0][c - 1] & 2))) {
#pragma unroll
                for (int m = 0; m < NM; ++m)
                    h.edge_out[m * h.estride + ((long long)k * ny + j) * (nx + 1) + c] = (double)sa[m];
            }
        }
    }
    if (st) atomicOr(&clk->status, st);
}
// The FP64 THINC form asks for four waves per SIMD: the allocator then fits it
// in 128 VGPRs (from 131, with a 12-byte spill): 0.151 -> 0.133 ms per pass
// on the 2048 x 1024 shock/inclusion level.  The same request on the FP32 and
// solid passes spills more than it gains.
template <int NM, bool TH, typename R = double, bool MX = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TH && !MX && sizeof(R) == 8 ? 4 : 1)))
void k_flux_x(Level lv, MF U, FacesT<R> F, Hist h, Clock* clk) {
    flux_x_body<NM, TH, R, MX>(lv, U, F, h, clk, (long long)blockIdx.x * (blockDim.x >> 6) + wave_in_block());
}

// ------------------------------------------------------------------ XCD-aware block order
// The 8 XCDs of the device take a grid's blocks round-robin (block b runs on
// XCD b mod 8) and each XCD has its own L2.  Renumbering blocks so that each
// XCD works through one contiguous range of cells keeps a 2D level's
// neighbouring rows, which share their faces, in one L2 (measured: the 2D
// update kernels 6-8 % faster).  In 3D the round-robin order is kept: with
// 256 cells per row the cell above in z is 256 blocks on, which round-robin
// puts on the same XCD one generation of resident blocks later, while the
// contiguous order would put it out of the L2 (measured slower).
__device__ __forceinline__ long long xcd_block(long long b, long long nb, int dim) {
    if (dim > 2) return b;
    const long long base = nb / 8, extra = nb % 8;
    const long long x = b % 8, q = b / 8;
    return x * base + (x < extra ? x : extra) + q;  // a bijection of [0, nb)
}

// ------------------------------------------------------------------ update helpers
// One direction of update (hllc.cpp:493-523): U1 += dt/dx (F_lo - F_hi), the
// alpha rows with the non-conservative term -alpha_U (S*_lo - S*_hi).
// The face values are FP64, or FP32 in the mixed form (widened here: the
// accumulation stays FP64).
template <int NM, typename T, bool MX = false>
__device__ __forceinline__ void accum_dir(double* al, double* ar, double* ru, double& E, const double* a0, double dtdx,
                                          const T* Flo, const T* Fhi, double* arl = nullptr) {
    const double s0 = (double)Flo[2 * NM + 4], s1 = (double)Fhi[2 * NM + 4];
#pragma unroll
    for (int m = 0; m < NM; ++m) al[m] += dtdx * ((double)Flo[m] - (double)Fhi[m] - a0[m] * (s0 - s1));
#pragma unroll
    for (int m = 0; m < NM; ++m) ar[m] += dtdx * ((double)Flo[NM + m] - (double)Fhi[NM + m]);
#pragma unroll
    for (int q = 0; q < 3; ++q) ru[q] += dtdx * ((double)Flo[2 * NM + q] - (double)Fhi[2 * NM + q]);
    E += dtdx * ((double)Flo[2 * NM + 3] - (double)Fhi[2 * NM + 3]);
    if (MX) *arl += dtdx * ((double)Flo[2 * NM + 5] - (double)Fhi[2 * NM + 5]);
}
template <int NM, typename T, bool MX = false>
__device__ __forceinline__ void load_faces(const FacesT<T>& Fd, int i, int j, int k, T* Fv) {
#pragma unroll
    for (int q = 0; q < 2 * NM + 5 + (MX ? 1 : 0); ++q) Fv[q] = Fd.at(q, i, j, k);
}
// A state store of the update kernels.  The end-of-step state (stage 2) goes
// out non-temporal: it is not read again until the next step's x pass, a
// level's state is larger than the MALL, and keeping it out of the caches
// leaves them to the stage's reads (measured on the 2048 x 1024
// shock/inclusion level: stage-2 update 0.194 -> 0.188 ms FP64, 0.153 ->
// 0.144 ms mixed, the flux passes 1-3 % faster; the same hint on the stage-1
// store, whose output the stage-2 passes read at once, costs 3-8 %).  The
// solid stage-2 update gains 1 % (FP64) and 1.5 % (mixed) on the 256^3 plate.
template <bool NT>
__device__ __forceinline__ void put_state(double& dst, double v) {
    if constexpr (NT) __builtin_nontemporal_store(v, &dst);
    else dst = v;
}
// A state read of the update kernels that nothing else in the kernel or the
// next one reads again: the conserved components of the stage's input and,
// in stage 2, the step's starting state for the RK average.  Non-temporal,
// which keeps the caches for the faces (same level: the base read in stage 2
// +1.7 % FP64 and +1.5 % mixed on the step; the conserved reads then the
// stage-1 update -5.6 % and stage-2 -3.5 % FP64, stage-2 mixed -8 %).
__device__ __forceinline__ double get_base(const double& src) { return __builtin_nontemporal_load(&src); }
template <int NM, bool MX = false>
__device__ __forceinline__ void load_cons(const Level& lv, const MF& U, int i, int j, int k, double* al, double* ar,
                                          double* ru, double& E, double* arl = nullptr) {
    if (MX) *arl = get_base(U.at(lv.cs[ALPHARHOLAMBDA], i, j, k));
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        al[m] = get_base(U.at(lv.cs[ALPHA] + m, i, j, k));
        ar[m] = get_base(U.at(lv.cs[ALPHARHO] + m, i, j, k));
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) ru[q] = get_base(U.at(lv.cs[RHOU] + q, i, j, k));
    E = get_base(U.at(lv.cs[TOTAL_E], i, j, k));
}
// geometricSourceTerm (radial.cpp:5-31): the factor dt * multiplier /
// ((i + 1/2) dx) of cell i, multiplier -2 in one dimension (spherical) and -1
// otherwise (cylindrical about the x = 0 axis).
__device__ __forceinline__ double radial_factor(const Level& lv, int i, double dt) {
    const double multiplier = lv.dim == 1 ? -2.0 : -1.0;
    const double inverse_r = multiplier / (((double)i + 0.5) * lv.dx[0]);
    return dt * inverse_r;
}
// The geometric fluxes of rhoU and E (flux.cpp:19-29, 51-77) added with
// factor f: s0 is stress column 0 (sigma_r0, r = 0..2); s01, s02, s22 the
// other stress entries the fluxes read.  Each conservative variable reads
// only its own old value and the primitives, so the in-place order of
// radial.cpp:13-28 does not matter.
__device__ __forceinline__ void radial_momentum_energy(double f, const double* u, const double* s0, double s01,
                                                       double s02, double s22, double* ru, double& E) {
    ru[0] += f * (ru[0] * u[0] - s0[0] + s22);
    ru[1] += f * (ru[1] * u[0] - s01);
    ru[2] += f * (-2.0 * s02);
    double vs = 0.0;
#pragma unroll
    for (int r = 0; r < 3; ++r) vs += u[r] * s0[r];
    E += f * (E * u[0] - vs);
}
// alpha has no geometric flux (flux.cpp:55): x + f * 0 leaves every x as it is.
// A mixture's alpha rho lambda takes alpha rho lambda u_x (flux.cpp:57).
template <int NM, bool MX = false>
__device__ __forceinline__ void radial_fluid(double f, const double* u, double p, double* ar, double* ru, double& E,
                                             double* arl = nullptr) {
#pragma unroll
    for (int m = 0; m < NM; ++m) ar[m] += f * (ar[m] * u[0]);
    if (MX) *arl += f * (*arl * u[0]);
    // sigma = -p I of a fluid cell (stressTensor, boxaccesscellarray.cpp:205-214)
    const double s0[3] = {-p * 1.0, -p * 0.0, -p * 0.0};
    radial_momentum_energy(f, u, s0, -p * 0.0, -p * 0.0, -p * 1.0, ru, E);
}

// pressureBased_UpdateMassFraction (reactive.cpp:3-24): one explicit step of
// the burn of alpha rho lambda y at alpha, alpha rho and p; no burn where
// alpha < 0.5, clamped at 0.
__device__ __forceinline__ double burn_step(double alpha, double ar, double y, double p, double dt) {
    const double sigma = 20.0, nu = 0.5, n = 1.2;
    if (alpha < 0.5) return y;
    double y1 = y - (dt) * (sigma * pm_pow(ar, 1.0 - nu) * pm_pow(y, nu) * pm_pow((p / 1E6), n));
    if (y1 < 0.0) y1 = 0.0;
    return y1;
}
// reactiveUpdate (reactive.cpp:60-103) of one cell: Heun's RK2 of the burn
// (the second stage reads the first's alpha rho lambda, the cell's alpha,
// alpha rho and p unchanged), then the average.
__device__ __forceinline__ double burn_rk2(double alpha, double ar, double y, double p, double dt) {
    const double y1 = burn_step(alpha, ar, y, p, dt);
    const double y2 = burn_step(alpha, ar, y1, p, dt);
    return ((y * (1.0 / 2.0)) + (y2 * (1.0 / 2.0)));
}

// conservativeToPrimitive of the updated cell (hllc.cpp:689), and in stage 2
// the RK average (hllc.cpp:705), the reactive source (main.cpp:149-152), the
// geometric source (main.cpp:159-162), cleanUpAlpha (main.cpp:169-171), the
// NaN scan (main.cpp:183-187) and the sound speed / CFL candidate of the next
// getTimeStep (main.cpp:128); then the store.  MX: the mixture's
// alpha rho lambda (arl) rides along, lambda and the phase densities come
// from rootFind in each c2p and are averaged with the rest.
template <int NM, int STAGE, bool MX = false>
__device__ __forceinline__ void finish_cell(const Level& lv, const MF& Ub, const MF& Uo, int i, int j, int k, double* al,
                                            double* ar, double* ru, double E, double dt, unsigned ops, double& cand,
                                            unsigned& st, double arl = 0.0) {
    double rk[NM], rho, u[3], p;
    MixV<double> x{};
    st |= c2p<NM, MX>(lv, al, ar, ru, E, rk, rho, u, p, arl, &x);
    if (STAGE == 2) {
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            al[m] = get_base(Ub.at(lv.cs[ALPHA] + m, i, j, k)) * (1.0 / 2.0) + al[m] * (1.0 / 2.0);
            ar[m] = get_base(Ub.at(lv.cs[ALPHARHO] + m, i, j, k)) * (1.0 / 2.0) + ar[m] * (1.0 / 2.0);
            rk[m] = get_base(Ub.at(lv.cs[RHO_K] + m, i, j, k)) * (1.0 / 2.0) + rk[m] * (1.0 / 2.0);
        }
        rho = get_base(Ub.at(lv.cs[RHO], i, j, k)) * (1.0 / 2.0) + rho * (1.0 / 2.0);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            ru[q] = get_base(Ub.at(lv.cs[RHOU] + q, i, j, k)) * (1.0 / 2.0) + ru[q] * (1.0 / 2.0);
            u[q] = get_base(Ub.at(lv.cs[VELOCITY] + q, i, j, k)) * (1.0 / 2.0) + u[q] * (1.0 / 2.0);
        }
        E = get_base(Ub.at(lv.cs[TOTAL_E], i, j, k)) * (1.0 / 2.0) + E * (1.0 / 2.0);
        p = get_base(Ub.at(lv.cs[PRESSURE], i, j, k)) * (1.0 / 2.0) + p * (1.0 / 2.0);
        if (MX) {
            arl = get_base(Ub.at(lv.cs[ALPHARHOLAMBDA], i, j, k)) * (1.0 / 2.0) + arl * (1.0 / 2.0);
            x.lam = get_base(Ub.at(lv.cs[LAMBDA], i, j, k)) * (1.0 / 2.0) + x.lam * (1.0 / 2.0);
            x.ra = get_base(Ub.at(lv.cs[RHO_MIX], i, j, k)) * (1.0 / 2.0) + x.ra * (1.0 / 2.0);
            x.rb = get_base(Ub.at(lv.cs[RHO_MIX] + 1, i, j, k)) * (1.0 / 2.0) + x.rb * (1.0 / 2.0);
            if (ops & 64u) {  // the burn on the averaged state, then c2p
#pragma unroll
                for (int m = 0; m < NM; ++m)
                    if (m == lv.mixm) arl = burn_rk2(al[m], ar[m], arl, p, dt);
                st |= c2p<NM, MX>(lv, al, ar, ru, E, rk, rho, u, p, arl, &x);
            }
        }
        if (ops & 32u) {  // the averaged sigma is -p I with the averaged p
            radial_fluid<NM, MX>(radial_factor(lv, i, dt), u, p, ar, ru, E, &arl);
            st |= c2p<NM, MX>(lv, al, ar, ru, E, rk, rho, u, p, arl, &x);
        }
        if (ops & 4u) st |= cleanup_alpha<NM>(al);
        if (ops & 16u) {
            bool bad = false;
#pragma unroll
            for (int m = 0; m < NM; ++m) bad |= __builtin_isnan(al[m]) || __builtin_isnan(ar[m]) || __builtin_isnan(rk[m]);
#pragma unroll
            for (int q = 0; q < 3; ++q) bad |= __builtin_isnan(ru[q]) || __builtin_isnan(u[q]);
            bad |= __builtin_isnan(E) || __builtin_isnan(p);
            if (MX) bad |= __builtin_isnan(arl) || __builtin_isnan(x.lam);
            if (bad) st |= 1u;
        }
        if (ops & 8u) {
            if (MX) {
#pragma unroll
                for (int m = 0; m < NM; ++m)
                    if (m == lv.mixm) x.rk = rk[m];
            }
            const double a = sound<NM, MX>(lv, al, ar, rk, rho, p, x);
            put_state<STAGE == 2>(Uo.at(lv.cs[SOUNDSPEED], i, j, k), a);
            #pragma unroll
            for (int d = 0; d < 3; ++d)  // unrolled: constant indices keep the state in registers
                if (d < lv.dim) cand = dmin(cand, lv.dx[d] / (a + __builtin_fabs(u[d])));
        }
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        put_state<STAGE == 2>(Uo.at(lv.cs[ALPHA] + m, i, j, k), al[m]);
        put_state<STAGE == 2>(Uo.at(lv.cs[ALPHARHO] + m, i, j, k), ar[m]);
        put_state<STAGE == 2>(Uo.at(lv.cs[RHO_K] + m, i, j, k), rk[m]);
    }
    put_state<STAGE == 2>(Uo.at(lv.cs[RHO], i, j, k), rho);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        put_state<STAGE == 2>(Uo.at(lv.cs[RHOU] + q, i, j, k), ru[q]);
        put_state<STAGE == 2>(Uo.at(lv.cs[VELOCITY] + q, i, j, k), u[q]);
    }
    put_state<STAGE == 2>(Uo.at(lv.cs[TOTAL_E], i, j, k), E);
    put_state<STAGE == 2>(Uo.at(lv.cs[PRESSURE], i, j, k), p);
    if (MX) {
        put_state<STAGE == 2>(Uo.at(lv.cs[ALPHARHOLAMBDA], i, j, k), arl);
        put_state<STAGE == 2>(Uo.at(lv.cs[LAMBDA], i, j, k), x.lam);
        put_state<STAGE == 2>(Uo.at(lv.cs[RHO_MIX], i, j, k), x.ra);
        put_state<STAGE == 2>(Uo.at(lv.cs[RHO_MIX] + 1, i, j, k), x.rb);
    }
}
// wave min of the positive CFL candidates (as ordered bits)
__device__ __forceinline__ unsigned long long wave_min_bits(double cand) {
    unsigned long long b = __builtin_bit_cast(unsigned long long, cand);
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(b, off, 64);
        b = o < b ? o : b;
    }
    return b;
}
// one atomic per wave (kernels whose waves each cover many cells)
__device__ __forceinline__ void wave_min_to(double cand, unsigned long long* dst) {
    const unsigned long long b = wave_min_bits(cand);
    if ((threadIdx.x & 63) == 0) atomicMin(dst, b);
}
// one atomic per 256-thread block (one-cell-per-thread kernels: keeps the
// number of same-address atomics low)
__device__ __forceinline__ void block_min_to(double cand, unsigned long long* dst) {
    __shared__ unsigned long long red[4];
    const unsigned long long b = wave_min_bits(cand);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long mn = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) mn = red[w] < mn ? red[w] : mn;
        atomicMin(dst, mn);
    }
}

// ------------------------------------------------------------------ y / z pass
// One wave covers 64 columns (lanes along x, coalesced) and marches along
// direction D over a segment of faces [f0, f1); the right state of row r-1
// is carried in registers to the face of row r.  With THINC the alpha
// reconstruction runs one row ahead so the BVD of row r sees rows r-1 and r+1.
// FUSE (last direction of the level): the segment owns rows [f0, f1) and
// completes each row as soon as its upper face is known (update in x, y, z
// order with the earlier directions' faces read from their planes, then
// finish_cell); the faces of this direction stay in registers.
// Per-wave LDS slices of the marching pass, slot-major and lane-minor: the
// right state of row r-1 (2 NM + 7 slots), the alpha reconstructions of rows
// r-1 and r (THINC, 4 NM + 1 slots each) and the lower-face flux of the row
// being completed (fused form, 2 NM + 5 slots).
template <int NM, bool TH, bool FUSE, bool MX = false>
struct MarchSlots {
    static constexpr int FST = 2 * NM + 7 + (MX ? 1 : 0);
    static constexpr int AREC = 4 * NM + 1;
    static constexpr int FLX = 2 * NM + 5;
    static constexpr int N = FST + (TH ? 2 * AREC : 0) + (FUSE ? FLX : 0);
};
template <int NM, typename R, bool MX = false>
__device__ __forceinline__ void fst_put(R* sh, int lane, const FSt<NM, R>& S) {
    int q = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        sh[(q++) * 64 + lane] = S.al[m];
        sh[(q++) * 64 + lane] = S.ar[m];
    }
    sh[(q++) * 64 + lane] = S.rho;
#pragma unroll
    for (int c = 0; c < 3; ++c) sh[(q++) * 64 + lane] = S.u[c];
    sh[(q++) * 64 + lane] = S.E;
    sh[(q++) * 64 + lane] = S.p;
    sh[(q++) * 64 + lane] = S.a;
    if (MX) sh[(q++) * 64 + lane] = S.arl;
}
template <int NM, typename R, bool MX = false>
__device__ __forceinline__ void fst_get(const R* sh, int lane, FSt<NM, R>& S) {
    int q = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        S.al[m] = sh[(q++) * 64 + lane];
        S.ar[m] = sh[(q++) * 64 + lane];
    }
    S.rho = sh[(q++) * 64 + lane];
#pragma unroll
    for (int c = 0; c < 3; ++c) S.u[c] = sh[(q++) * 64 + lane];
    S.E = sh[(q++) * 64 + lane];
    S.p = sh[(q++) * 64 + lane];
    S.a = sh[(q++) * 64 + lane];
    if (MX) S.arl = sh[(q++) * 64 + lane];
}
template <int NM, typename R>
__device__ __forceinline__ void arec_put(R* sh, int lane, const ARec<NM, R>& A) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        sh[(4 * m + 0) * 64 + lane] = A.ula[m];
        sh[(4 * m + 1) * 64 + lane] = A.ura[m];
        sh[(4 * m + 2) * 64 + lane] = A.thl[m];
        sh[(4 * m + 3) * 64 + lane] = A.thr[m];
    }
    sh[(4 * NM) * 64 + lane] = (R)A.mixed;  // a small integer: exact in either type
}
template <int NM, typename R>
__device__ __forceinline__ void arec_get(const R* sh, int lane, ARec<NM, R>& A) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        A.ula[m] = sh[(4 * m + 0) * 64 + lane];
        A.ura[m] = sh[(4 * m + 1) * 64 + lane];
        A.thl[m] = sh[(4 * m + 2) * 64 + lane];
        A.thr[m] = sh[(4 * m + 3) * 64 + lane];
    }
    A.mixed = (int)sh[(4 * NM) * 64 + lane];
}

// One wave covers 64 columns (lanes along x, coalesced) and marches along
// direction D over a segment of faces [f0, f1).  What passes from row r-1 to
// row r -- the right face state, the THINC lookahead reconstructions, the
// fused form's lower face -- waits in the wave's own LDS slice rather than in
// registers, and the primitive rows are reloaded from cache, so the pass
// keeps the register footprint of a single face.  With THINC the alpha
// reconstruction runs one row ahead so the BVD of row r sees rows r-1 and
// r+1.  FUSE (last direction of the level): the segment owns rows [f0, f1)
// and completes each row as soon as its upper face is known (update in
// x, y, z order with the earlier directions' faces read from their planes,
// then finish_cell); the faces of this direction never reach HBM.
template <int NM, bool TH, int D, bool FUSE, int STAGE, typename R, bool MX = false>
__device__ __forceinline__ void flux_march_body(const Level& lv, const MF& U, const FacesT<R>& F, const Hist& h, int seglen,
                                                Clock* clk, const MF& Ub, const MF& Uo, const FacesT<R>& Fx,
                                                const FacesT<R>& Fy, unsigned ops, long long wave) {
    static_assert(!(FUSE && MX), "the fused update has no mixture form");
    constexpr int NF = 2 * NM + 5 + (MX ? 1 : 0);  // flux components per face
    using Slots = MarchSlots<NM, TH, FUSE, MX>;
    __shared__ R shm[4 * Slots::N * 64];
    if (clk->done) return;
    const int lane = threadIdx.x & 63;
    R* const shR = shm + wave_in_block() * Slots::N * 64;  // right state of row r-1
    R* const shAp = shR + Slots::FST * 64;                      // alpha reconstruction of row r-1
    R* const shAc = shAp + Slots::AREC * 64;                    // alpha reconstruction of row r
    R* const shF = shR + (Slots::FST + (TH ? 2 * Slots::AREC : 0)) * 64;  // lower face flux (fused)
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const int nd = lv.n[D];
    const int no = (D == 1) ? nz : ny;
    const int nstrip = (nx + 63) / 64;
    const int nseg = FUSE ? (nd + seglen - 1) / seglen : (nd + 1 + seglen - 1) / seglen;
    if (wave >= (long long)nstrip * nseg * no) return;  // uniform per wave; no block barrier below
    const int strip = (int)(wave % nstrip);
    const long long rest = wave / nstrip;
    const int seg = (int)(rest % nseg);
    const int o = (int)(rest / nseg);
    const int i0 = strip * 64 + lane;
    const bool act = i0 < nx;
    const int i = act ? i0 : nx - 1;
    const int f0 = seg * seglen;
    // faces computed: [f0, f1) (unfused) or [f0, f1] (fused: rows [f0, f1) owned)
    const int f1 = FUSE ? min(f0 + seglen, nd) : min(f0 + seglen, nd + 1);
    const int rb = max(f0 - 1, 0), re = FUSE ? min(f1, nd - 1) : min(f1 - 1, nd - 1);
    const int ftop = FUSE ? f1 : f1 - 1;  // last face of the segment
    auto cj = [&](int r) { return D == 1 ? r : o; };
    auto ck = [&](int r) { return D == 2 ? r : o; };
    auto cellidx = [&](int r) { return ((long long)ck(r) * ny + cj(r)) * nx + i; };
    auto faceidx = [&](int r) {  // lower face r along D
        return D == 1 ? ((long long)ck(r) * (ny + 1) + r) * nx + i : ((long long)r * ny + cj(r)) * nx + i;
    };
    double cand = __builtin_huge_val();
    unsigned st = 0;
    const R ynx = (R)lv.yn[0], yny = (R)lv.yn[1];

    // complete row q (faces q and q+1 known) in the fused form
    auto finish_row = [&](int q, const R* Flo, const R* Fhi) {
        const int jj = cj(q), kk = ck(q);
        double al[NM], ar[NM], ru[3], E, a0[NM];
        load_cons<NM>(lv, U, i, jj, kk, al, ar, ru, E);
#pragma unroll
        for (int m = 0; m < NM; ++m) a0[m] = al[m];
        const double dt = clk->dt;
        {
            R Fa[2 * NM + 5], Fb[2 * NM + 5];
            load_faces<NM, R>(Fx, i, jj, kk, Fa);
            load_faces<NM, R>(Fx, i + 1, jj, kk, Fb);
            accum_dir<NM, R>(al, ar, ru, E, a0, dt / lv.dx[0], Fa, Fb);
        }
        if (D == 2) {
            R Fa[2 * NM + 5], Fb[2 * NM + 5];
            load_faces<NM, R>(Fy, i, jj, kk, Fa);
            load_faces<NM, R>(Fy, i, jj + 1, kk, Fb);
            accum_dir<NM, R>(al, ar, ru, E, a0, dt / lv.dx[1], Fa, Fb);
        }
        accum_dir<NM, R>(al, ar, ru, E, a0, dt / lv.dx[D], Flo, Fhi);
        if (act) finish_cell<NM, STAGE>(lv, Ub, Uo, i, jj, kk, al, ar, ru, E, dt, ops, cand, st);
    };
    auto emit = [&](int face, const FSt<NM, R>& A, const FSt<NM, R>& B, R* Fv) {
        R sa[NM];
        hllc<NM, R, MX>(A, B, D, Fv, sa);
        if (!act) return;
        const long long fi = faceidx(face);
        if (!FUSE) {
            const long long fcs = F.cs;
#pragma unroll
            for (int q = 0; q < NF; ++q) F.p[q * fcs + fi] = Fv[q];
        }
        if (TH) {
            if (face < nd) {
#pragma unroll
                for (int m = 0; m < NM; ++m) h.star_out[m * h.pstride + cellidx(face)] = (double)sa[m];
            }
            if (face == nd || (face > 0 && (lv.edge[D][face - 1] & 2))) {
#pragma unroll
                for (int m = 0; m < NM; ++m) h.edge_out[m * h.estride + fi] = (double)sa[m];
            }
        }
    };

    bool curValid = false;
    FSt<NM, R> SLs, SRs;
    for (int r = rb; r <= re; ++r) {
        Prim<NM, R> PL, PR;
        {
            Prim<NM, R> Pm, Pc, Pp;
            load_prim<NM, R, MX>(lv, U, i, cj(r - 1), ck(r - 1), Pm);
            load_prim<NM, R, MX>(lv, U, i, cj(r), ck(r), Pc);
            load_prim<NM, R, MX>(lv, U, i, cj(r + 1), ck(r + 1), Pp);
            muscl<NM, R, MX>(Pm, Pc, Pp, PL, PR);
        }
        if (TH) {
            const uint8_t ef = lv.edge[D][r];
            const bool blo = ef & 1, bhi = ef & 2;
            if (!curValid) {
                ARec<NM, R> A;
                arec<NM, R>(lv, U, i, cj(r), ck(r), D, ynx, yny, A);
                arec_put<NM, R>(shAc, lane, A);
                if (!blo) {
                    arec<NM, R>(lv, U, i, cj(r - 1), ck(r - 1), D, ynx, yny, A);
                    arec_put<NM, R>(shAp, lane, A);
                }
            }
            ARec<NM, R> Anext;
            if (!bhi) arec<NM, R>(lv, U, i, cj(r + 1), ck(r + 1), D, ynx, yny, Anext);
            ARec<NM, R> Acur;
            arec_get<NM, R>(shAc, lane, Acur);
            const int mixedFlag = (Acur.mixed >> (NM - 1)) & 1;
            const long long cr = cellidx(r);
            const long long cstride = (D == 1) ? nx : (long long)nx * ny;
            if (mixedFlag) {
                ARec<NM, R> Aprev;
                if (!blo) arec_get<NM, R>(shAp, lane, Aprev);
                R URl[NM], TRl[NM], ULr[NM], TLr[NM], thl[NM], thr[NM];
#pragma unroll
                for (int m = 0; m < NM; ++m) {
                    if (blo) {
                        URl[m] = load_alpha<R>(lv, U, m, i, cj(r - 1), ck(r - 1));
                        TRl[m] = R(0);
                    } else {
                        URl[m] = Aprev.ura[m];
                        TRl[m] = ((Aprev.mixed >> m) & 1) ? Aprev.thr[m] : (R)h.urstar[m * h.pstride + cr - cstride];
                    }
                    if (bhi) {
                        ULr[m] = load_alpha<R>(lv, U, m, i, cj(r + 1), ck(r + 1));
                        TLr[m] = (R)h.edge_in[m * h.estride + faceidx(r + 1)];
                    } else {
                        ULr[m] = Anext.ula[m];
                        TLr[m] = ((Anext.mixed >> m) & 1) ? Anext.thl[m] : (R)h.star_in[m * h.pstride + cr + cstride];
                    }
                    if ((Acur.mixed >> m) & 1) {
                        thl[m] = Acur.thl[m];
                        thr[m] = Acur.thr[m];
                    } else {
                        thl[m] = (R)h.star_in[m * h.pstride + cr];
                        thr[m] = (R)h.urstar[m * h.pstride + cr];
                    }
                }
                bvd<NM, R>(1, URl, TRl, ULr, TLr, thl, thr, PL.a, PR.a);
            }
            st |= cleanup_alpha<NM, R>(PL.a);
            st |= cleanup_alpha<NM, R>(PR.a);
            if (act && r >= f0 && r < (FUSE ? f1 : nd)) {
#pragma unroll
                for (int m = 0; m < NM; ++m)
                    if ((Acur.mixed >> m) & 1) h.urstar[m * h.pstride + cr] = (double)Acur.thr[m];
            }
            // rows move up by one: Aprev <- Acur, Acur <- Anext
            arec_put<NM, R>(shAp, lane, Acur);
            if (!bhi) arec_put<NM, R>(shAc, lane, Anext);
            curValid = !bhi;
        }
        p2c_sound<NM, R, MX>(lv, PL, SLs);
        p2c_sound<NM, R, MX>(lv, PR, SRs);
        if (r >= f0 && r <= ftop) {  // face r: left UR(r-1) (the image of UR(0) at the domain face), right UL(r)
            FSt<NM, R> A;
            if (r == 0) {
                A = SRs;
                if (lv.bc_lo[D]) flip_state<NM, R>(A, D);
            } else {
                fst_get<NM, R, MX>(shR, lane, A);
            }
            R Fcur[NF];
            emit(r, A, SLs, Fcur);
            if (FUSE) {
                if (r - 1 >= f0) {
                    R Flo[2 * NM + 5];
#pragma unroll
                    for (int q = 0; q < 2 * NM + 5; ++q) Flo[q] = shF[q * 64 + lane];
                    finish_row(r - 1, Flo, Fcur);
                }
#pragma unroll
                for (int q = 0; q < 2 * NM + 5; ++q) shF[q * 64 + lane] = Fcur[q];
            }
        }
        fst_put<NM, R, MX>(shR, lane, SRs);
    }
    if (ftop == nd) {  // domain face nd: UR(nd-1) and the image of UL(nd-1), from the last row
        if (lv.bc_hi[D]) flip_state<NM, R>(SLs, D);
        R Ftop[NF];
        emit(nd, SRs, SLs, Ftop);
        if (FUSE) {
            R Flo[2 * NM + 5];
#pragma unroll
            for (int q = 0; q < 2 * NM + 5; ++q) Flo[q] = shF[q * 64 + lane];
            finish_row(nd - 1, Flo, Ftop);
        }
    }
    if (FUSE && STAGE == 2 && (ops & 8u)) wave_min_to(cand, dt_slot(lv, clk));
    if (st) atomicOr(&clk->status, st);
}
template <int NM, bool TH, int D, bool FUSE = false, int STAGE = 1, typename R = double, bool MX = false>
__global__ __launch_bounds__(256) void k_flux_march(Level lv, MF U, FacesT<R> F, Hist h, int seglen, Clock* clk,
                                                    MF Ub = MF{}, MF Uo = MF{}, FacesT<R> Fx = FacesT<R>{},
                                                    FacesT<R> Fy = FacesT<R>{}, unsigned ops = 0u) {
    flux_march_body<NM, TH, D, FUSE, STAGE, R, MX>(lv, U, F, h, seglen, clk, Ub, Uo, Fx, Fy, ops,
                                                   (long long)blockIdx.x * (blockDim.x >> 6) + wave_in_block());
}
// The same pass held to two waves per SIMD (the FP64 THINC pass: its few
// registers beyond 256 spill to scratch on rarely taken paths, and the second
// wave hides the FP64 latency; measured faster than one wave per SIMD).
template <int NM, bool TH, int D, typename R = double, bool MX = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_flux_march_w2(
    Level lv, MF U, FacesT<R> F, Hist h, int seglen, Clock* clk) {
    flux_march_body<NM, TH, D, false, 1, R, MX>(lv, U, F, h, seglen, clk, MF{}, MF{}, FacesT<R>{}, FacesT<R>{}, 0u,
                                                (long long)blockIdx.x * (blockDim.x >> 6) + wave_in_block());
}

// ------------------------------------------------------------------ update
// update (hllc.cpp:493-523) for every direction in x, y, z order, then
// conservativeToPrimitive (hllc.cpp:689).  Stage 2 adds the RK average of
// advance (hllc.cpp:705), cleanUpAlpha (main.cpp:169-171), the NaN scan
// (main.cpp:183-187) and getTimeStep of the next step (main.cpp:128).
// (one block of 256 cells: `block` is the block's index within the level)
template <int NM, int STAGE, typename T, bool MX = false>
__device__ __forceinline__ void update_body(const Level& lv, const MF& Uin, const MF& Ub, const MF& Uo, const FacesT<T>& F0,
                                            const FacesT<T>& F1, const FacesT<T>& F2, Clock* clk, unsigned ops,
                                            long long block) {
    constexpr int NF = 2 * NM + 5 + (MX ? 1 : 0);
    if (clk->done) return;
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = block * blockDim.x + threadIdx.x;
    const long long ncell = (long long)nx * ny * nz;
    double cand = __builtin_huge_val();
    unsigned st = 0;
    if (id < ncell) {
        const int i = (int)(id % nx);
        const int j = (int)((id / nx) % ny);
        const int k = (int)(id / ((long long)nx * ny));
        const double dt = clk->dt;
        double al[NM], ar[NM], ru[3], E, a0[NM], arl = 0.0;
        load_cons<NM, MX>(lv, Uin, i, j, k, al, ar, ru, E, &arl);
#pragma unroll
        for (int m = 0; m < NM; ++m) a0[m] = al[m];
        for (int d = 0; d < lv.dim; ++d) {
            const FacesT<T>& Fd = d == 0 ? F0 : (d == 1 ? F1 : F2);
            T Fa[NF], Fb[NF];
            load_faces<NM, T, MX>(Fd, i, j, k, Fa);
            load_faces<NM, T, MX>(Fd, i + (d == 0), j + (d == 1), k + (d == 2), Fb);
            accum_dir<NM, T, MX>(al, ar, ru, E, a0, dt / lv.dx[d], Fa, Fb, &arl);
        }
        finish_cell<NM, STAGE, MX>(lv, Ub, Uo, i, j, k, al, ar, ru, E, dt, ops, cand, st, arl);
    }
    if (STAGE == 2 && (ops & 8u)) block_min_to(cand, dt_slot(lv, clk));
    if (st) atomicOr(&clk->status, st);
}
template <int NM, int STAGE, typename T = double, bool MX = false>
__global__ __launch_bounds__(256) void k_update(Level lv, MF Uin, MF Ub, MF Uo, FacesT<T> F0, FacesT<T> F1,
                                                FacesT<T> F2, Clock* clk, unsigned ops) {
    update_body<NM, STAGE, T, MX>(lv, Uin, Ub, Uo, F0, F1, F2, clk, ops, xcd_block(blockIdx.x, gridDim.x, lv.dim));
}

// getSoundSpeed + boxTimeStepFunction + min over the level (timestep.cpp:7-70).
// A mixture reads its stored lambda and phase densities (MixtureEOS::
// mixtureSoundSpeed, equationofstate.cpp:212-226).
template <int NM, bool MX = false>
__device__ __forceinline__ void sound_dt_body(const Level& lv, const MF& U, Clock* clk, long long block) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = block * blockDim.x + threadIdx.x;
    const long long ncell = (long long)nx * ny * nz;
    double cand = __builtin_huge_val();
    if (id < ncell) {
        const int i = (int)(id % nx);
        const int j = (int)((id / nx) % ny);
        const int k = (int)(id / ((long long)nx * ny));
        double al[NM], ar[NM], rk[NM];
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            al[m] = U.at(lv.cs[ALPHA] + m, i, j, k);
            ar[m] = U.at(lv.cs[ALPHARHO] + m, i, j, k);
            rk[m] = U.at(lv.cs[RHO_K] + m, i, j, k);
        }
        const double rho = U.at(lv.cs[RHO], i, j, k), p = U.at(lv.cs[PRESSURE], i, j, k);
        MixV<double> x{};
        if (MX) {
#pragma unroll
            for (int m = 0; m < NM; ++m)
                if (m == lv.mixm) x.rk = rk[m];
            x.lam = U.at(lv.cs[LAMBDA], i, j, k);
            x.ra = U.at(lv.cs[RHO_MIX], i, j, k);
            x.rb = U.at(lv.cs[RHO_MIX] + 1, i, j, k);
        }
        const double a = sound<NM, MX>(lv, al, ar, rk, rho, p, x);
        U.at(lv.cs[SOUNDSPEED], i, j, k) = a;
        for (int d = 0; d < lv.dim; ++d)
            cand = dmin(cand, lv.dx[d] / (a + __builtin_fabs(U.at(lv.cs[VELOCITY] + d, i, j, k))));
    }
    block_min_to(cand, dt_slot(lv, clk));
}
template <int NM, bool MX = false>
__global__ __launch_bounds__(256) void k_sound_dt(Level lv, MF U, Clock* clk) {
    sound_dt_body<NM, MX>(lv, U, clk, blockIdx.x);
}

#ifndef TJ_TEMPLATE_KERNELS_ONLY  // non-template kernels: one definition, in tj_lib.hip
// dt = CFL * min (timestep.cpp:61-69); with `driver` the loop test and the
// clamps of main.cpp:121, 135-143.  Re-arms the min word and, with `part`
// (DT_PARTS words, launched with one wave), first folds the partial slots
// into it and re-arms them.
__global__ void k_clock_dt(Clock* clk, double cfl, int driver, unsigned long long* part) {
    if (blockIdx.x != 0) return;
    if (part) {
        unsigned long long b = DT_NONE;
        for (int q = (int)threadIdx.x; q < DT_PARTS; q += (int)blockDim.x) {
            const unsigned long long v = part[q];
            b = v < b ? v : b;
            part[q] = DT_NONE;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(b, off, 64);
            b = o < b ? o : b;
        }
        if (threadIdx.x == 0 && b < clk->dtmin_bits) clk->dtmin_bits = b;
    }
    if (threadIdx.x != 0) return;
    if (driver) {
        if (clk->done) return;
        if (!(clk->t < clk->final_time)) {
            clk->done = 1;
            return;
        }
    }
    const double mn = __builtin_bit_cast(double, clk->dtmin_bits);
    double dt = cfl * mn;
    if (driver) {
        if (clk->t + dt > clk->final_time) dt = clk->final_time - clk->t;
        if (clk->n < 5) dt *= 0.2;
    }
    clk->dt = dt;
    clk->dtmin_bits = DT_NONE;
}
// fill of the partial CFL slots with the empty value
__global__ void k_dtpart_reset(unsigned long long* part) {
    for (int q = (int)threadIdx.x; q < DT_PARTS; q += (int)blockDim.x) part[q] = DT_NONE;
}
// end of a loop iteration: the NaN stop of main.cpp:183-187 (before t, n
// advance), else t += dt, n++ (main.cpp:121).
__global__ void k_clock_advance(Clock* clk) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (clk->done) return;
    if (clk->status & ST_STOP) {
        clk->done = 1;
        return;
    }
    clk->t += clk->dt;
    clk->n += 1;
}
__global__ void k_clock_final(Clock* clk, double final_time) {
    if (threadIdx.x == 0 && blockIdx.x == 0) clk->final_time = final_time;
}
__global__ void k_clock_clear_status(Clock* clk) {
    if (threadIdx.x == 0 && blockIdx.x == 0) clk->status = 0;
}
#endif

// stand-alone cleanUpAlpha over the valid cells (boxaccesscellarray.cpp:530-599)
template <int NM>
__global__ __launch_bounds__(256) void k_cleanup_alpha(Level lv, MF U, Clock* clk) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)nx * ny * nz) return;
    const int i = (int)(id % nx);
    const int j = (int)((id / nx) % ny);
    const int k = (int)(id / ((long long)nx * ny));
    double a[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) a[m] = U.at(lv.cs[ALPHA] + m, i, j, k);
    const unsigned st = cleanup_alpha<NM>(a);
#pragma unroll
    for (int m = 0; m < NM; ++m) U.at(lv.cs[ALPHA] + m, i, j, k) = a[m];
    if (st) atomicOr(&clk->status, st);
}

// Mixture variables of a cell (MixtureEOS): alpha rho lambda, lambda, RHO_MIX.
template <int NM>
__device__ __forceinline__ void load_mix(const Level& lv, const MF& U, int i, int j, int k, const double* rk, double& arl,
                                         MixV<double>& x) {
    arl = U.at(lv.cs[ALPHARHOLAMBDA], i, j, k);
#pragma unroll
    for (int m = 0; m < NM; ++m)
        if (m == lv.mixm) x.rk = rk[m];
    x.lam = U.at(lv.cs[LAMBDA], i, j, k);
    x.ra = U.at(lv.cs[RHO_MIX], i, j, k);
    x.rb = U.at(lv.cs[RHO_MIX] + 1, i, j, k);
}
__device__ __forceinline__ void store_mix(const Level& lv, const MF& U, int i, int j, int k, double arl, const MixV<double>& x) {
    U.at(lv.cs[ALPHARHOLAMBDA], i, j, k) = arl;
    U.at(lv.cs[LAMBDA], i, j, k) = x.lam;
    U.at(lv.cs[RHO_MIX], i, j, k) = x.ra;
    U.at(lv.cs[RHO_MIX] + 1, i, j, k) = x.rb;
}
// a fluid cell after c2p: the conserved and primitive variables it changed
template <int NM>
__device__ __forceinline__ void store_c2p(const Level& lv, const MF& U, int i, int j, int k, const double* ar, const double* rk,
                                          double rho, const double* ru, const double* u, double E, double p) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        U.at(lv.cs[ALPHARHO] + m, i, j, k) = ar[m];
        U.at(lv.cs[RHO_K] + m, i, j, k) = rk[m];
    }
    U.at(lv.cs[RHO], i, j, k) = rho;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        U.at(lv.cs[RHOU] + q, i, j, k) = ru[q];
        U.at(lv.cs[VELOCITY] + q, i, j, k) = u[q];
    }
    U.at(lv.cs[TOTAL_E], i, j, k) = E;
    U.at(lv.cs[PRESSURE], i, j, k) = p;
}

// stand-alone geometricSourceTerm of a fluid level (radial.cpp:37-54) with the
// clock's dt: the source, then conservativeToPrimitive
template <int NM, bool MX = false>
__global__ __launch_bounds__(256) void k_radial_fluid(Level lv, MF U, Clock* clk) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)nx * ny * nz) return;
    const int i = (int)(id % nx);
    const int j = (int)((id / nx) % ny);
    const int k = (int)(id / ((long long)nx * ny));
    double al[NM], ar[NM], ru[3], E, rk[NM], rho, u[3], p, arl = 0.0;
    load_cons<NM, MX>(lv, U, i, j, k, al, ar, ru, E, &arl);
#pragma unroll
    for (int q = 0; q < 3; ++q) u[q] = U.at(lv.cs[VELOCITY] + q, i, j, k);
    p = U.at(lv.cs[PRESSURE], i, j, k);
    radial_fluid<NM, MX>(radial_factor(lv, i, clk->dt), u, p, ar, ru, E, &arl);
    MixV<double> x{};
    const unsigned st = c2p<NM, MX>(lv, al, ar, ru, E, rk, rho, u, p, arl, &x);
    store_c2p<NM>(lv, U, i, j, k, ar, rk, rho, ru, u, E, p);
    if (MX) store_mix(lv, U, i, j, k, arl, x);
    if (st) atomicOr(&clk->status, st);
}

// stand-alone reactiveUpdate of a level with a mixture (reactive.cpp:60-103)
// with the clock's dt: Heun's RK2 of the burn, then conservativeToPrimitive
template <int NM>
__global__ __launch_bounds__(256) void k_reactive_fluid(Level lv, MF U, Clock* clk) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)nx * ny * nz) return;
    const int i = (int)(id % nx);
    const int j = (int)((id / nx) % ny);
    const int k = (int)(id / ((long long)nx * ny));
    double al[NM], ar[NM], ru[3], E, rk[NM], rho, u[3], p, arl = 0.0;
    load_cons<NM, true>(lv, U, i, j, k, al, ar, ru, E, &arl);
    p = U.at(lv.cs[PRESSURE], i, j, k);
#pragma unroll
    for (int m = 0; m < NM; ++m)
        if (m == lv.mixm) arl = burn_rk2(al[m], ar[m], arl, p, clk->dt);
    MixV<double> x{};
    const unsigned st = c2p<NM, true>(lv, al, ar, ru, E, rk, rho, u, p, arl, &x);
    store_c2p<NM>(lv, U, i, j, k, ar, rk, rho, ru, u, E, p);
    store_mix(lv, U, i, j, k, arl, x);
    if (st) atomicOr(&clk->status, st);
}

// contains_nan over conserved and primitive components (boxaccesscellarray.cpp:601-645)
template <int NM, bool MX = false>
__global__ __launch_bounds__(256) void k_nancheck(Level lv, MF U, Clock* clk) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)nx * ny * nz) return;
    const int i = (int)(id % nx);
    const int j = (int)((id / nx) % ny);
    const int k = (int)(id / ((long long)nx * ny));
    bool bad = false;
    const int vars[6] = {ALPHA, ALPHARHO, RHO_K, RHOU, VELOCITY, TOTAL_E};
    const int cnt[6] = {NM, NM, NM, 3, 3, 1};
    for (int v = 0; v < 6; ++v)
        for (int q = 0; q < cnt[v]; ++q) bad |= __builtin_isnan(U.at(lv.cs[vars[v]] + q, i, j, k));
    bad |= __builtin_isnan(U.at(lv.cs[PRESSURE], i, j, k));
    if (MX) bad |= __builtin_isnan(U.at(lv.cs[LAMBDA], i, j, k)) || __builtin_isnan(U.at(lv.cs[ALPHARHOLAMBDA], i, j, k));
    if (bad) atomicOr(&clk->status, 1u);
}

// setInitialConditions' pressure shift and primitiveToConservative
// (initialise.cpp:104, boxaccesscellarray.cpp:124-174) over valid cells.  A
// mixture's pressure shift reads the phase densities the initial state holds
// (initialise.cpp:66-71); p2c then defines them from the shifted p.
template <int NM, bool MX = false>
__device__ __forceinline__ void init_p2c_body(const Level& lv, const MF& U, int shift, long long block) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = block * blockDim.x + threadIdx.x;
    if (id >= (long long)nx * ny * nz) return;
    const int i = (int)(id % nx);
    const int j = (int)((id / nx) % ny);
    const int k = (int)(id / ((long long)nx * ny));
    Prim<NM> P;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        P.a[m] = U.at(lv.cs[ALPHA] + m, i, j, k);
        P.r[m] = U.at(lv.cs[RHO_K] + m, i, j, k);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) P.u[q] = U.at(lv.cs[VELOCITY] + q, i, j, k);
    P.p = U.at(lv.cs[PRESSURE], i, j, k);
    if (MX) P.lam = U.at(lv.cs[LAMBDA], i, j, k);
    if (shift) {
        double ar[NM], G, Ent, Pnt, arl;
#pragma unroll
        for (int m = 0; m < NM; ++m) ar[m] = P.a[m] * P.r[m];
        MixV<double> x{};
        if (MX) load_mix<NM>(lv, U, i, j, k, P.r, arl, x);
        eos_sums<NM, double, MX>(lv, P.a, ar, G, Ent, Pnt, x);
        P.p += Pnt / G;
        U.at(lv.cs[PRESSURE], i, j, k) = P.p;
    }
    FSt<NM> S;
    MixV<double> xo{};
    p2c_sound<NM, double, MX>(lv, P, S, &xo);
#pragma unroll
    for (int m = 0; m < NM; ++m) U.at(lv.cs[ALPHARHO] + m, i, j, k) = S.ar[m];
    U.at(lv.cs[RHO], i, j, k) = S.rho;
#pragma unroll
    for (int q = 0; q < 3; ++q) U.at(lv.cs[RHOU] + q, i, j, k) = S.u[q] * S.rho;
    U.at(lv.cs[TOTAL_E], i, j, k) = S.E;
    if (MX) {
        U.at(lv.cs[ALPHARHOLAMBDA], i, j, k) = S.arl;
        U.at(lv.cs[RHO_MIX], i, j, k) = xo.ra;
        U.at(lv.cs[RHO_MIX] + 1, i, j, k) = xo.rb;
    }
}
template <int NM, bool MX = false>
__global__ __launch_bounds__(256) void k_init_p2c(Level lv, MF U, int shift) {
    init_p2c_body<NM, MX>(lv, U, shift, blockIdx.x);
}

// conservativeToPrimitive over the valid cells.
template <int NM, bool MX = false>
__global__ __launch_bounds__(256) void k_c2p(Level lv, MF U, Clock* clk) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)nx * ny * nz) return;
    const int i = (int)(id % nx);
    const int j = (int)((id / nx) % ny);
    const int k = (int)(id / ((long long)nx * ny));
    double al[NM], ar[NM], ru[3], rk[NM], u[3], rho, p, E, arl = 0.0;
    load_cons<NM, MX>(lv, U, i, j, k, al, ar, ru, E, &arl);
    MixV<double> x{};
    const unsigned st = c2p<NM, MX>(lv, al, ar, ru, E, rk, rho, u, p, arl, &x);
#pragma unroll
    for (int m = 0; m < NM; ++m) U.at(lv.cs[RHO_K] + m, i, j, k) = rk[m];
    U.at(lv.cs[RHO], i, j, k) = rho;
#pragma unroll
    for (int q = 0; q < 3; ++q) U.at(lv.cs[VELOCITY] + q, i, j, k) = u[q];
    U.at(lv.cs[PRESSURE], i, j, k) = p;
    if (MX) store_mix(lv, U, i, j, k, arl, x);
    if (st) atomicOr(&clk->status, st);
}

#ifndef TJ_TEMPLATE_KERNELS_ONLY  // non-template kernels: one definition, in tj_lib.hip
// sigma = -p I for fluids (stressTensor, boxaccesscellarray.cpp:176-216).
__global__ __launch_bounds__(256) void k_derive_fluid(Level lv, MF U) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)nx * ny * nz) return;
    const int i = (int)(id % nx);
    const int j = (int)((id / nx) % ny);
    const int k = (int)(id / ((long long)nx * ny));
    const double p = U.at(lv.cs[PRESSURE], i, j, k);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) U.at(lv.cs[SIGMA] + 3 * r + c, i, j, k) = -p * (r == c ? 1.0 : 0.0);
}

// CellArray RK algebra (cellarray.cpp:80-99): Z = (X * a) + (Y * b) over
// every component and valid cell, each product rounded before the sum as in
// MultiFab::mult followed by MultiFab::Add.  Without Y: Z = X * a.
__global__ __launch_bounds__(256) void k_lincomb(MF X, double a, MF Y, double b, int hasY, MF Z, int ncomp, int nx,
                                                 int ny, int nz) {
    const long long ncell = (long long)nx * ny * nz;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ncell * ncomp) return;
    const int c = (int)(id / ncell);
    const long long cell = id % ncell;
    const int i = (int)(cell % nx);
    const int j = (int)((cell / nx) % ny);
    const int k = (int)(cell / ((long long)nx * ny));
    const double xa = X.at(c, i, j, k) * a;
    Z.at(c, i, j, k) = hasY ? xa + Y.at(c, i, j, k) * b : xa;
}

// zero-fill for scratch that a caller may capture in a graph (no memset nodes)
__global__ void k_zero(double* p, long long n) {
    long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; id < n; id += stride) p[id] = 0.0;
}
#endif

}  // inline namespace TJ_VIEW_NS
}  // namespace tj
