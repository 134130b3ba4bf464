// C ABI of the tangent-linear form of the solid path (include/tenjin.h,
// "Tangent-linear (differentiated) form").
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#define TJ_TEMPLATE_KERNELS_ONLY  // the plain kernels of tj_kernels.h live in tj_lib.hip
#include "tj_ctx.h"
#include "tj_tangent.h"

struct tj_tangent {
    tj_ctx* c = nullptr;
    int ndir = 0, chunk = 0;
    std::vector<int> mat, par;
    void* d_mat = nullptr;            // MatT<Dual<chunk>>[MAXM] of the group being run
    double* d_tan[2] = {nullptr, nullptr};  // derivative blocks of S and T: chunk x ncomp x ncells
    double* d_flux[3] = {nullptr, nullptr, nullptr};
    Faces F[3]{};
    double* d_cand = nullptr;         // CFL candidates: (1 + chunk) x ncells
    unsigned long long* d_arg = nullptr;
    double* d_tdt = nullptr;          // dt and its derivatives (1 + chunk)
    double* d_tt = nullptr;           // derivatives of t (chunk)
    double* d_trace = nullptr;
    long long trace_cap = 0;
    // THINC history planes with derivative parts ((1 + chunk) x NM planes each)
    double* d_urstar = nullptr;
    double* d_star[2] = {nullptr, nullptr};
    double* d_estar[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    long long estride[3] = {0, 0, 0};
    int ncomp = 0;
    int group_first = -1, group_cnt = 0;
};

namespace {

void tg_free(tj_tangent* tg) {
    auto f = [](void* p) {
        if (p) (void)hipFree(p);
    };
    f(tg->d_mat);
    f(tg->d_tan[0]);
    f(tg->d_tan[1]);
    for (int a = 0; a < 3; ++a) f(tg->d_flux[a]);
    f(tg->d_cand);
    f(tg->d_arg);
    f(tg->d_tdt);
    f(tg->d_tt);
    f(tg->d_trace);
    f(tg->d_urstar);
    for (int q = 0; q < 2; ++q) {
        f(tg->d_star[q]);
        for (int a = 0; a < 3; ++a) f(tg->d_estar[a][q]);
    }
}

// The materials of a group: values from the bound level; derivative part q
// seeded 1 in the parameter of direction first + q (Gamma0 enters as Gg,
// Gg + 1 and 1/Gg).
template <int K>
void fill_mats(const tj_tangent* tg, int first, int cnt, MatT<Dual<K>>* M) {
    const Level& lv = tg->c->lv;
    for (int m = 0; m < MAXM; ++m) {
        const MatC& s = lv.m[m];
        MatT<Dual<K>>& d = M[m];
        d.eos = s.eos;
        d.solid = s.solid;
        d.plastic = s.plastic;
        d.mix = s.mix;
        d.Gg = s.Gg;
        d.gp1 = s.gp1;
        d.xi = s.xi;
        d.pref = s.pref;
        d.eref = s.eref;
        d.CV = s.CV;
        d.rho0 = s.rho0;
        d.c0 = s.c0;
        d.s = s.s;
        d.G0 = s.G0;
        d.yield = s.yield;
        d.ga = s.ga;
        d.g2 = s.g2;
        d.Gg2 = s.Gg2;
        d.pref2 = s.pref2;
        d.eref2 = s.eref2;
        d.CV2 = s.CV2;
        d.K0 = s.K0;
        d.ea = s.ea;
        d.eb = s.eb;
        d.e1 = s.e1;
        d.e2 = s.e2;
        d.e3 = s.e3;
        d.e4 = s.e4;
        d.e5 = s.e5;
    }
    for (int q = 0; q < cnt; ++q) {
        const int m = tg->mat[first + q];
        const MatC& s = lv.m[m];
        switch (tg->par[first + q]) {
            case TJ_TP_GAMMA0:
                M[m].Gg.d[q] = 1.0;
                M[m].gp1.d[q] = 1.0;
                M[m].xi.d[q] = -1.0 / (s.Gg * s.Gg);
                break;
            case TJ_TP_S:
                M[m].s.d[q] = 1.0;
                break;
            case TJ_TP_PREF:
                M[m].pref.d[q] = 1.0;
                break;
            default:
                M[m].yield.d[q] = 1.0;
                break;
        }
    }
}

template <int K>
int run_group(tj_tangent* tg, int first, int cnt, const tj_mf* P0, int shift, tj_mf* S, tj_mf* T, double final_time,
              int nsteps, const int probe[4], double* trace, int* steps_done) {
    tj_ctx* c = tg->c;
    const Level& lv0 = c->lv;
    const int nx = lv0.n[0], ny = lv0.n[1], nz = lv0.n[2];
    const long long nc = lv0.ncells;
    MatT<Dual<K>> hm[MAXM];
    fill_mats<K>(tg, first, cnt, hm);
    HIPCHK(c, hipMemcpyAsync(tg->d_mat, hm, sizeof(hm), hipMemcpyHostToDevice, c->stream));
    Level lv = lv0;
    lv.dtpart = nullptr;  // the tangent kernels reduce into the clock's word (k_tclock_dt reads it alone)
    lv.md = tg->d_mat;
    const long long bs = (long long)tg->ncomp * nc;
    DMF<K> dS{view(S), MF{tg->d_tan[0], nx, (long long)nx * ny, nc}, bs};
    DMF<K> dT{view(T), MF{tg->d_tan[1], nx, (long long)nx * ny, nc}, bs};
    const hipStream_t st = c->stream;
    const unsigned gc = grid1(nc, 256);
    // P0 -> S, zero derivatives, setInitialConditions
    hipLaunchKernelGGL(k_copy_mf, dim3(grid1(nc * tg->ncomp, 256)), dim3(256), 0, st, view(P0), view(S), tg->ncomp, nx,
                       ny, nz);
    // tj_reset_history: THINC history planes zero, pass counters at 0
    const bool thinc = c->desc.THINC != 0;
    const long long hp = (long long)c->nmat * nc;
    if (thinc) {
        HIPCHK(c, hipMemsetAsync(tg->d_urstar, 0, sizeof(double) * hp * (K + 1), st));
        for (int q = 0; q < 2; ++q) {
            HIPCHK(c, hipMemsetAsync(tg->d_star[q], 0, sizeof(double) * hp * (K + 1), st));
            for (int a = 0; a < lv0.dim; ++a)
                HIPCHK(c, hipMemsetAsync(tg->d_estar[a][q], 0, sizeof(double) * c->nmat * tg->estride[a] * (K + 1), st));
        }
    }
    long long pass = 0, dpass[3] = {0, 0, 0};
    HIPCHK(c, hipMemsetAsync(tg->d_tan[0], 0, sizeof(double) * bs * K, st));
    HIPCHK(c, hipMemsetAsync(tg->d_tan[1], 0, sizeof(double) * bs * K, st));
    const bool solid = lv.solid != 0;
    if (solid) hipLaunchKernelGGL((k_tinit_p2c_s<2, K>), dim3(gc), dim3(256), 0, st, lv, dS, shift);
    else hipLaunchKernelGGL((k_tinit_p2c_f<2, K>), dim3(gc), dim3(256), 0, st, lv, dS, shift);
    // tj_begin_run: clock reset, final time, getTimeStep of the initial state
    int r;
    if ((r = tj_set_dt(c, 0.0, 1))) return r;
    hipLaunchKernelGGL(k_tclock_begin<K>, dim3(1), dim3(64), 0, st, c->d_clk, final_time, tg->d_arg, tg->d_tt);
    if (solid) hipLaunchKernelGGL((k_tsound_dt_s<2, K>), dim3(gc), dim3(256), 0, st, lv, dS, c->d_clk, tg->d_cand);
    else hipLaunchKernelGGL((k_tsound_dt_f<2, K>), dim3(gc), dim3(256), 0, st, lv, dS, c->d_clk, tg->d_cand);
    hipLaunchKernelGGL(k_targmin, dim3(gc), dim3(256), 0, st, c->d_clk, tg->d_cand, nc, tg->d_arg);
    unsigned ops = TJ_OP_CLEAN_ALPHA | TJ_OP_DT | TJ_OP_NANCHECK;
    if (solid) ops |= TJ_OP_CLEAN_V;
    if (c->desc.PLASTIC) ops |= TJ_OP_PLASTIC;
    auto fluxes = [&](const DMF<K>& U) {
        for (int d = 0; d < lv.dim; ++d) {
            const Faces& F = tg->F[d];
            THist h{};
            if (thinc) {  // hist_for of tj_lib.hip
                const int q = (int)(pass & 1), e = (int)(dpass[d] & 1);
                h.urstar = tg->d_urstar;
                h.star_in = tg->d_star[q];
                h.star_out = tg->d_star[q ^ 1];
                h.edge_in = tg->d_estar[d][e];
                h.edge_out = tg->d_estar[d][e ^ 1];
                h.pstride = nc;
                h.estride = tg->estride[d];
                h.pblk = hp;
                h.eblk = (long long)c->nmat * tg->estride[d];
                pass += 1;
                dpass[d] += 1;
            }
            // tiles of 64 x 1 (x pass) or 8 x 8 cells giving 63 or 8 x 7 faces, 4 waves per block
            const int per = d == 0 ? 63 : 7;
            const long long nseg = (lv.n[d] + 1 + per - 1) / per;
            const long long nw = d == 0 ? nseg * lv.n[1] * lv.n[2]
                                        : (long long)((lv.n[0] + 7) / 8) * nseg * (d == 1 ? lv.n[2] : lv.n[1]);
            const unsigned gt = grid1(nw, 4);
            if (solid) {
                if (d == 0) hipLaunchKernelGGL((k_tflux_ts<2, K, 0>), dim3(gt), dim3(256), 0, st, lv, U, F, c->d_clk);
                else if (d == 1) hipLaunchKernelGGL((k_tflux_ts<2, K, 1>), dim3(gt), dim3(256), 0, st, lv, U, F, c->d_clk);
                else hipLaunchKernelGGL((k_tflux_ts<2, K, 2>), dim3(gt), dim3(256), 0, st, lv, U, F, c->d_clk);
            } else if (thinc) {  // THINC levels are 1D or 2D
                if (d == 0) hipLaunchKernelGGL((k_tflux_tf<2, K, 0, true>), dim3(gt), dim3(256), 0, st, lv, U, F, h, c->d_clk);
                else hipLaunchKernelGGL((k_tflux_tf<2, K, 1, true>), dim3(gt), dim3(256), 0, st, lv, U, F, h, c->d_clk);
            } else {
                if (d == 0) hipLaunchKernelGGL((k_tflux_tf<2, K, 0, false>), dim3(gt), dim3(256), 0, st, lv, U, F, h, c->d_clk);
                else if (d == 1) hipLaunchKernelGGL((k_tflux_tf<2, K, 1, false>), dim3(gt), dim3(256), 0, st, lv, U, F, h, c->d_clk);
                else hipLaunchKernelGGL((k_tflux_tf<2, K, 2, false>), dim3(gt), dim3(256), 0, st, lv, U, F, h, c->d_clk);
            }
        }
    };
    const Faces& F0 = tg->F[0];
    const Faces& F1 = lv.dim > 1 ? tg->F[1] : tg->F[0];
    const Faces& F2 = lv.dim > 2 ? tg->F[2] : tg->F[0];
    for (int q = 0; q < nsteps; ++q) {
        hipLaunchKernelGGL(k_tclock_dt<K>, dim3(1), dim3(64), 0, st, c->d_clk, lv.cfl, tg->d_cand, nc, tg->d_arg,
                           tg->d_tdt, tg->d_tt);
        fluxes(dS);
        if (solid)
            hipLaunchKernelGGL((k_tupdate_s<2, 1, K>), dim3(gc), dim3(256), 0, st, lv, dS, dS, dT, F0, F1, F2,
                               c->d_clk, tg->d_tdt, 0u, tg->d_cand);
        else
            hipLaunchKernelGGL((k_tupdate_f<2, 1, K>), dim3(gc), dim3(256), 0, st, lv, dS, dS, dT, F0, F1, F2,
                               c->d_clk, tg->d_tdt, 0u, tg->d_cand);
        fluxes(dT);
        if (solid)
            hipLaunchKernelGGL((k_tupdate_s<2, 2, K>), dim3(gc), dim3(256), 0, st, lv, dT, dS, dS, F0, F1, F2,
                               c->d_clk, tg->d_tdt, ops, tg->d_cand);
        else
            hipLaunchKernelGGL((k_tupdate_f<2, 2, K>), dim3(gc), dim3(256), 0, st, lv, dT, dS, dS, F0, F1, F2,
                               c->d_clk, tg->d_tdt, ops, tg->d_cand);
        hipLaunchKernelGGL(k_targmin, dim3(gc), dim3(256), 0, st, c->d_clk, tg->d_cand, nc, tg->d_arg);
        hipLaunchKernelGGL(k_tclock_advance<K>, dim3(1), dim3(64), 0, st, c->d_clk, tg->d_tdt, tg->d_tt);
        hipLaunchKernelGGL(k_tprobe<K>, dim3(1), dim3(64), 0, st, dS, probe[0], probe[1], probe[2], probe[3],
                           tg->d_trace + (long long)q * (K + 1));
        HIPCHK(c, hipGetLastError());
    }
    std::vector<double> h((size_t)nsteps * (K + 1));
    if (nsteps > 0)
        HIPCHK(c, hipMemcpyAsync(h.data(), tg->d_trace, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    Clock ck;
    HIPCHK(c, hipMemcpyAsync(&ck, c->d_clk, sizeof(Clock), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    if (steps_done) *steps_done = ck.n;
    if (trace) {
        const int w = 1 + tg->ndir;
        for (int q = 0; q < nsteps; ++q) {
            trace[(long long)q * w] = h[(size_t)q * (K + 1)];
            for (int t = 0; t < cnt; ++t) trace[(long long)q * w + 1 + first + t] = h[(size_t)q * (K + 1) + 1 + t];
        }
    }
    tg->group_first = first;
    tg->group_cnt = cnt;
    return TJ_OK;
}

}  // namespace

extern "C" {

int tj_tangent_create(tj_ctx* c, int ndir, const int* mat, const int* param, int chunk, tj_tangent** out) {
    if (!out) return TJ_EINVAL;
    *out = nullptr;
    if (!c || !c->bound) return TJ_ENOLEVEL;
    if (c->desc.RADIAL || c->lv.mixm >= 0) {
        c->err = "tangent: solid levels, and fluid levels with MUSCL(+THINC) + HLLC (no mixture or geometric source)";
        return TJ_EUNSUPPORTED;
    }
    if (ndir < 1 || !mat || !param) { c->err = "tangent: no directions"; return TJ_EINVAL; }
    if (chunk == 0) chunk = ndir <= 3 ? ndir : 6;
    if (chunk == 4 || chunk == 5) chunk = 6;
    if (chunk != 1 && chunk != 2 && chunk != 3 && chunk != 6) {
        c->err = "tangent: chunk must be 0, 1, 2, 3 or 6";
        return TJ_EINVAL;
    }
    for (int m = 0; m < c->nmat; ++m)
        if (c->lv.m[m].eos >= 2) {
            c->err = "tangent: the Romenskii and Wilkins solid EOSs are not part of the tangent-linear form";
            return TJ_EUNSUPPORTED;
        }
    for (int q = 0; q < ndir; ++q) {
        const int m = mat[q], p = param[q];
        if (m < 0 || m >= c->nmat || p < TJ_TP_GAMMA0 || p > TJ_TP_PREF) {
            c->err = "tangent: bad direction (material or parameter out of range)";
            return TJ_EINVAL;
        }
        const MatC& M = c->lv.m[m];
        if ((p == TJ_TP_S || p == TJ_TP_YIELD) && (M.eos != 1 || !M.solid || (p == TJ_TP_YIELD && !M.plastic))) {
            c->err = "tangent: s and the yield stress belong to Us-up solid materials (yield stress: plastic ones)";
            return TJ_EINVAL;
        }
        if (p == TJ_TP_PREF && M.eos != 0) {
            c->err = "tangent: pref belongs to MieGruneisen (stiffened-gas) materials";
            return TJ_EINVAL;
        }
    }
    tj_tangent* tg = new tj_tangent();
    tg->c = c;
    tg->ndir = ndir;
    tg->chunk = chunk;
    tg->mat.assign(mat, mat + ndir);
    tg->par.assign(param, param + ndir);
    tg->ncomp = tj_ncomp(c);
    const long long nc = c->lv.ncells;
    const int NF = c->desc.SOLID ? 3 * c->nmat + 17 : 2 * c->nmat + 5;
    auto fail = [&](hipError_t e, const char* what) {
        c->err = std::string("tangent: ") + what + ": " + hipGetErrorString(e);
        tg_free(tg);
        delete tg;
        return TJ_EHIP;
    };
    hipError_t e;
    if ((e = hipSetDevice(c->device)) != hipSuccess) return fail(e, "hipSetDevice");
    if ((e = hipMalloc(&tg->d_mat, sizeof(MatT<Dual<6>>) * MAXM)) != hipSuccess) return fail(e, "hipMalloc");
    for (int b = 0; b < 2; ++b)
        if ((e = hipMalloc(&tg->d_tan[b], sizeof(double) * chunk * tg->ncomp * nc)) != hipSuccess)
            return fail(e, "hipMalloc");
    for (int a = 0; a < c->lv.dim; ++a) {
        Faces& F = tg->F[a];
        F.nx = c->lv.n[0] + (a == 0);
        F.ny = c->lv.n[1] + (a == 1);
        F.nz = c->lv.n[2] + (a == 2);
        F.cs = (long long)F.nx * F.ny * F.nz;
        if ((e = hipMalloc(&tg->d_flux[a], sizeof(double) * F.cs * NF * (chunk + 1))) != hipSuccess)
            return fail(e, "hipMalloc");
        F.p = tg->d_flux[a];
    }
    if ((e = hipMalloc(&tg->d_cand, sizeof(double) * nc * (chunk + 1))) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&tg->d_arg, sizeof(unsigned long long))) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&tg->d_tdt, sizeof(double) * (chunk + 1))) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&tg->d_tt, sizeof(double) * chunk)) != hipSuccess) return fail(e, "hipMalloc");
    if (c->desc.THINC) {
        const long long hp = (long long)c->nmat * nc * (chunk + 1);
        if ((e = hipMalloc(&tg->d_urstar, sizeof(double) * hp)) != hipSuccess) return fail(e, "hipMalloc");
        for (int q = 0; q < 2; ++q) {
            if ((e = hipMalloc(&tg->d_star[q], sizeof(double) * hp)) != hipSuccess) return fail(e, "hipMalloc");
            for (int a = 0; a < c->lv.dim; ++a) {
                tg->estride[a] = tg->F[a].cs;
                if ((e = hipMalloc(&tg->d_estar[a][q], sizeof(double) * c->nmat * tg->estride[a] * (chunk + 1))) !=
                    hipSuccess)
                    return fail(e, "hipMalloc");
            }
        }
    }
    *out = tg;
    return TJ_OK;
}

void tj_tangent_destroy(tj_tangent* tg) {
    if (!tg) return;
    if (tg->c) {
        (void)hipSetDevice(tg->c->device);
        (void)hipStreamSynchronize(tg->c->stream);
    }
    tg_free(tg);
    delete tg;
}

int tj_tangent_run(tj_tangent* tg, const tj_mf* P0, int shift, tj_mf* S, tj_mf* T, double final_time, int nsteps,
                   int probe_comp, int pi, int pj, int pk, double* trace, int* steps_done) {
    if (!tg || !tg->c) return TJ_EINVAL;
    tj_ctx* c = tg->c;
    int r;
    if ((r = check_mf(c, P0)) || (r = check_mf(c, S)) || (r = check_mf(c, T))) return r;
    if (S->data == T->data || P0->data == S->data || P0->data == T->data) {
        c->err = "tangent: P0, S and T must be distinct buffers (every group restarts from P0)";
        return TJ_EINVAL;
    }
    if (nsteps < 0) { c->err = "tangent: nsteps < 0"; return TJ_EINVAL; }
    const Level& lv = c->lv;
    if (probe_comp < 0 || probe_comp >= tg->ncomp || pi < 0 || pi >= lv.n[0] || pj < 0 || pj >= lv.n[1] || pk < 0 ||
        pk >= lv.n[2]) {
        c->err = "tangent: probe outside the level";
        return TJ_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->device));
    const long long need = (long long)nsteps * (tg->chunk + 1);
    if (need > tg->trace_cap) {
        if (tg->d_trace) HIPCHK(c, hipFree(tg->d_trace));
        tg->d_trace = nullptr;
        tg->trace_cap = 0;
        HIPCHK(c, hipMalloc(&tg->d_trace, sizeof(double) * need));
        tg->trace_cap = need;
    }
    const int probe[4] = {probe_comp, pi, pj, pk};
    int done = 0;
    for (int first = 0; first < tg->ndir; first += tg->chunk) {
        const int cnt = std::min(tg->chunk, tg->ndir - first);
        switch (tg->chunk) {
            case 1: r = run_group<1>(tg, first, cnt, P0, shift, S, T, final_time, nsteps, probe, trace, &done); break;
            case 2: r = run_group<2>(tg, first, cnt, P0, shift, S, T, final_time, nsteps, probe, trace, &done); break;
            case 3: r = run_group<3>(tg, first, cnt, P0, shift, S, T, final_time, nsteps, probe, trace, &done); break;
            default: r = run_group<6>(tg, first, cnt, P0, shift, S, T, final_time, nsteps, probe, trace, &done); break;
        }
        if (r) return r;
    }
    if (steps_done) *steps_done = done;
    return TJ_OK;
}

int tj_tangent_state(tj_tangent* tg, int dir, tj_mf* D) {
    if (!tg || !tg->c) return TJ_EINVAL;
    tj_ctx* c = tg->c;
    int r;
    if ((r = check_mf(c, D))) return r;
    if (dir < tg->group_first || dir >= tg->group_first + tg->group_cnt) {
        c->err = "tangent: direction not in the last group run";
        return TJ_EINVAL;
    }
    const Level& lv = c->lv;
    const long long nc = lv.ncells;
    const MF X{tg->d_tan[0] + (long long)(dir - tg->group_first) * tg->ncomp * nc, lv.n[0], (long long)lv.n[0] * lv.n[1], nc};
    hipLaunchKernelGGL(k_copy_mf, dim3(grid1(nc * tg->ncomp, 256)), dim3(256), 0, c->stream, X, view(D), tg->ncomp,
                       lv.n[0], lv.n[1], lv.n[2]);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TJ_OK;
}

}  // extern "C"
