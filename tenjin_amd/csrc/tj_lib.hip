// C ABI of the gfx950 per-cell update (include/tenjin.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <map>
#include <queue>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "tj_ctx.h"

namespace {
hipEvent_t ev_get(tj_ctx* c) {
    if (!c->pool.empty()) { hipEvent_t e = c->pool.back(); c->pool.pop_back(); return e; }
    hipEvent_t e;
    hipEventCreate(&e);
    return e;
}
// brackets one launch with events when timing is on
struct Timed {
    tj_ctx* c;
    const char* name;
    hipEvent_t a{}, b{};
    bool on;
    Timed(tj_ctx* c_, const char* n)
        : c(c_), name(n), on(c_->prof && (c_->prof_only.empty() || c_->prof_only == n) && sample(c_)) {
        if (on) { a = ev_get(c); hipEventRecord(a, c->stream); }
    }
    static bool sample(tj_ctx* c) { return c->prof_every <= 1 || c->prof_seen++ % c->prof_every == 0; }
    ~Timed() {
        if (on) { b = ev_get(c); hipEventRecord(b, c->stream); c->pending.push_back({name, a, b}); }
    }
};
void prof_drain(tj_ctx* c) {
    for (auto& r : c->pending) {
        hipEventSynchronize(r.b);
        float ms = 0.f;
        hipEventElapsedTime(&ms, r.a, r.b);
        auto& e = c->acc[r.name];
        e.first += 1;
        e.second += ms;
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->pending.clear();
}
}  // namespace


namespace {

// BoxArray::maxSize of one domain box along one dimension: reduce by powers
// of two, cut into nearly equal blocks, larger blocks first from the high end.
std::vector<int> chop_dim(int n, int chunk) {
    std::vector<int> starts;
    if (n <= chunk) return {0};
    int ratio = 1, bs = chunk, nlen = n;
    while (bs % 2 == 0 && nlen % 2 == 0) { ratio *= 2; bs /= 2; nlen /= 2; }
    int numblk = nlen / bs + (nlen % bs ? 1 : 0);
    int sz = nlen / numblk, extra = nlen % numblk;
    int hi = n - 1;
    for (int k = 0; k < numblk - 1; ++k) {
        int ksize = (k < extra ? sz + 1 : sz) * ratio;
        int pos = hi - ksize + 1;
        starts.push_back(pos);
        hi = pos - 1;
    }
    starts.push_back(0);
    std::sort(starts.begin(), starts.end());
    return starts;
}

// AccessPattern (accesspattern.cpp:44-79); the mixture block (RHO_MIX,
// LAMBDA, ALPHARHOLAMBDA) when a material is a MixtureEOS.
int build_layout(const tj_level_desc& d, int cs[NVAR]) {
    for (int v = 0; v < NVAR; ++v) cs[v] = -1;
    int n = 0;
    auto add = [&](Var v, int cnt) { cs[v] = n; n += cnt; };
    add(ALPHA, d.nmat);
    add(ALPHARHO, d.nmat);
    add(RHO_K, d.nmat);
    add(RHO, 1);
    add(RHOU, 3);
    add(VELOCITY, 3);
    add(TOTAL_E, 1);
    add(PRESSURE, 1);
    add(SOUNDSPEED, 1);
    add(USTAR, 1);
    add(SIGMA, 9);
    int nmix = 0;
    for (int m = 0; m < d.nmat; ++m) nmix |= d.mat[m].mixture != 0;
    if (nmix) {  // every mixture uses mixtureIndex 0 (initialise.cpp:187)
        add(RHO_MIX, 2);
        add(LAMBDA, 1);
        add(ALPHARHOLAMBDA, 1);
    }
    if (d.SOLID) {
        add(V_TENSOR, 9);
        add(VSTAR, 9);
        add(DEVH, 9);
        add(HJ2, 1);
        if (d.PLASTIC) {
            add(EPSILON, d.nmat);
            add(ALPHARHOEPSILON, d.nmat);
        }
    }
    return n;
}

void free_all(tj_ctx* c) {
    for (int d = 0; d < 3; ++d) {
        if (c->d_edge[d]) hipFree(c->d_edge[d]);
        if (c->d_flux[d]) hipFree(c->d_flux[d]);
        if (c->d_fluxf[d]) hipFree(c->d_fluxf[d]);
        for (int q = 0; q < 2; ++q)
            if (c->d_estar[d][q]) hipFree(c->d_estar[d][q]);
        c->d_edge[d] = nullptr;
        c->d_flux[d] = nullptr;
        c->d_fluxf[d] = nullptr;
        c->d_estar[d][0] = c->d_estar[d][1] = nullptr;
    }
    if (c->d_urstar) hipFree(c->d_urstar);
    for (int q = 0; q < 2; ++q)
        if (c->d_star[q]) hipFree(c->d_star[q]);
    c->d_urstar = c->d_star[0] = c->d_star[1] = nullptr;
    if (c->d_ux) hipFree(c->d_ux);
    c->d_ux = nullptr;
    c->bound = false;
}

int zero(tj_ctx* c, double* p, long long n) {
    if (!p || n <= 0) return TJ_OK;
    unsigned g = grid1(n, 256);
    if (g > 65535u * 16u) g = 65535u * 16u;
    hipLaunchKernelGGL(k_zero, dim3(g), dim3(256), 0, c->stream, p, n);
    HIPCHK(c, hipGetLastError());
    return TJ_OK;
}

}  // namespace

// The library is compiled twice (see tj_device.h): this file as it stands, and
// with TJ_FAB_VIEWS, where MF is a MultiFab of one FAB per box.  In the second
// compilation the entry points below are C++ functions of namespace tjfab;
// here each entry point taking state views forwards to its tjfab twin when the
// views name tj_fabsets.
// An entry point calling another: the one of its own compilation (qualified,
// since argument-dependent lookup on tj_ctx* also finds the global ones).
#ifdef TJ_FAB_VIEWS
#define TJ_SELF(f) tjfab::f
#else
#define TJ_SELF(f) ::f
#endif
#ifdef TJ_FAB_VIEWS
#define TJ_FORWARD_FABS(c, call, ...) \
    do {                              \
    } while (0)
namespace tjfab {
#else
namespace tjfab {
int tj_compute_dt(tj_ctx*, const tj_mf*, double*);
int tj_rk_stage(tj_ctx*, int, const tj_mf*, const tj_mf*, tj_mf*, int);
int tj_post_step(tj_ctx*, tj_mf*, unsigned);
int tj_step(tj_ctx*, tj_mf*, tj_mf*, int);
int tj_begin_run(tj_ctx*, const tj_mf*, double);
int tj_advance_steps(tj_ctx*, tj_mf*, tj_mf*, int, int);
int tj_graph_capture(tj_ctx*, tj_mf*, tj_mf*, int, int);
int tj_run(tj_ctx*, tj_mf*, tj_mf*, double, int, int, int, double*, int*);
int tj_init_from_prims(tj_ctx*, tj_mf*, int);
int tj_c2p(tj_ctx*, tj_mf*);
int tj_lincomb(tj_ctx*, const tj_mf*, double, const tj_mf*, double, tj_mf*);
int tj_derive(tj_ctx*, tj_mf*);
}  // namespace tjfab
namespace {
// 0: every view is a level array, 1: every view names a tj_fabset, -1: mixed
int view_form(std::initializer_list<const tj_mf*> ms) {
    int f = -2;
    for (const tj_mf* m : ms) {
        if (!m) continue;
        const int g = m->fabs ? 1 : 0;
        if (f == -2) f = g;
        else if (f != g) return -1;
    }
    return f < 0 ? 0 : f;
}
}  // namespace
#define TJ_FORWARD_FABS(c, call, ...)                                                         \
    do {                                                                                     \
        const int f_ = view_form({__VA_ARGS__});                                             \
        if (f_ < 0) {                                                                        \
            (c)->err = "the views of one call must all be level arrays or all tj_fabsets";   \
            return TJ_EINVAL;                                                                \
        }                                                                                    \
        if (f_ > 0) return tjfab::call;                                                      \
    } while (0)
extern "C" {
#endif

int tj_create(tj_ctx** out, int device) {
    if (!out) return TJ_EINVAL;
    tj_ctx* c = new tj_ctx();
    c->device = device;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) { delete c; return TJ_EHIP; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return TJ_EHIP; }
    if (hipMalloc(&c->d_clk, sizeof(Clock)) != hipSuccess) { delete c; return TJ_EHIP; }
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c->ncu < 1)
        c->ncu = 1;
    Clock init{};
    init.final_time = __builtin_huge_val();
    init.dtmin_bits = 0x7FEFFFFFFFFFFFFFull;
    if (hipMemcpy(c->d_clk, &init, sizeof(Clock), hipMemcpyHostToDevice) != hipSuccess) { delete c; return TJ_EHIP; }
    if (hipMalloc(&c->d_dtpart, sizeof(unsigned long long) * DT_PARTS) != hipSuccess) { delete c; return TJ_EHIP; }
    hipLaunchKernelGGL(k_dtpart_reset, dim3(1), dim3(64), 0, c->stream, c->d_dtpart);
    if (hipMalloc(&c->d_ticket, sizeof(unsigned long long)) != hipSuccess) { delete c; return TJ_EHIP; }
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, c->stream, reinterpret_cast<double*>(c->d_ticket), 1ll);
    if (hipStreamSynchronize(c->stream) != hipSuccess) { delete c; return TJ_EHIP; }
    *out = c;
    return TJ_OK;
}

static void graph_clear(tj_ctx* c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

void tj_destroy(tj_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    graph_clear(c);
    free_all(c);
    prof_drain(c);
    for (hipEvent_t e : c->pool) hipEventDestroy(e);
    if (c->d_clk) hipFree(c->d_clk);
    if (c->d_dtpart) hipFree(c->d_dtpart);
    if (c->d_ticket) hipFree(c->d_ticket);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char* tj_last_error(const tj_ctx* c) { return c ? c->err.c_str() : "null context"; }
void* tj_stream(tj_ctx* c) { return c ? (void*)c->stream : nullptr; }

// Validates a level description and fills the device-side Level (without
// the edge-flag pointers).  Allocates nothing: the error is a static string.
static int level_from_desc(const char*& err, const tj_level_desc* d, Level& lv) {
    if (!d) return TJ_EINVAL;
    if (d->dim < 1 || d->dim > 3 || d->nmat < 1 || d->nmat > TJ_MAX_MATERIALS) {
        err = "bind_level: bad dim or nmat";
        return TJ_EINVAL;
    }
    if (d->nmat != 2) {
        err = "bind_level: this build instantiates two-material levels";
        return TJ_EUNSUPPORTED;
    }
    if (d->SOLID && d->THINC) {
        err = "bind_level: THINC with SOLID is not available in this build";
        return TJ_EUNSUPPORTED;
    }
    if (d->THINC && d->dim == 3) {
        err = "bind_level: THINC is defined for 1D and 2D levels only (thinc.cpp:151-162)";
        return TJ_EUNSUPPORTED;
    }
    if (!d->MUSCL) {
        err = "bind_level: MUSCL = 0 is not available in this build";
        return TJ_EUNSUPPORTED;
    }
    int nmixm = 0;
    for (int m = 0; m < d->nmat; ++m) nmixm += d->mat[m].mixture != 0;
    if (nmixm > 1) {
        // the reference addresses every mixture's lambda and phase densities
        // at mixtureIndex 0, so a second mixture would alias the first
        err = "bind_level: at most one mixture material (all mixtures share mixtureIndex 0, initialise.cpp:187)";
        return TJ_EUNSUPPORTED;
    }
    for (int m = 0; m < d->nmat; ++m) {
        const tj_material& M = d->mat[m];
        if (M.eos < 0 || M.eos > 3 || (M.eos >= 1 && !M.solid)) {
            err = "bind_level: EOS kinds are 0 (MieGruneisen), 1 (Us-up), 2 (Romenskii), 3 (Wilkins); 1-3 are solids";
            return TJ_EINVAL;
        }
        if (M.eos >= 1 && !d->SOLID) {
            err = "bind_level: a solid EOS needs a SOLID level";
            return TJ_EINVAL;
        }
    }
    if (nmixm && d->SOLID) {
        err = "bind_level: mixtures are available on fluid levels";
        return TJ_EUNSUPPORTED;
    }
    if (d->RADIAL && d->SOLID && !d->PLASTIC) {
        // the source reads the stored stress, which on an elastic level is the
        // RK average of two states' stresses (main.cpp:149-162), not the stress
        // of the averaged state this library can recompute
        err = "bind_level: RADIAL on a solid level needs PLASTIC in this build";
        return TJ_EUNSUPPORTED;
    }
    lv = Level{};
    lv.dim = d->dim;
    for (int a = 0; a < 3; ++a) {
        lv.n[a] = a < d->dim ? d->n[a] : 1;
        lv.dx[a] = a < d->dim ? d->L[a] / d->n[a] : 1.0;
        lv.bc_lo[a] = a < d->dim ? d->bc_lo[a] : 0;
        lv.bc_hi[a] = a < d->dim ? d->bc_hi[a] : 0;
        if (lv.n[a] < 1) { err = "bind_level: n < 1"; return TJ_EINVAL; }
    }
    lv.yn[0] = -1.0 / (8.0 * lv.dx[0]);
    lv.yn[1] = -1.0 / (8.0 * lv.dx[1]);
    lv.nmat = d->nmat;
    lv.solid = d->SOLID;
    lv.plastic = d->PLASTIC;
    lv.thinc = d->THINC;
    lv.beta0 = d->THINCbeta;
    lv.cfl = d->CFL;
    lv.mixm = -1;
    for (int m = 0; m < d->nmat; ++m) {
        const tj_material& M = d->mat[m];
        MatC& C = lv.m[m];
        C.eos = M.eos;
        C.solid = M.solid;
        C.plastic = M.plastic;
        C.Gg = M.gamma - 1.0;
        C.gp1 = C.Gg + 1.0;
        C.xi = 1.0 / C.Gg;
        C.pref = M.pref;
        C.eref = M.eref;
        C.CV = M.CV;
        C.rho0 = M.rho0;
        C.c0 = M.c0;
        C.s = M.s;
        C.G0 = M.G0;
        C.yield = M.yield_stress;
        C.mix = M.mixture != 0;
        C.ga = M.gamma;
        C.g2 = M.gamma_mix;
        C.Gg2 = M.gamma_mix - 1.0;
        C.pref2 = M.pref_mix;
        C.eref2 = M.eref_mix;
        C.CV2 = M.CV_mix;
        C.K0 = M.K0;
        C.ea = M.eos_alpha;
        C.eb = M.eos_beta;
        C.e1 = M.e1;
        C.e2 = M.e2;
        C.e3 = M.e3;
        C.e4 = M.e4;
        C.e5 = M.e5;
        if (C.mix) lv.mixm = m;
        MatT<float>& F = lv.mf[m];
        F.eos = C.eos;
        F.solid = C.solid;
        F.plastic = C.plastic;
        F.Gg = (float)C.Gg;
        F.gp1 = (float)C.gp1;
        F.xi = (float)C.xi;
        F.pref = (float)C.pref;
        F.eref = (float)C.eref;
        F.CV = (float)C.CV;
        F.rho0 = (float)C.rho0;
        F.c0 = (float)C.c0;
        F.s = (float)C.s;
        F.G0 = (float)C.G0;
        F.yield = (float)C.yield;
        F.mix = C.mix;
        F.ga = (float)C.ga;
        F.g2 = (float)C.g2;
        F.Gg2 = (float)C.Gg2;
        F.pref2 = (float)C.pref2;
        F.eref2 = (float)C.eref2;
        F.CV2 = (float)C.CV2;
        F.K0 = (float)C.K0;
        F.ea = (float)C.ea;
        F.eb = (float)C.eb;
        F.e1 = (float)C.e1;
        F.e2 = (float)C.e2;
        F.e3 = (float)C.e3;
        F.e4 = (float)C.e4;
        F.e5 = (float)C.e5;
    }
    build_layout(*d, lv.cs);
    lv.ncells = (long long)lv.n[0] * lv.n[1] * lv.n[2];
    return TJ_OK;
}

// level_from_desc, and edges[a] receives the box-edge flags of coordinate a
// (bit 0: first cell of a box, bit 1: last cell of a box).
static int make_level(std::string& err, const tj_level_desc* d, Level& lv, std::vector<uint8_t> edges[3]) {
    const char* e = nullptr;
    const int r = level_from_desc(e, d, lv);
    if (r) {
        if (e) err = e;
        return r;
    }
    for (int a = 0; a < 3; ++a) {
        std::vector<uint8_t>& fl = edges[a];
        fl.assign(lv.n[a], 0);
        std::vector<int> st = a < d->dim ? chop_dim(lv.n[a], d->max_grid_size) : std::vector<int>{0};
        for (size_t b = 0; b < st.size(); ++b) {
            int lo = st[b];
            int hi = (b + 1 < st.size() ? st[b + 1] : lv.n[a]) - 1;
            fl[lo] |= 1;
            fl[hi] |= 2;
        }
    }
    return TJ_OK;
}

// A level the stage kernels serve: two-dimensional fluid, no mixture.
static bool stage2d_level(const tj_ctx* c) { return !c->lv.solid && c->lv.dim == 2 && c->lv.mixm < 0; }

int tj_bind_level(tj_ctx* c, const tj_level_desc* d) {
    if (!c || !d) return TJ_EINVAL;
    Level lvn{};
    std::vector<uint8_t> edges[3];
    int r;
    if ((r = make_level(c->err, d, lvn, edges))) return r;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_all(c);
    c->seg_cache.clear();
    c->desc = *d;
    Level& lv = c->lv;
    lv = lvn;
    lv.dtpart = c->d_dtpart;
    c->nmat = d->nmat;
    for (int a = 0; a < 3; ++a) {
        const std::vector<uint8_t>& fl = edges[a];
        HIPCHK(c, hipMalloc(&c->d_edge[a], fl.size()));
        HIPCHK(c, hipMemcpy(c->d_edge[a], fl.data(), fl.size(), hipMemcpyHostToDevice));
        lv.edge[a] = c->d_edge[a];
    }
    c->nfc = d->SOLID ? 3 * d->nmat + 17 : 2 * d->nmat + 5 + (lv.mixm >= 0 ? 1 : 0);
    for (int a = 0; a < d->dim; ++a) {
        Faces& F = c->F[a];
        F.nx = lv.n[0] + (a == 0);
        F.ny = lv.n[1] + (a == 1);
        F.nz = lv.n[2] + (a == 2);
        F.cs = (long long)F.nx * F.ny * F.nz;
        HIPCHK(c, hipMalloc(&c->d_flux[a], sizeof(double) * F.cs * c->nfc));
        F.p = c->d_flux[a];
        c->estride[a] = F.cs;
    }
    if (stage2d_level(c)) HIPCHK(c, hipMalloc(&c->d_ux, sizeof(double) * lv.ncells * (2 * d->nmat + 4)));
    if (d->THINC) {
        const long long pl = lv.ncells * d->nmat;
        HIPCHK(c, hipMalloc(&c->d_urstar, sizeof(double) * pl));
        for (int q = 0; q < 2; ++q) HIPCHK(c, hipMalloc(&c->d_star[q], sizeof(double) * pl));
        for (int a = 0; a < d->dim; ++a)
            for (int q = 0; q < 2; ++q) HIPCHK(c, hipMalloc(&c->d_estar[a][q], sizeof(double) * c->estride[a] * d->nmat));
    }
    c->bound = true;
    return ::tj_reset_history(c);
}

int tj_reset_history(tj_ctx* c) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    c->pass = 0;
    c->dpass[0] = c->dpass[1] = c->dpass[2] = 0;
    if (!c->desc.THINC) return TJ_OK;
    const long long pl = c->lv.ncells * c->nmat;
    int r;
    if ((r = zero(c, c->d_urstar, pl))) return r;
    for (int q = 0; q < 2; ++q)
        if ((r = zero(c, c->d_star[q], pl))) return r;
    for (int a = 0; a < c->desc.dim; ++a)
        for (int q = 0; q < 2; ++q)
            if ((r = zero(c, c->d_estar[a][q], c->estride[a] * c->nmat))) return r;
    return TJ_OK;
}

int tj_ncomp(const tj_ctx* c) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    int cs[NVAR];
    return build_layout(c->desc, cs);
}

int tj_comp_start(const tj_ctx* c, int v) {
    if (!c || !c->bound || v < 0 || v >= NVAR) return -1;
    return c->lv.cs[v];
}


// A solid level holding a Romenskii or Wilkins material runs the solid kernels'
// RW instantiation; Us-up levels run the one without their pow/log code.
static bool rw_level(const Level& lv) {
    for (int m = 0; m < lv.nmat; ++m)
        if (lv.m[m].eos >= 2) return true;
    return false;
}

static void launch_sound_dt(tj_ctx* c, const MF& U) {
    const unsigned g = grid1(c->lv.ncells, 256);
    if (c->lv.solid && rw_level(c->lv)) hipLaunchKernelGGL((k_sound_dt_s<2, true>), dim3(g), dim3(256), 0, c->stream, c->lv, U, c->d_clk);
    else if (c->lv.solid) hipLaunchKernelGGL((k_sound_dt_s<2, false>), dim3(g), dim3(256), 0, c->stream, c->lv, U, c->d_clk);
    else if (c->lv.mixm >= 0) hipLaunchKernelGGL((k_sound_dt<2, true>), dim3(g), dim3(256), 0, c->stream, c->lv, U, c->d_clk);
    else hipLaunchKernelGGL(k_sound_dt<2>, dim3(g), dim3(256), 0, c->stream, c->lv, U, c->d_clk);
}

extern "C++" {
template <typename T>
static void launch_update_t(tj_ctx* c, int stage, const MF& in, const MF& base, const MF& out, unsigned ops,
                            const FacesT<T>* F) {
    const unsigned g = grid1(c->lv.ncells, 256);
    if (c->lv.solid) {
        const bool rw = rw_level(c->lv);
        if (stage == 1 && rw)
            hipLaunchKernelGGL((k_update_s<2, 1, true, T>), dim3(g), dim3(256), 0, c->stream, c->lv, in, base, out, F[0], F[1], F[2], c->d_clk, ops);
        else if (stage == 1)
            hipLaunchKernelGGL((k_update_s<2, 1, false, T>), dim3(g), dim3(256), 0, c->stream, c->lv, in, base, out, F[0], F[1], F[2], c->d_clk, ops);
        else if (rw)
            hipLaunchKernelGGL((k_update_s<2, 2, true, T>), dim3(g), dim3(256), 0, c->stream, c->lv, in, base, out, F[0], F[1], F[2], c->d_clk, ops);
        else
            hipLaunchKernelGGL((k_update_s<2, 2, false, T>), dim3(g), dim3(256), 0, c->stream, c->lv, in, base, out, F[0], F[1], F[2], c->d_clk, ops);
    } else if (c->lv.mixm >= 0) {
        if (stage == 1)
            hipLaunchKernelGGL((k_update<2, 1, T, true>), dim3(g), dim3(256), 0, c->stream, c->lv, in, base, out, F[0], F[1], F[2], c->d_clk, ops);
        else
            hipLaunchKernelGGL((k_update<2, 2, T, true>), dim3(g), dim3(256), 0, c->stream, c->lv, in, base, out, F[0], F[1], F[2], c->d_clk, ops);
    } else {
        if (stage == 1)
            hipLaunchKernelGGL((k_update<2, 1, T>), dim3(g), dim3(256), 0, c->stream, c->lv, in, base, out, F[0], F[1], F[2], c->d_clk, ops);
        else
            hipLaunchKernelGGL((k_update<2, 2, T>), dim3(g), dim3(256), 0, c->stream, c->lv, in, base, out, F[0], F[1], F[2], c->d_clk, ops);
    }
}

}  // extern "C++"

static void launch_update(tj_ctx* c, int stage, const MF& in, const MF& base, const MF& out, unsigned ops, int prec) {
    if (prec == TJ_PREC_MIXED) launch_update_t<float>(c, stage, in, base, out, ops, c->Ff);
    else launch_update_t<double>(c, stage, in, base, out, ops, c->F);
}

// FP32 flux planes for the mixed-precision form, allocated on first use.
static int ensure_mixed(tj_ctx* c) {
    if (c->d_fluxf[0]) return TJ_OK;
    for (int a = 0; a < c->lv.dim; ++a) {
        FacesT<float>& F = c->Ff[a];
        F.nx = c->F[a].nx;
        F.ny = c->F[a].ny;
        F.nz = c->F[a].nz;
        F.cs = c->F[a].cs;
        HIPCHK(c, hipMalloc(&c->d_fluxf[a], sizeof(float) * F.cs * c->nfc));
        F.p = c->d_fluxf[a];
    }
    return TJ_OK;
}

static int check_prec(tj_ctx* c, int prec) {
    if (prec != TJ_PREC_FP64 && prec != TJ_PREC_MIXED) { c->err = "unknown precision form"; return TJ_EINVAL; }
    return prec == TJ_PREC_MIXED ? ensure_mixed(c) : TJ_OK;
}

// post-step operations of the reference's loop body for the bound level
static unsigned step_ops(const tj_ctx* c) {
    unsigned ops = TJ_OP_CLEAN_ALPHA | TJ_OP_DT | TJ_OP_NANCHECK;
    if (c->desc.PLASTIC) ops |= TJ_OP_PLASTIC;
    if (c->desc.SOLID) ops |= TJ_OP_CLEAN_V;
    if (c->desc.RADIAL) ops |= TJ_OP_RADIAL;
    if (c->desc.REACTIVE && c->lv.mixm >= 0) ops |= TJ_OP_REACTIVE;  // reactive.cpp:64
    return ops;
}

int tj_compute_dt(tj_ctx* c, const tj_mf* U, double* dt_out) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    TJ_FORWARD_FABS(c, tj_compute_dt(c, U, dt_out), U);
    int r = check_mf(c, U);
    if (r) return r;
    HIPCHK(c, hipSetDevice(c->device));
    launch_sound_dt(c, view(U));
    hipLaunchKernelGGL(k_clock_dt, dim3(1), dim3(64), 0, c->stream, c->d_clk, c->lv.cfl, 0, c->d_dtpart);
    HIPCHK(c, hipGetLastError());
    if (dt_out) {
        Clock h;
        HIPCHK(c, hipMemcpyAsync(&h, c->d_clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        *dt_out = h.dt;
    }
    return TJ_OK;
}

int tj_set_dt(tj_ctx* c, double dt, int reset_clock) {
    if (!c) return TJ_EINVAL;
    Clock h;
    HIPCHK(c, hipMemcpyAsync(&h, c->d_clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    h.dt = dt;
    if (reset_clock) {
        h.t = 0.0;
        h.n = 0;
        h.done = 0;
        h.status = 0;
        h.final_time = __builtin_huge_val();
        h.dtmin_bits = 0x7FEFFFFFFFFFFFFFull;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_clk, &h, sizeof(Clock), hipMemcpyHostToDevice, c->stream));
    if (reset_clock && c->d_dtpart) hipLaunchKernelGGL(k_dtpart_reset, dim3(1), dim3(64), 0, c->stream, c->d_dtpart);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TJ_OK;
}

int tj_get_clock(tj_ctx* c, double* t, double* dt, int* n) {
    if (!c) return TJ_EINVAL;
    Clock h;
    HIPCHK(c, hipMemcpyAsync(&h, c->d_clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (t) *t = h.t;
    if (dt) *dt = h.dt;
    if (n) *n = h.n;
    return TJ_OK;
}

int tj_driver_dt(tj_ctx* c, double final_time) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    hipLaunchKernelGGL(k_clock_final, dim3(1), dim3(64), 0, c->stream, c->d_clk, final_time);
    hipLaunchKernelGGL(k_clock_dt, dim3(1), dim3(64), 0, c->stream, c->d_clk, c->lv.cfl, 1, c->d_dtpart);
    HIPCHK(c, hipGetLastError());
    return TJ_OK;
}

// The update folded into the last directional pass: stage, RK base, output
// and post-step ops of the stage.
struct FuseArgs {
    int stage;
    MF base, out;
    unsigned ops;
};

// THINC history planes of direction d for the current pass.
static Hist hist_for(tj_ctx* c, int d) {
    Hist h{};
    if (!c->desc.THINC) return h;
    const int q = (int)(c->pass & 1), e = (int)(c->dpass[d] & 1);
    h.urstar = c->d_urstar;
    h.star_in = c->d_star[q];
    h.star_out = c->d_star[q ^ 1];
    h.edge_in = c->d_estar[d][e];
    h.edge_out = c->d_estar[d][e ^ 1];
    h.pstride = c->lv.ncells;
    h.estride = c->estride[d];
    return h;
}

// Rows per segment of a marching pass along d.  The pass's waves are numbered
// strip fastest, then segment, then the outer index, four to a block, and a
// block holds one of `slots` block slots: compute units x blocks per unit at
// the kernel's occupancy, at most two (eight waves per unit; measured on
// gfx950, these latency-bound passes gain nothing from more resident waves).
// A wave's time is taken as its segment length + 2 (the row above the segment
// and its set-up), a block's as its slowest wave's.  Of the seglen values
// 8..24 the one with the shortest estimated pass time is returned (ties: the
// one closest to 16):
//  * up to two blocks per slot at seglen 16, max(total work / slots, longest
//    block): blocks of unequal work (interface cells) let a short tail round
//    fill in behind the first, and a level too small for the device gets
//    shorter segments so that more waves run;
//  * beyond that, the makespan of list scheduling in launch order, which
//    favours segment counts that divide into whole rounds;
//  * a pass much larger than the device keeps 16.
// The choice changes no value a pass computes, only how rows are split among
// waves.
static int pick_seglen(long long nstrip, int units, long long no, long long slots) {
    const long long nb16 = (nstrip * ((units + 15) / 16) * no + 3) / 4;
    if (nb16 > 64 * slots) return 16;
    const bool few_rounds = nb16 <= 2 * slots;
    int best = 16;
    double best_t = -1.0;
    for (int sl = 8; sl <= 24; ++sl) {
        const long long nseg = (units + sl - 1) / sl;
        const long long nw = nstrip * nseg * no;
        const long long nb = (nw + 3) / 4;
        auto work = [&](long long b) {
            long long w = 0;
            for (long long q = 4 * b; q < std::min(4 * b + 4, nw); ++q) {
                const long long seg = (q / nstrip) % nseg;
                w = std::max(w, std::min<long long>(sl, units - seg * sl) + 2);
            }
            return w;
        };
        double t = 0.0;
        if (few_rounds) {
            long long total = 0, longest = 0;
            for (long long b = 0; b < nb; ++b) {
                const long long w = work(b);
                total += w;
                longest = std::max(longest, w);
            }
            t = std::max((double)total / (double)slots, (double)longest);
        } else {
            std::priority_queue<long long, std::vector<long long>, std::greater<long long>> free_at;
            for (long long q = 0; q < slots; ++q) free_at.push(0);
            long long end_max = 0;
            for (long long b = 0; b < nb; ++b) {
                const long long start = free_at.top();
                free_at.pop();
                const long long end = start + work(b);
                end_max = std::max(end_max, end);
                free_at.push(end);
            }
            t = (double)end_max;
        }
        if (best_t < 0.0 || t < best_t || (t == best_t && std::abs(sl - 16) < std::abs(best - 16))) {
            best = sl;
            best_t = t;
        }
    }
    return best;
}

// seglen of one marching kernel along d: the caller's fixed value, or
// pick_seglen at the kernel's occupancy (blocks of 256 threads, static LDS).
static int seg_for(tj_ctx* c, const void* kernel, int d, bool fused) {
    if (c->seglen > 0) return c->seglen;
    const auto key = std::make_pair(kernel, d * 2 + (fused ? 1 : 0));
    const auto it = c->seg_cache.find(key);
    if (it != c->seg_cache.end()) return it->second;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    const Level& lv = c->lv;
    const long long nstrip = (lv.n[0] + 63) / 64;
    const long long no = d == 1 ? lv.n[2] : lv.n[1];
    const int units = fused ? lv.n[d] : lv.n[d] + 1;  // rows owned (fused) or faces computed
    const int sl = pick_seglen(nstrip, units, no, (long long)c->ncu * std::min(per_cu, 2));
    c->seg_cache[key] = sl;
    return sl;
}
// grid of a marching pass along d with segments of sl rows (faces, or rows
// owned when the update is fused)
static unsigned march_grid(const Level& lv, int d, int sl, bool fused) {
    const long long nstrip = (lv.n[0] + 63) / 64;
    const long long nseg = fused ? (lv.n[d] + sl - 1) / sl : (lv.n[d] + 1 + sl - 1) / sl;
    const long long no = d == 1 ? lv.n[2] : lv.n[1];
    return grid1(nstrip * nseg * no, 4);
}
#define TJ_KPTR(...) reinterpret_cast<const void*>(&__VA_ARGS__)

// One directional flux pass (not fused) with face values of type R.
extern "C++" {
template <typename R, bool RW>
static void launch_pass_solid(tj_ctx* c, const MF& Uin, int d, const FacesT<R>* F) {
    const Level& lv = c->lv;
    if (d == 0) {
        const long long nseg = (lv.n[0] + 1 + 62) / 63;
        const unsigned g = grid1(nseg * lv.n[1] * lv.n[2], 4);
        hipLaunchKernelGGL((k_flux_x_s<2, RW, R>), dim3(g), dim3(256), 0, c->stream, lv, Uin, F[0], c->d_clk);
    } else if (d == 1) {
        const int sl = seg_for(c, TJ_KPTR(k_flux_lmarch_s<2, 1, RW, R>), 1, false);
        hipLaunchKernelGGL((k_flux_lmarch_s<2, 1, RW, R>), dim3(march_grid(lv, 1, sl, false)), dim3(256), 0, c->stream,
                           lv, Uin, F[1], sl, c->d_clk);
    } else {
        const int sl = seg_for(c, TJ_KPTR(k_flux_lmarch_s<2, 2, RW, R>), 2, false);
        hipLaunchKernelGGL((k_flux_lmarch_s<2, 2, RW, R>), dim3(march_grid(lv, 2, sl, false)), dim3(256), 0, c->stream,
                           lv, Uin, F[2], sl, c->d_clk);
    }
}
template <typename R, bool MX>
static void launch_pass_t(tj_ctx* c, const MF& Uin, int d, const FacesT<R>* F, const Hist& h) {
    const Level& lv = c->lv;
    const bool th = c->desc.THINC != 0;
    if (lv.solid) {
        if (rw_level(lv)) launch_pass_solid<R, true>(c, Uin, d, F);
        else launch_pass_solid<R, false>(c, Uin, d, F);
    } else if (d == 0) {
        const int per = th ? 61 : 63;
        const long long nseg = (lv.n[0] + 1 + per - 1) / per;
        const unsigned g = grid1(nseg * lv.n[1] * lv.n[2], 4);
        if (th) hipLaunchKernelGGL((k_flux_x<2, true, R, MX>), dim3(g), dim3(256), 0, c->stream, lv, Uin, F[0], h, c->d_clk);
        else hipLaunchKernelGGL((k_flux_x<2, false, R, MX>), dim3(g), dim3(256), 0, c->stream, lv, Uin, F[0], h, c->d_clk);
    } else {
        const MF z{};
        const FacesT<R> zf{};
        if (d == 1 && th && sizeof(R) == 8) {
            const int sl = seg_for(c, TJ_KPTR(k_flux_march_w2<2, true, 1, R, MX>), 1, false);
            hipLaunchKernelGGL((k_flux_march_w2<2, true, 1, R, MX>), dim3(march_grid(lv, 1, sl, false)), dim3(256), 0,
                               c->stream, lv, Uin, F[1], h, sl, c->d_clk);
        } else if (d == 1 && th) {
            const int sl = seg_for(c, TJ_KPTR(k_flux_march<2, true, 1, false, 1, R, MX>), 1, false);
            hipLaunchKernelGGL((k_flux_march<2, true, 1, false, 1, R, MX>), dim3(march_grid(lv, 1, sl, false)), dim3(256), 0,
                               c->stream, lv, Uin, F[1], h, sl, c->d_clk, z, z, zf, zf, 0u);
        } else if (d == 1) {
            const int sl = seg_for(c, TJ_KPTR(k_flux_march<2, false, 1, false, 1, R, MX>), 1, false);
            hipLaunchKernelGGL((k_flux_march<2, false, 1, false, 1, R, MX>), dim3(march_grid(lv, 1, sl, false)), dim3(256), 0,
                               c->stream, lv, Uin, F[1], h, sl, c->d_clk, z, z, zf, zf, 0u);
        } else {
            const int sl = seg_for(c, TJ_KPTR(k_flux_march<2, false, 2, false, 1, R, MX>), 2, false);
            hipLaunchKernelGGL((k_flux_march<2, false, 2, false, 1, R, MX>), dim3(march_grid(lv, 2, sl, false)), dim3(256), 0,
                               c->stream, lv, Uin, F[2], h, sl, c->d_clk, z, z, zf, zf, 0u);
        }
    }
}

// the pass for the level's material set (MX: a mixture material)
template <typename R>
static void launch_pass(tj_ctx* c, const MF& Uin, int d, const FacesT<R>* F, const Hist& h) {
    if (c->lv.mixm >= 0) launch_pass_t<R, true>(c, Uin, d, F, h);
    else launch_pass_t<R, false>(c, Uin, d, F, h);
}

}  // extern "C++"

// The last fluid pass with the update folded in (FP64 form).
static void launch_fused(tj_ctx* c, const MF& Uin, int d, const Hist& h, const FuseArgs* fz) {
    const Level& lv = c->lv;
    const bool th = c->desc.THINC != 0;
    const MF b = fz->base, o = fz->out;
    const unsigned ops = fz->ops;
    const bool s2 = fz->stage == 2;
#define TJ_FUSED(TH, D, ST)                                                                                     \
    do {                                                                                                        \
        const int sl = seg_for(c, TJ_KPTR(k_flux_march<2, TH, D, true, ST>), D, true);                          \
        hipLaunchKernelGGL((k_flux_march<2, TH, D, true, ST>), dim3(march_grid(lv, D, sl, true)), dim3(256), 0, \
                           c->stream, lv, Uin, c->F[D], h, sl, c->d_clk, b, o, c->F[0], c->F[1], ops);          \
    } while (0)
    if (d == 1) {
        if (th && s2) TJ_FUSED(true, 1, 2);
        else if (th) TJ_FUSED(true, 1, 1);
        else if (s2) TJ_FUSED(false, 1, 2);
        else TJ_FUSED(false, 1, 1);
    } else {
        if (s2) TJ_FUSED(false, 2, 2);
        else TJ_FUSED(false, 2, 1);
    }
#undef TJ_FUSED
}

// Directional flux passes of one RK stage.  With fz (FP64 form, fluid
// levels, dim >= 2, output distinct from the input, fusion switched on) the
// last pass also completes the update and *fused is set; otherwise the caller
// launches the update kernel.
static int launch_fluxes(tj_ctx* c, const MF& Uin, const FuseArgs* fz, bool* fused, int prec) {
    const Level& lv = c->lv;
    const bool mixed = prec == TJ_PREC_MIXED;
    const bool fuse = fz && c->fuse && !mixed && !lv.solid && lv.mixm < 0 && lv.dim >= 2 && !same_view(fz->out, Uin);
    if (fused) *fused = fuse;
    for (int d = 0; d < lv.dim; ++d) {
        const Hist h = hist_for(c, d);
        const bool last = fuse && d == lv.dim - 1;
        static const char* names[2][3] = {{"flux_x", "flux_y", "flux_z"}, {"flux_x_mixed", "flux_y_mixed", "flux_z_mixed"}};
        static const char* fnames[2][3] = {{"", "flux_y_update1", "flux_z_update1"},
                                           {"", "flux_y_update2_post", "flux_z_update2_post"}};
        Timed tm(c, last ? fnames[fz->stage - 1][d] : names[mixed][d]);
        if (last) launch_fused(c, Uin, d, h, fz);
        else if (mixed) launch_pass<float>(c, Uin, d, c->Ff, h);
        else launch_pass<double>(c, Uin, d, c->F, h);
        HIPCHK(c, hipGetLastError());
        if (c->desc.THINC) {
            c->pass += 1;
            c->dpass[d] += 1;
        }
    }
    return TJ_OK;
}

// One RK stage of a two-dimensional fluid level by the stage kernels
// (tj_stage2d.h): the x kernel, then the y kernel with the update's
// completion (stage 2: RK average and the post-step ops).
extern "C++" {
#ifndef TJ_STAGE_TC
#define TJ_STAGE_TC 16
#endif
// y tiles of 16 columns x 16-lane segments (blocks of four waves): measured
// faster than 16 x 32, 8 x 32 and 32 x 16 on the 2048 x 1024 level
#ifndef TJ_STAGE_LPS
#define TJ_STAGE_LPS 16
#endif
template <bool TH, typename R>
static void launch_stage2d_t(tj_ctx* c, int stage, const MF& in, const MF& base, const MF& out, unsigned ops,
                             const Hist& hx, const Hist& hy, const char* nx_name, const char* ny_name) {
    constexpr int TC = TJ_STAGE_TC;
    const Level& lv = c->lv;
    const UxPlanes ux{c->d_ux, lv.ncells};
    {
        Timed tm(c, nx_name);
        constexpr int CPW = XGeom<TH>::CPW;
        const long long waves = (long long)((lv.n[0] + CPW - 1) / CPW) * lv.n[1];
        hipLaunchKernelGGL((k_xstage<2, TH, R>), dim3(grid1(waves, 4)), dim3(256), 0, c->stream, lv, in, ux, hx, c->d_clk);
    }
    if (TH) {
        c->pass += 1;
        c->dpass[0] += 1;
    }
    {
        Timed tm(c, ny_name);
        constexpr int LPS = TJ_STAGE_LPS;
        constexpr int CPW = YTile<2, TH, TC, LPS, R>::CPW;
        const long long tiles = (long long)((lv.n[0] + TC - 1) / TC) * ((lv.n[1] + CPW - 1) / CPW);
        if (stage == 1)
            hipLaunchKernelGGL((k_ystage<2, TH, TC, LPS, 1, R>), dim3((unsigned)tiles), dim3(TC * LPS), 0, c->stream, lv,
                               in, ux, base, out, hy, c->d_clk, ops);
        else
            hipLaunchKernelGGL((k_ystage<2, TH, TC, LPS, 2, R>), dim3((unsigned)tiles), dim3(TC * LPS), 0, c->stream, lv,
                               in, ux, base, out, hy, c->d_clk, ops);
    }
    if (TH) {
        c->pass += 1;
        c->dpass[1] += 1;
    }
}
}  // extern "C++"

static int launch_stage2d(tj_ctx* c, int stage, const MF& in, const MF& base, const MF& out, unsigned ops, int prec) {
    const bool mixed = prec == TJ_PREC_MIXED;
    static const char* xn[2][2] = {{"xstage1", "xstage2"}, {"xstage1_mixed", "xstage2_mixed"}};
    static const char* yn[2][2] = {{"ystage1", "ystage2_post"}, {"ystage1_mixed", "ystage2_post_mixed"}};
    const Hist hx = hist_for(c, 0);
    // the y kernel's history: the planes after the x kernel's pass
    const long long p0 = c->pass, d1 = c->dpass[1];
    if (c->desc.THINC) c->pass += 1;
    const Hist hy = hist_for(c, 1);
    c->pass = p0;
    c->dpass[1] = d1;
    const char* a = xn[mixed][stage - 1];
    const char* b = yn[mixed][stage - 1];
    if (c->desc.THINC) {
        if (mixed) launch_stage2d_t<true, float>(c, stage, in, base, out, ops, hx, hy, a, b);
        else launch_stage2d_t<true, double>(c, stage, in, base, out, ops, hx, hy, a, b);
    } else {
        if (mixed) launch_stage2d_t<false, float>(c, stage, in, base, out, ops, hx, hy, a, b);
        else launch_stage2d_t<false, double>(c, stage, in, base, out, ops, hx, hy, a, b);
    }
    HIPCHK(c, hipGetLastError());
    return TJ_OK;
}

static bool use_stage2d(const tj_ctx* c) { return c->stage2d && c->d_ux && stage2d_level(c); }

// One-dimensional fluid levels without a mixture: each RK stage is one launch
// of k_xstage1d (the x stage completing its cells); with `tail` stage 2 also
// advances the clock (k_clock_advance's work).
static bool use_stage1d(const tj_ctx* c) {
    return c->stage2d && c->d_ticket && !c->lv.solid && c->lv.dim == 1 && c->lv.mixm < 0;
}
extern "C++" {
template <bool TH, typename R>
static void launch_stage1d_t(tj_ctx* c, int stage, const MF& in, const MF& base, const MF& out, unsigned ops, bool tail,
                             const char* name) {
    const Level& lv = c->lv;
    const Hist h = hist_for(c, 0);
    {
        Timed tm(c, name);
        constexpr int CPW = XGeom<TH>::CPW;
        const unsigned g = grid1((long long)((lv.n[0] + CPW - 1) / CPW) * lv.n[1], 4);
        if (stage == 1)
            hipLaunchKernelGGL((k_xstage1d<2, TH, 1, false, R>), dim3(g), dim3(256), 0, c->stream, lv, in, base, out, h,
                               c->d_clk, ops, c->d_ticket);
        else if (tail)
            hipLaunchKernelGGL((k_xstage1d<2, TH, 2, true, R>), dim3(g), dim3(256), 0, c->stream, lv, in, base, out, h,
                               c->d_clk, ops, c->d_ticket);
        else
            hipLaunchKernelGGL((k_xstage1d<2, TH, 2, false, R>), dim3(g), dim3(256), 0, c->stream, lv, in, base, out, h,
                               c->d_clk, ops, c->d_ticket);
    }
    if (TH) {
        c->pass += 1;
        c->dpass[0] += 1;
    }
}
}  // extern "C++"
static int launch_stage1d(tj_ctx* c, int stage, const MF& in, const MF& base, const MF& out, unsigned ops, bool tail,
                          int prec) {
    const bool mixed = prec == TJ_PREC_MIXED;
    static const char* nm[2][2] = {{"xstage1d_1", "xstage1d_2_post"}, {"xstage1d_1_mixed", "xstage1d_2_post_mixed"}};
    const char* name = nm[mixed][stage - 1];
    if (c->desc.THINC) {
        if (mixed) launch_stage1d_t<true, float>(c, stage, in, base, out, ops, tail, name);
        else launch_stage1d_t<true, double>(c, stage, in, base, out, ops, tail, name);
    } else {
        if (mixed) launch_stage1d_t<false, float>(c, stage, in, base, out, ops, tail, name);
        else launch_stage1d_t<false, double>(c, stage, in, base, out, ops, tail, name);
    }
    HIPCHK(c, hipGetLastError());
    return TJ_OK;
}

int tj_rk_stage(tj_ctx* c, int stage, const tj_mf* Uin, const tj_mf* Ubase, tj_mf* Uout, int prec) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    TJ_FORWARD_FABS(c, tj_rk_stage(c, stage, Uin, Ubase, Uout, prec), Uin, Ubase, Uout);
    int r;
    if ((r = check_prec(c, prec))) return r;
    if ((r = check_mf(c, Uin)) || (r = check_mf(c, Uout))) return r;
    if (stage == 2 && (r = check_mf(c, Ubase))) return r;
    HIPCHK(c, hipSetDevice(c->device));
    const MF in = view(Uin);
    const MF base = stage == 2 ? view(Ubase) : in;
    // the y kernel reads rows of Uin around each tile while other tiles write
    // Uout: an in-place stage (Uout == Uin) takes the flux-plane passes
    if (use_stage2d(c) && !same_view(view(Uout), in)) return launch_stage2d(c, stage, in, base, view(Uout), 0u, prec);
    if (use_stage1d(c) && !same_view(view(Uout), in)) return launch_stage1d(c, stage, in, base, view(Uout), 0u, false, prec);
    const FuseArgs fz{stage, base, view(Uout), 0u};
    bool fused = false;
    if ((r = launch_fluxes(c, in, &fz, &fused, prec))) return r;
    if (!fused) {
        static const char* names[2][2] = {{"update_stage1", "update_stage2"}, {"update_stage1_mixed", "update_stage2_mixed"}};
        Timed tm(c, names[prec == TJ_PREC_MIXED][stage - 1]);
        launch_update(c, stage, in, base, view(Uout), 0u, prec);
    }
    HIPCHK(c, hipGetLastError());
    return TJ_OK;
}

int tj_post_step(tj_ctx* c, tj_mf* U, unsigned ops) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    TJ_FORWARD_FABS(c, tj_post_step(c, U, ops), U);
    int r;
    if ((r = check_mf(c, U))) return r;
    if ((ops & (TJ_OP_PLASTIC | TJ_OP_CLEAN_V)) && !c->lv.solid) {
        c->err = "post_step: PLASTIC and CLEAN_V apply to solid levels";
        return TJ_EINVAL;
    }
    if ((ops & TJ_OP_RADIAL) && c->lv.solid && !c->lv.plastic) {
        c->err = "post_step: RADIAL on a solid level needs a PLASTIC level";
        return TJ_EINVAL;
    }
    const unsigned g = grid1(c->lv.ncells, 256);
    if (c->lv.solid) {
        if (rw_level(c->lv)) hipLaunchKernelGGL((k_post_s<2, true>), dim3(g), dim3(256), 0, c->stream, c->lv, view(U), c->d_clk, ops);
        else hipLaunchKernelGGL((k_post_s<2, false>), dim3(g), dim3(256), 0, c->stream, c->lv, view(U), c->d_clk, ops);
    } else {
        const bool mx = c->lv.mixm >= 0;
        if ((ops & TJ_OP_REACTIVE) && mx)  // reactiveUpdate acts on mixtures only (reactive.cpp:64)
            hipLaunchKernelGGL(k_reactive_fluid<2>, dim3(g), dim3(256), 0, c->stream, c->lv, view(U), c->d_clk);
        if ((ops & TJ_OP_RADIAL) && mx)
            hipLaunchKernelGGL((k_radial_fluid<2, true>), dim3(g), dim3(256), 0, c->stream, c->lv, view(U), c->d_clk);
        else if (ops & TJ_OP_RADIAL)
            hipLaunchKernelGGL(k_radial_fluid<2>, dim3(g), dim3(256), 0, c->stream, c->lv, view(U), c->d_clk);
        if (ops & TJ_OP_CLEAN_ALPHA)
            hipLaunchKernelGGL(k_cleanup_alpha<2>, dim3(g), dim3(256), 0, c->stream, c->lv, view(U), c->d_clk);
        if (ops & TJ_OP_NANCHECK)
        {
            if (mx) hipLaunchKernelGGL((k_nancheck<2, true>), dim3(g), dim3(256), 0, c->stream, c->lv, view(U), c->d_clk);
            else hipLaunchKernelGGL(k_nancheck<2>, dim3(g), dim3(256), 0, c->stream, c->lv, view(U), c->d_clk);
        }
        if (ops & TJ_OP_DT) launch_sound_dt(c, view(U));
    }
    HIPCHK(c, hipGetLastError());
    return TJ_OK;
}

int tj_step(tj_ctx* c, tj_mf* S, tj_mf* T, int prec) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    TJ_FORWARD_FABS(c, tj_step(c, S, T, prec), S, T);
    int r;
    if ((r = check_prec(c, prec))) return r;
    if ((r = check_mf(c, S)) || (r = check_mf(c, T))) return r;
    HIPCHK(c, hipSetDevice(c->device));
    const MF s = view(S), t = view(T);
    if (same_view(s, t)) { c->err = "step: S and T must be distinct buffers"; return TJ_EINVAL; }
    const bool mixed = prec == TJ_PREC_MIXED;
    if (use_stage2d(c)) {
        if ((r = launch_stage2d(c, 1, s, s, t, 0u, prec))) return r;
        if ((r = launch_stage2d(c, 2, t, s, s, step_ops(c), prec))) return r;
        hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, c->stream, c->d_clk);
        HIPCHK(c, hipGetLastError());
        return TJ_OK;
    }
    if (use_stage1d(c)) {  // two launches; the second advances the clock
        if ((r = launch_stage1d(c, 1, s, s, t, 0u, false, prec))) return r;
        return launch_stage1d(c, 2, t, s, s, step_ops(c), true, prec);
    }
    bool fused = false;
    const FuseArgs f1{1, s, t, 0u};
    if ((r = launch_fluxes(c, s, &f1, &fused, prec))) return r;
    if (!fused) {
        Timed tm(c, mixed ? "update_stage1_mixed" : "update_stage1");
        launch_update(c, 1, s, s, t, 0u, prec);
    }
    const FuseArgs f2{2, s, s, step_ops(c)};
    if ((r = launch_fluxes(c, t, &f2, &fused, prec))) return r;
    if (!fused) {
        Timed tm(c, mixed ? "update_stage2_post_mixed" : "update_stage2_post");
        launch_update(c, 2, t, s, s, step_ops(c), prec);
    }
    hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, c->stream, c->d_clk);
    HIPCHK(c, hipGetLastError());
    return TJ_OK;
}

int tj_begin_run(tj_ctx* c, const tj_mf* S, double final_time) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    TJ_FORWARD_FABS(c, tj_begin_run(c, S, final_time), S);
    int r;
    if ((r = check_mf(c, S))) return r;
    if ((r = TJ_SELF(tj_set_dt)(c, 0.0, 1))) return r;
    hipLaunchKernelGGL(k_clock_final, dim3(1), dim3(64), 0, c->stream, c->d_clk, final_time);
    // getTimeStep of the initial state feeds the first driver dt
    launch_sound_dt(c, view(S));
    HIPCHK(c, hipGetLastError());
    return TJ_OK;
}

int tj_advance_steps(tj_ctx* c, tj_mf* S, tj_mf* T, int nsteps, int prec) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    TJ_FORWARD_FABS(c, tj_advance_steps(c, S, T, nsteps, prec), S, T);
    int r;
    for (int q = 0; q < nsteps; ++q) {
        hipLaunchKernelGGL(k_clock_dt, dim3(1), dim3(64), 0, c->stream, c->d_clk, c->lv.cfl, 1, c->d_dtpart);
        if ((r = TJ_SELF(tj_step)(c, S, T, prec))) return r;
    }
    return TJ_OK;
}

int tj_graph_capture(tj_ctx* c, tj_mf* S, tj_mf* T, int nsteps, int prec) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    TJ_FORWARD_FABS(c, tj_graph_capture(c, S, T, nsteps, prec), S, T);
    if (nsteps < 1) { c->err = "graph_capture: nsteps < 1"; return TJ_EINVAL; }
    int r;
    if ((r = check_mf(c, S)) || (r = check_mf(c, T))) return r;
    if ((r = check_prec(c, prec))) return r;  // scratch allocations happen before the capture
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    graph_clear(c);
    const bool prof = c->prof;
    c->prof = false;
    const long long p0 = c->pass, d0[3] = {c->dpass[0], c->dpass[1], c->dpass[2]};
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rr = TJ_SELF(tj_advance_steps)(c, S, T, nsteps, prec);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->prof = prof;
    c->g_pass0 = p0;
    c->g_npass = c->pass - p0;
    for (int a = 0; a < 3; ++a) {
        c->g_dpass0[a] = d0[a];
        c->g_ndpass[a] = c->dpass[a] - d0[a];
        c->dpass[a] = d0[a];
    }
    c->pass = p0;  // the counters advance when the graph is launched
    if (rr) {
        if (e == hipSuccess && g) (void)hipGraphDestroy(g);
        return rr;
    }
    HIPCHK(c, e);
    c->graph = g;
    HIPCHK(c, hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
    return TJ_OK;
}

int tj_graph_launch(tj_ctx* c) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    if (!c->gexec) { c->err = "graph_launch: nothing captured"; return TJ_EINVAL; }
    // the captured passes read and write the THINC history planes of a fixed
    // ping-pong phase: the level must be at that phase again
    bool same = ((c->pass - c->g_pass0) & 1) == 0;
    for (int a = 0; a < 3; ++a) same = same && ((c->dpass[a] - c->g_dpass0[a]) & 1) == 0;
    if (c->desc.THINC && !same) {
        c->err = "graph_launch: THINC history phase differs from the captured one (capture again)";
        return TJ_EINVAL;
    }
    HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    c->pass += c->g_npass;
    for (int a = 0; a < 3; ++a) c->dpass[a] += c->g_ndpass[a];
    return TJ_OK;
}

int tj_run(tj_ctx* c, tj_mf* S, tj_mf* T, double final_time, int max_steps, int prec, int check_every,
           double* t_out, int* n_out) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    TJ_FORWARD_FABS(c, tj_run(c, S, T, final_time, max_steps, prec, check_every, t_out, n_out), S, T);
    int r;
    if ((r = check_mf(c, T))) return r;
    if ((r = TJ_SELF(tj_begin_run)(c, S, final_time))) return r;
    if (check_every < 1) check_every = 1;
    int steps = 0;
    while (steps < max_steps) {
        const int chunk = std::min(check_every, max_steps - steps);
        if ((r = TJ_SELF(tj_advance_steps)(c, S, T, chunk, prec))) return r;
        steps += chunk;
        Clock h;
        HIPCHK(c, hipMemcpyAsync(&h, c->d_clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (h.done) break;
    }
    Clock h;
    HIPCHK(c, hipMemcpyAsync(&h, c->d_clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (t_out) *t_out = h.t;
    if (n_out) *n_out = h.n;
    return TJ_OK;
}

int tj_init_from_prims(tj_ctx* c, tj_mf* U, int shift) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    TJ_FORWARD_FABS(c, tj_init_from_prims(c, U, shift), U);
    int r;
    if ((r = check_mf(c, U))) return r;
    if (c->lv.solid) {
        if (rw_level(c->lv)) hipLaunchKernelGGL((k_init_p2c_s<2, true>), dim3(grid1(c->lv.ncells, 256)), dim3(256), 0, c->stream, c->lv, view(U), shift);
        else hipLaunchKernelGGL((k_init_p2c_s<2, false>), dim3(grid1(c->lv.ncells, 256)), dim3(256), 0, c->stream, c->lv, view(U), shift);
    } else
        if (c->lv.mixm >= 0)
            hipLaunchKernelGGL((k_init_p2c<2, true>), dim3(grid1(c->lv.ncells, 256)), dim3(256), 0, c->stream, c->lv, view(U), shift);
        else
            hipLaunchKernelGGL(k_init_p2c<2>, dim3(grid1(c->lv.ncells, 256)), dim3(256), 0, c->stream, c->lv, view(U), shift);
    HIPCHK(c, hipGetLastError());
    return TJ_OK;
}

int tj_c2p(tj_ctx* c, tj_mf* U) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    TJ_FORWARD_FABS(c, tj_c2p(c, U), U);
    int r;
    if ((r = check_mf(c, U))) return r;
    if (c->lv.solid) {
        if (rw_level(c->lv)) hipLaunchKernelGGL((k_c2p_s<2, true>), dim3(grid1(c->lv.ncells, 256)), dim3(256), 0, c->stream, c->lv, view(U));
        else hipLaunchKernelGGL((k_c2p_s<2, false>), dim3(grid1(c->lv.ncells, 256)), dim3(256), 0, c->stream, c->lv, view(U));
    } else
        if (c->lv.mixm >= 0)
            hipLaunchKernelGGL((k_c2p<2, true>), dim3(grid1(c->lv.ncells, 256)), dim3(256), 0, c->stream, c->lv, view(U), c->d_clk);
        else
            hipLaunchKernelGGL(k_c2p<2>, dim3(grid1(c->lv.ncells, 256)), dim3(256), 0, c->stream, c->lv, view(U), c->d_clk);
    HIPCHK(c, hipGetLastError());
    return TJ_OK;
}

int tj_lincomb(tj_ctx* c, const tj_mf* X, double a, const tj_mf* Y, double b, tj_mf* Z) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    TJ_FORWARD_FABS(c, tj_lincomb(c, X, a, Y, b, Z), X, Y, Z);
    int r;
    if ((r = check_mf(c, X)) || (r = check_mf(c, Z))) return r;
    if (Y && (r = check_mf(c, Y))) return r;
    const int nc = std::min(X->ncomp, Z->ncomp);
    if (Y && Y->ncomp < nc) { c->err = "lincomb: Y has fewer components than X and Z"; return TJ_EINVAL; }
    const long long n = c->lv.ncells * nc;
    const MF y = Y ? view(Y) : view(X);
    hipLaunchKernelGGL(k_lincomb, dim3(grid1(n, 256)), dim3(256), 0, c->stream, view(X), a, y, b, Y ? 1 : 0, view(Z), nc,
                       c->lv.n[0], c->lv.n[1], c->lv.n[2]);
    HIPCHK(c, hipGetLastError());
    return TJ_OK;
}

int tj_derive(tj_ctx* c, tj_mf* U) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    TJ_FORWARD_FABS(c, tj_derive(c, U), U);
    int r;
    if ((r = check_mf(c, U))) return r;
    if (c->lv.solid) {
        if (rw_level(c->lv)) hipLaunchKernelGGL((k_derive_s<2, true>), dim3(grid1(c->lv.ncells, 256)), dim3(256), 0, c->stream, c->lv, view(U));
        else hipLaunchKernelGGL((k_derive_s<2, false>), dim3(grid1(c->lv.ncells, 256)), dim3(256), 0, c->stream, c->lv, view(U));
    } else
        hipLaunchKernelGGL(k_derive_fluid, dim3(grid1(c->lv.ncells, 256)), dim3(256), 0, c->stream, c->lv, view(U));
    HIPCHK(c, hipGetLastError());
    return TJ_OK;
}

int tj_status(tj_ctx* c, uint32_t* status, int clear) {
    if (!c) return TJ_EINVAL;
    Clock h;
    HIPCHK(c, hipMemcpyAsync(&h, c->d_clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (status) *status = h.status;
    if (clear) {
        hipLaunchKernelGGL(k_clock_clear_status, dim3(1), dim3(64), 0, c->stream, c->d_clk);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return TJ_OK;
}

int tj_set_tuning(tj_ctx* c, int seglen, int fuse) {
    if (!c) return TJ_EINVAL;
    if (seglen >= 0) c->seglen = seglen;
    if (fuse >= 0) {
        c->stage2d = fuse == 2;
        c->fuse = fuse == 1;
    }
    return TJ_OK;
}

int tj_prof_enable(tj_ctx* c, int on) {
    if (!c) return TJ_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_drain(c);
    c->acc.clear();
    c->prof = on != 0;
    return TJ_OK;
}

int tj_prof_only(tj_ctx* c, const char* name) {
    if (!c) return TJ_EINVAL;
    c->prof_only = name ? name : "";
    c->prof_seen = 0;
    return TJ_OK;
}

int tj_prof_sample(tj_ctx* c, int every) {
    if (!c || every < 1) return TJ_EINVAL;
    c->prof_every = every;
    c->prof_seen = 0;
    return TJ_OK;
}

int tj_prof_report(tj_ctx* c, char* buf, int len) {
    if (!c || !buf || len < 1) return TJ_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_drain(c);
    std::string out;
    for (auto& kv : c->acc) {
        char line[256];
        snprintf(line, sizeof line, "%s %lld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
        out += line;
    }
    snprintf(buf, (size_t)len, "%s", out.c_str());
    return TJ_OK;
}

int tj_synchronize(tj_ctx* c) {
    if (!c) return TJ_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TJ_OK;
}

#ifndef TJ_FAB_VIEWS
// A MultiFab of one FAB per box (BoxAccessCellArray(const Box&, FArrayBox&,
// CellArray&), boxaccesscellarray.cpp:3-23, for every box at once).
int tj_fabset_create(tj_ctx* c, int nfab, const tj_fab* fabs, int ncomp, tj_fabset** out) {
    if (!out) return TJ_EINVAL;
    *out = nullptr;
    if (!c || !c->bound) return TJ_ENOLEVEL;
    if (nfab < 1 || !fabs) { c->err = "fabset: no FABs"; return TJ_EINVAL; }
    if (ncomp < TJ_SELF(tj_ncomp)(c)) { c->err = "fabset: fewer components than the bound layout"; return TJ_EINVAL; }
    // the level's BoxArray::maxSize decomposition, as tj_bind_level cuts it
    std::vector<int> starts[3];
    int nb[3];
    for (int d = 0; d < 3; ++d) {
        starts[d] = d < c->lv.dim ? chop_dim(c->lv.n[d], c->desc.max_grid_size) : std::vector<int>{0};
        nb[d] = (int)starts[d].size();
    }
    const long long nbox = (long long)nb[0] * nb[1] * nb[2];
    if (nfab != nbox) {
        c->err = "fabset: the FABs must be the level's maxSize boxes (" + std::to_string(nbox) + ")";
        return TJ_EINVAL;
    }
    auto box_end = [&](int d, int b) { return b + 1 < nb[d] ? starts[d][b + 1] - 1 : c->lv.n[d] - 1; };
    std::vector<int> slot(nbox, -1);
    std::vector<FabRec> rec(nbox);
    for (int q = 0; q < nfab; ++q) {
        const tj_fab& f = fabs[q];
        int bc[3];
        for (int d = 0; d < 3; ++d) {
            const auto it = std::lower_bound(starts[d].begin(), starts[d].end(), f.lo[d]);
            if (it == starts[d].end() || *it != f.lo[d] || box_end(d, (int)(it - starts[d].begin())) != f.hi[d]) {
                c->err = "fabset: FAB " + std::to_string(q) + " is not a box of the level's decomposition";
                return TJ_EINVAL;
            }
            bc[d] = (int)(it - starts[d].begin());
            if (f.glo[d] > f.lo[d]) {
                c->err = "fabset: FAB " + std::to_string(q) + " does not store its valid box (glo > lo)";
                return TJ_EINVAL;
            }
        }
        const long long e0 = f.hi[0] - f.glo[0] + 1, e1 = f.hi[1] - f.glo[1] + 1, e2 = f.hi[2] - f.glo[2] + 1;
        if (!f.data || f.jstride < e0 || f.kstride < f.jstride * e1 || f.nstride < f.kstride * e2) {
            c->err = "fabset: FAB " + std::to_string(q) + " has no data or strides that do not cover its box";
            return TJ_EINVAL;
        }
        const long long b = ((long long)bc[2] * nb[1] + bc[1]) * nb[0] + bc[0];
        if (slot[b] >= 0) {
            c->err = "fabset: two FABs for one box";
            return TJ_EINVAL;
        }
        slot[b] = q;
        // address of level cell (0,0,0) in this FAB's strides (integer arithmetic:
        // the address may lie outside the FAB; only the box's own cells are formed)
        const long long off = f.glo[0] + f.glo[1] * f.jstride + f.glo[2] * f.kstride;
        rec[b] = FabRec{reinterpret_cast<double*>(reinterpret_cast<uintptr_t>(f.data) - (uintptr_t)(off * 8)),
                        f.jstride, f.kstride, f.nstride};
    }
    tj_fabset* fs = new tj_fabset();
    fs->ctx = c;
    fs->ncomp = ncomp;
    for (int d = 0; d < 3; ++d) {
        fs->nb[d] = nb[d];
        fs->n[d] = c->lv.n[d];
    }
    for (long long b = 0; b < nbox; ++b) fs->fabs.push_back(fabs[slot[b]]);
    // the maps carry kMapPad entries beyond each end that repeat the edge box, so
    // an index a few cells outside the level still names a FAB (no kernel forms
    // one: every stencil index is mapped into the domain first)
    constexpr int kMapPad = 8;
    std::vector<int> map[3];
    const int mul[3] = {1, nb[0], nb[0] * nb[1]};
    for (int d = 0; d < 3; ++d) {
        const int n = c->lv.n[d];
        map[d].resize(n + 2 * kMapPad);
        for (int b = 0; b < nb[d]; ++b)
            for (int i = starts[d][b]; i <= box_end(d, b); ++i) map[d][kMapPad + i] = b * mul[d];
        for (int q = 0; q < kMapPad; ++q) {
            map[d][q] = map[d][kMapPad];
            map[d][kMapPad + n + q] = map[d][kMapPad + n - 1];
        }
    }
    auto fail = [&](hipError_t e) {
        c->err = std::string("fabset: ") + hipGetErrorString(e);
        TJ_SELF(tj_fabset_destroy)(fs);
        return TJ_EHIP;
    };
    hipError_t e;
    if ((e = hipSetDevice(c->device)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&fs->d_rec, sizeof(FabRec) * nbox)) != hipSuccess) return fail(e);
    if ((e = hipMemcpy(fs->d_rec, rec.data(), sizeof(FabRec) * nbox, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    for (int d = 0; d < 3; ++d) {
        if ((e = hipMalloc(&fs->d_mapbase[d], sizeof(int) * map[d].size())) != hipSuccess) return fail(e);
        if ((e = hipMemcpy(fs->d_mapbase[d], map[d].data(), sizeof(int) * map[d].size(), hipMemcpyHostToDevice)) !=
            hipSuccess)
            return fail(e);
        fs->d_map[d] = fs->d_mapbase[d] + kMapPad;
    }
    *out = fs;
    return TJ_OK;
}

int tj_level_boxes(const tj_ctx* c, int* nbox, int* lohi) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    if (!nbox) return TJ_EINVAL;
    std::vector<int> starts[3];
    for (int d = 0; d < 3; ++d)
        starts[d] = d < c->lv.dim ? chop_dim(c->lv.n[d], c->desc.max_grid_size) : std::vector<int>{0};
    *nbox = (int)(starts[0].size() * starts[1].size() * starts[2].size());
    if (!lohi) return TJ_OK;
    int q = 0;
    for (size_t bk = 0; bk < starts[2].size(); ++bk)
        for (size_t bj = 0; bj < starts[1].size(); ++bj)
            for (size_t bi = 0; bi < starts[0].size(); ++bi, ++q) {
                const size_t b[3] = {bi, bj, bk};
                for (int d = 0; d < 3; ++d) {
                    lohi[6 * q + d] = starts[d][b[d]];
                    lohi[6 * q + 3 + d] = b[d] + 1 < starts[d].size() ? starts[d][b[d] + 1] - 1 : c->lv.n[d] - 1;
                }
            }
    return TJ_OK;
}

int tj_box_op(tj_ctx* c, tj_mf* U, const int lo[3], const int hi[3], unsigned ops, double* dt_out) {
    if (!c || !c->bound) return TJ_ENOLEVEL;
    if (!U || !lo || !hi) return TJ_EINVAL;
    for (int d = 0; d < 3; ++d)
        if (lo[d] < 0 || hi[d] >= c->lv.n[d] || lo[d] > hi[d]) {
            c->err = "box_op: the box is not inside the level";
            return TJ_EINVAL;
        }
    const unsigned post = TJ_OP_PLASTIC | TJ_OP_CLEAN_V | TJ_OP_CLEAN_ALPHA | TJ_OP_NANCHECK | TJ_OP_REACTIVE;
    if (ops & TJ_OP_RADIAL) {
        c->err = "box_op: the geometric source needs each cell's radius on the level (tj_post_step)";
        return TJ_EUNSUPPORTED;
    }
    if (ops & ~(post | TJ_OP_DT | TJ_BOX_P2C | TJ_BOX_C2P | TJ_BOX_DERIVE)) {
        c->err = "box_op: unknown op bits";
        return TJ_EINVAL;
    }
    // the box as a level array whose cell (0,0,0) is lo: inside the level
    // array, or inside the FAB that owns the box
    tj_mf sub{};
    sub.ncomp = U->ncomp;
    if (U->fabs) {
        const tj_fab* own = nullptr;
        for (const tj_fab& f : U->fabs->fabs) {
            bool in = true;
            for (int d = 0; d < 3; ++d) in = in && f.lo[d] <= lo[d] && hi[d] <= f.hi[d];
            if (in) { own = &f; break; }
        }
        if (!own || U->fabs->ctx != c) {
            c->err = "box_op: with a tj_fabset the box must lie in one FAB of this context's level";
            return TJ_EINVAL;
        }
        sub.data = own->data + (lo[0] - own->glo[0]) + (lo[1] - own->glo[1]) * own->jstride +
                   (lo[2] - own->glo[2]) * own->kstride;
        sub.jstride = own->jstride;
        sub.kstride = own->kstride;
        sub.nstride = own->nstride;
        sub.ncomp = U->fabs->ncomp;
    } else {
        int r;
        if ((r = check_mf(c, U))) return r;
        sub = *U;
        sub.data = U->data + lo[0] + lo[1] * U->jstride + lo[2] * U->kstride;
    }
    HIPCHK(c, hipSetDevice(c->device));
    Clock saved;
    HIPCHK(c, hipMemcpyAsync(&saved, c->d_clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // the cell-wise kernels index cells from the level's extent only: narrow it
    const Level whole = c->lv;
    for (int d = 0; d < 3; ++d) c->lv.n[d] = hi[d] - lo[d] + 1;
    c->lv.ncells = (long long)c->lv.n[0] * c->lv.n[1] * c->lv.n[2];
    int r = TJ_OK;
    if (!r && (ops & TJ_BOX_P2C)) r = ::tj_init_from_prims(c, &sub, 0);
    if (!r && (ops & TJ_BOX_C2P)) r = ::tj_c2p(c, &sub);
    if (!r && (ops & post)) r = ::tj_post_step(c, &sub, ops & post);
    if (!r && (ops & TJ_BOX_DERIVE)) r = ::tj_derive(c, &sub);
    if (!r && (ops & TJ_OP_DT)) r = ::tj_compute_dt(c, &sub, dt_out);
    c->lv = whole;
    // the level's clock as it was; status bits raised meanwhile stay
    Clock now;
    HIPCHK(c, hipMemcpyAsync(&now, c->d_clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    saved.status = now.status;
    HIPCHK(c, hipMemcpyAsync(c->d_clk, &saved, sizeof(Clock), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return r;
}

int tj_fabset_view(const tj_fabset* fs, tj_mf* out) {
    if (!fs || !out) return TJ_EINVAL;
    *out = tj_mf{};
    out->ncomp = fs->ncomp;
    out->fabs = fs;
    return TJ_OK;
}

void tj_fabset_destroy(tj_fabset* fs) {
    if (!fs) return;
    if (fs->ctx) {
        (void)hipSetDevice(fs->ctx->device);
        (void)hipStreamSynchronize(fs->ctx->stream);
    }
    (void)hipFree(fs->d_rec);
    for (int d = 0; d < 3; ++d) (void)hipFree(fs->d_mapbase[d]);
    delete fs;
}
#endif

}  // extern "C" (namespace tjfab with TJ_FAB_VIEWS)

#ifndef TJ_FAB_VIEWS
#include "tj_batch.h"

// ---------------------------------------------------------------------------
// tj_run_static: the time loop of tj_run for callers that cannot allocate
// device memory.  Every buffer tj_bind_level allocates -- the clock words,
// the box-edge flags, the flux planes (FP64, and FP32 for the mixed form) and
// the THINC history planes -- is carved out of `scratch`, a file-scope device
// array of TJ_STATIC_SCRATCH_BYTES.  Each GPU the library runs on holds its
// own copy until the process ends.  Host code never uses the array's name as
// a device address: the address comes from hipGetSymbolAddress, once per
// device with that device current, and is kept per device.
#include <mutex>

namespace {

__device__ __align__(256) unsigned char scratch[TJ_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void* scratch_addr[kMaxDevices] = {};

// The device address of the current device's scratch (looked up on the first
// call on that device).
int scratch_for_device(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) return TJ_EINVAL;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_addr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TJ_EHIP;
        scratch_addr[dev] = p;
    }
    *out = static_cast<unsigned char*>(scratch_addr[dev]);
    return TJ_OK;
}

// Does v belong to {ratio (k sz + min(k, extra)) : klo <= k <= khi}, the
// cumulative sizes of the first k boxes counted from the high end?
__device__ __forceinline__ bool chop_sum(long long v, int ratio, int sz, int extra, int klo, int khi) {
    if (v < 0 || v % ratio) return false;
    const long long w = v / ratio;
    const long long big = (long long)extra * (sz + 1);
    long long k;
    if (w <= big) {
        if (w % (sz + 1)) return false;
        k = w / (sz + 1);
    } else {
        if ((w - big) % sz) return false;
        k = extra + (w - big) / sz;
    }
    return k >= klo && k <= khi;
}

// Box-edge flags of one coordinate (bit 0: first cell of a box, bit 1: last
// cell) by the BoxArray::maxSize rule of chop_dim, evaluated per cell: boxes
// k = 0, 1, ... from the high end hold (sz + 1) ratio cells while k < extra,
// then sz ratio, and the last box starts at 0.
__global__ void k_edge_flags(uint8_t* fl, int n, int chunk) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    uint8_t f = 0;
    if (n <= chunk) {
        f = (uint8_t)((i == 0 ? 1 : 0) | (i == n - 1 ? 2 : 0));
    } else {
        int ratio = 1, bs = chunk, nlen = n;
        while (bs % 2 == 0 && nlen % 2 == 0) { ratio *= 2; bs /= 2; nlen /= 2; }
        const int numblk = nlen / bs + (nlen % bs ? 1 : 0);
        const int sz = nlen / numblk, extra = nlen % numblk;
        if (chop_sum((long long)n - 1 - i, ratio, sz, extra, 0, numblk - 1)) f |= 2;  // a box's last cell
        if (i == 0 || chop_sum((long long)n - i, ratio, sz, extra, 1, numblk - 1)) f |= 1;  // a box's first cell
    }
    fl[i] = f;
}

// Bump allocation of the scratch: 256-byte aligned pieces; with base == null
// it only measures.
struct Carve {
    unsigned char* base;
    size_t off = 0;
    template <typename T>
    T* take(long long count) {
        off = (off + 255) & ~(size_t)255;
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += (size_t)count * sizeof(T);
        return p;
    }
};

// The buffers of tj_bind_level for a level, carved from `base` (or only
// measured, base == null): returns the bytes used.
size_t carve_level(tj_ctx* c, const tj_level_desc* d, unsigned char* base, int prec) {
    Carve cv{base};
    Level& lv = c->lv;
    c->d_clk = cv.take<Clock>(1);
    c->d_dtpart = cv.take<unsigned long long>(DT_PARTS);
    lv.dtpart = c->d_dtpart;
    c->d_ticket = reinterpret_cast<unsigned*>(cv.take<unsigned long long>(1));
    for (int a = 0; a < 3; ++a) {
        c->d_edge[a] = cv.take<uint8_t>(lv.n[a]);
        lv.edge[a] = c->d_edge[a];
    }
    c->nfc = d->SOLID ? 3 * d->nmat + 17 : 2 * d->nmat + 5 + (lv.mixm >= 0 ? 1 : 0);
    for (int a = 0; a < d->dim; ++a) {
        Faces& F = c->F[a];
        F.nx = lv.n[0] + (a == 0);
        F.ny = lv.n[1] + (a == 1);
        F.nz = lv.n[2] + (a == 2);
        F.cs = (long long)F.nx * F.ny * F.nz;
        c->d_flux[a] = cv.take<double>(F.cs * c->nfc);
        F.p = c->d_flux[a];
        c->estride[a] = F.cs;
        if (prec == TJ_PREC_MIXED) {
            FacesT<float>& G = c->Ff[a];
            G.nx = F.nx;
            G.ny = F.ny;
            G.nz = F.nz;
            G.cs = F.cs;
            c->d_fluxf[a] = cv.take<float>(F.cs * c->nfc);
            G.p = c->d_fluxf[a];
        }
    }
    if (stage2d_level(c)) c->d_ux = cv.take<double>(lv.ncells * (2 * d->nmat + 4));
    if (d->THINC) {
        const long long pl = lv.ncells * d->nmat;
        c->d_urstar = cv.take<double>(pl);
        for (int q = 0; q < 2; ++q) c->d_star[q] = cv.take<double>(pl);
        for (int a = 0; a < d->dim; ++a)
            for (int q = 0; q < 2; ++q) c->d_estar[a][q] = cv.take<double>(c->estride[a] * d->nmat);
    }
    return cv.off;
}

// The library-owned pointers of a context over the scratch are not its own:
// clear them so nothing frees them.
struct StaticCtx {
    tj_ctx c;
    ~StaticCtx() {
        for (int a = 0; a < 3; ++a) {
            c.d_edge[a] = nullptr;
            c.d_flux[a] = nullptr;
            c.d_fluxf[a] = nullptr;
            c.d_estar[a][0] = c.d_estar[a][1] = nullptr;
        }
        c.d_urstar = c.d_star[0] = c.d_star[1] = nullptr;
        c.d_ux = nullptr;
        c.d_clk = nullptr;
        c.d_dtpart = nullptr;
        c.d_ticket = nullptr;
        c.stream = nullptr;
    }
};

}  // namespace

extern "C" int tj_run_static(const tj_level_desc* desc, tj_mf* S, tj_mf* T, double final_time, int max_steps,
                             int prec, int check_every, void* stream, double* t_out, int* n_out) {
    if (!desc || !S || !T || !S->data || !T->data || S->data == T->data) return TJ_EINVAL;
    if (prec != TJ_PREC_FP64 && prec != TJ_PREC_MIXED) return TJ_EINVAL;
    StaticCtx sc;
    tj_ctx* c = &sc.c;
    const char* err = nullptr;
    int r = level_from_desc(err, desc, c->lv);
    if (r) return r;
    c->desc = *desc;
    c->nmat = desc->nmat;
    if (carve_level(c, desc, nullptr, prec) > (size_t)TJ_STATIC_SCRATCH_BYTES) return TJ_ESCRATCH;
    int ncomp_cs[NVAR];
    const int ncomp = build_layout(*desc, ncomp_cs);
    const Level& lv = c->lv;
    for (const tj_mf* m : {S, T}) {  // the checks of check_mf, without its error text
        if (m->ncomp < ncomp || m->jstride < lv.n[0] || m->kstride < m->jstride * lv.n[1] ||
            m->nstride < m->kstride * lv.n[2])
            return TJ_EINVAL;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return TJ_EHIP;
    unsigned char* base = nullptr;
    if ((r = scratch_for_device(dev, &base))) return r;
    carve_level(c, desc, base, prec);
    c->device = dev;
    c->stream = static_cast<hipStream_t>(stream);
    c->seglen = 16;  // fixed segment length: the per-pass pick keeps a host-side cache
    c->bound = true;
    for (int a = 0; a < 3; ++a) {
        hipLaunchKernelGGL(k_edge_flags, dim3(grid1(lv.n[a], 256)), dim3(256), 0, c->stream, c->d_edge[a], lv.n[a],
                           a < desc->dim ? desc->max_grid_size : lv.n[a]);
        if (hipGetLastError() != hipSuccess) return TJ_EHIP;
    }
    // the history planes start at zero on every call (they hold the last
    // call's contents otherwise), as after tj_bind_level
    if ((r = TJ_SELF(tj_reset_history)(c))) return r;
    // the 1D stage's wave tickets start at zero (a fill kernel: no memset node)
    if ((r = zero(c, reinterpret_cast<double*>(c->d_ticket), 1))) return r;
    return tj_run(c, S, T, final_time, max_steps, prec, check_every, t_out, n_out);
}

#endif  // !TJ_FAB_VIEWS
