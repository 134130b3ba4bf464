// Library context shared by the translation units of the C ABI (the FP64 /
// mixed forms in tj_lib.hip, the tangent-linear form in tj_tangent.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/tenjin.h"
#include "tj_kernels.h"
#include "tj_solid_kernels.h"
#include "tj_stage2d.h"

using namespace tj;

struct tj_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool bound = false;
    tj_level_desc desc{};
    Level lv{};
    int nmat = 0;
    uint8_t* d_edge[3] = {nullptr, nullptr, nullptr};
    double* d_flux[3] = {nullptr, nullptr, nullptr};
    Faces F[3]{};
    // FP32 flux planes of the mixed-precision form (allocated on first use)
    float* d_fluxf[3] = {nullptr, nullptr, nullptr};
    FacesT<float> Ff[3]{};
    int nfc = 0;  // flux components per face
    double* d_urstar = nullptr;
    double* d_star[2] = {nullptr, nullptr};
    double* d_estar[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    long long estride[3] = {0, 0, 0};
    Clock* d_clk = nullptr;
    unsigned long long* d_dtpart = nullptr;  // DT_PARTS partial slots of the CFL minimum
    unsigned* d_ticket = nullptr;            // wave tickets of the one-launch 1D stage (k_xstage1d), 0 between launches
    unsigned* d_status = nullptr;
    long long pass = 0;
    long long dpass[3] = {0, 0, 0};
    // rows per marching segment: 0 = chosen per pass from the launch shape
    // (pick_seglen, cached per kernel and direction), > 0 = fixed by the caller
    int seglen = 0;
    int ncu = 0;  // compute units of the device
    std::map<std::pair<const void*, int>, int> seg_cache;
    bool fuse = false;
    // two-dimensional fluid levels without a mixture: each RK stage as the x
    // and y stage kernels (tj_stage2d.h) over the partially updated state
    // d_ux, with no flux planes (on by default; off: the flux-plane passes)
    bool stage2d = true;
    double* d_ux = nullptr;
    // a captured run of loop bodies (tj_graph_capture) and the THINC pass
    // counters it advances each launch
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    long long g_pass0 = 0, g_dpass0[3] = {0, 0, 0};
    long long g_npass = 0, g_ndpass[3] = {0, 0, 0};
    // per-kernel event timing
    bool prof = false;
    std::string prof_only;  // when set, only launches of this name are timed
    int prof_every = 1;     // of the launches that qualify, every prof_every-th is timed
    long long prof_seen = 0;
    struct Rec { const char* name; hipEvent_t a, b; };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<long long, double>> acc;
};

#define HIPCHK(ctx, call)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                 \
            return TJ_EHIP;                                                                 \
        }                                                                                   \
    } while (0)

// A MultiFab of one FAB per box (tj_fabset_create): the FAB records in
// box-grid order and the per-dimension box maps, in device memory.
struct tj_fabset {
    tj_ctx* ctx = nullptr;
    int ncomp = 0;
    int nb[3] = {1, 1, 1};  // boxes per dimension
    int n[3] = {1, 1, 1};   // the level's cells per dimension when the set was made
    tj::FabRec* d_rec = nullptr;
    int* d_map[3] = {nullptr, nullptr, nullptr};      // cell index -> box offset (d_mapbase + pad)
    int* d_mapbase[3] = {nullptr, nullptr, nullptr};  // the allocations
    std::vector<tj_fab> fabs;  // the caller's records, in box-grid order
};

inline unsigned grid1(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

namespace tj {
inline namespace TJ_VIEW_NS {
#ifdef TJ_FAB_VIEWS
inline MF view(const tj_mf* m) {
    const tj_fabset* f = m->fabs;
    return MF{f->d_rec, f->d_map[0], f->d_map[1], f->d_map[2]};
}
inline int check_mf(tj_ctx* c, const tj_mf* m) {
    if (!m || !m->fabs) { c->err = "null tj_mf"; return TJ_EINVAL; }
    const tj_fabset* f = m->fabs;
    if (f->ctx != c) { c->err = "tj_fabset made for another context"; return TJ_EINVAL; }
    if (f->n[0] != c->lv.n[0] || f->n[1] != c->lv.n[1] || f->n[2] != c->lv.n[2]) {
        c->err = "tj_fabset made for another level";
        return TJ_EINVAL;
    }
    if (m->ncomp < tj_ncomp(c) || f->ncomp < tj_ncomp(c)) {
        c->err = "tj_mf has fewer components than the bound layout";
        return TJ_EINVAL;
    }
    return TJ_OK;
}
#else
inline MF view(const tj_mf* m) { return MF{m->data, m->jstride, m->kstride, m->nstride}; }
inline int check_mf(tj_ctx* c, const tj_mf* m) {
    if (!m || !m->data) { c->err = "null tj_mf"; return TJ_EINVAL; }
    if (m->fabs) { c->err = "tj_mf names a tj_fabset where a level array is expected"; return TJ_EINVAL; }
    if (m->ncomp < tj_ncomp(c)) { c->err = "tj_mf has fewer components than the bound layout"; return TJ_EINVAL; }
    if (m->jstride < c->lv.n[0] || m->kstride < m->jstride * c->lv.n[1] || m->nstride < m->kstride * c->lv.n[2]) {
        c->err = "tj_mf strides do not cover the level";
        return TJ_EINVAL;
    }
    return TJ_OK;
}
#endif
}  // inline namespace TJ_VIEW_NS
}  // namespace tj
