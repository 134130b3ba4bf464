// tenjin_host.cpp -- the reference's hot-path C++ interface over the C ABI of
// the gfx950 library.  Host code only: every computation is a tj_* call.
#include "tenjin_host.h"

#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <cerrno>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

tenjin_host::Device& dev(ParameterStruct& parameters) {
    if (!parameters.gpu)
        throw std::runtime_error("tenjin_host: level not bound (call tenjin_host::bindLevel after the set-up)");
    return *parameters.gpu;
}

void tj_check(ParameterStruct& parameters, int rc, const char* what) {
    if (rc != TJ_OK) {
        const char* msg = parameters.gpu ? tj_last_error(parameters.gpu->ctx) : "";
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + (msg ? msg : ""));
    }
}

// Starts of the pieces of [0, n) cut by BoxArray::maxSize(chunk): common
// factors of two are taken out of n and chunk, the reduced length is split
// into ceil(n'/chunk') pieces whose sizes differ by at most one, and the
// pieces are laid out from the high end, the larger ones first (the rule the
// library applies in tj_bind_level).
Vector<int> chop(int n, int chunk) {
    if (n <= chunk) return {0, n};
    int ratio = 1, bs = chunk, len = n;
    while (bs % 2 == 0 && len % 2 == 0) {
        ratio *= 2;
        bs /= 2;
        len /= 2;
    }
    const int pieces = len / bs + (len % bs ? 1 : 0);
    const int sz = len / pieces, extra = len % pieces;
    Vector<int> starts;
    int hi = n - 1;
    for (int p = 0; p < pieces - 1; ++p) {
        const int size = (p < extra ? sz + 1 : sz) * ratio;
        starts.push_back(hi - size + 1);
        hi -= size;
    }
    starts.push_back(0);
    std::sort(starts.begin(), starts.end());
    starts.push_back(n);
    return starts;
}

}  // namespace

// ------------------------------------------------------------------ amrex_lite
namespace amrex_lite {

void BoxArray::maxSize(int chunk) {
    Vector<Box> out;
    for (const Box& b : boxes_) {
        Vector<int> cut[3];
        for (int d = 0; d < 3; ++d) {
            cut[d] = chop(b.length(d), chunk);
            for (int& v : cut[d]) v += b.lo[d];
        }
        for (size_t k = 0; k + 1 < cut[2].size(); ++k)
            for (size_t j = 0; j + 1 < cut[1].size(); ++j)
                for (size_t i = 0; i + 1 < cut[0].size(); ++i)
                    out.push_back(Box(IntVect(cut[0][i], cut[1][j], cut[2][k]),
                                      IntVect(cut[0][i + 1] - 1, cut[1][j + 1] - 1, cut[2][k + 1] - 1)));
    }
    boxes_ = out;
}

Box BoxArray::minimalBox() const {
    Box m = boxes_.at(0);
    for (const Box& b : boxes_)
        for (int d = 0; d < 3; ++d) {
            m.lo[d] = std::min(m.lo[d], b.lo[d]);
            m.hi[d] = std::max(m.hi[d], b.hi[d]);
        }
    return m;
}

// ------------------------------------------------------------------ FArrayBox
FArrayBox::FArrayBox(const Box& grown, int ncomp) : box_(grown), ncomp_(ncomp) {
    const size_t n = sizeof(Real) * (size_t)ncomp * (size_t)grown.numPts();
    hip_check(hipMalloc(&data_, n), "FArrayBox: hipMalloc");
    hip_check(hipMemset(data_, 0, n), "FArrayBox: hipMemset");
}
FArrayBox::~FArrayBox() {
    if (data_) (void)hipFree(data_);
}
void FArrayBox::copyToHost() {
    host.resize((size_t)ncomp_ * (size_t)nstride());
    hip_check(hipDeviceSynchronize(), "FArrayBox::copyToHost");
    hip_check(hipMemcpy(host.data(), data_, sizeof(Real) * host.size(), hipMemcpyDeviceToHost), "FArrayBox::copyToHost");
}
void FArrayBox::copyFromHost() {
    if (host.size() != (size_t)ncomp_ * (size_t)nstride()) throw std::runtime_error("FArrayBox::copyFromHost: stage first");
    hip_check(hipDeviceSynchronize(), "FArrayBox::copyFromHost");
    hip_check(hipMemcpy(data_, host.data(), sizeof(Real) * host.size(), hipMemcpyHostToDevice),
              "FArrayBox::copyFromHost");
}
Array4<Real> FArrayBox::hostArray() {
    Array4<Real> a;
    a.p = host.data();
    a.jstride = jstride();
    a.kstride = kstride();
    a.nstride = nstride();
    a.begin = box_.lo;
    a.end = IntVect(box_.hi[0] + 1, box_.hi[1] + 1, box_.hi[2] + 1);
    a.ncomp = ncomp_;
    return a;
}

// ------------------------------------------------------------------ MultiFab
void MultiFab::define(const BoxArray& ba, const DistributionMapping&, int ncomp, int ngrow, const MFInfo& info) {
    ba_ = ba;
    dom_ = ba.minimalBox();
    ncomp_ = ncomp;
    ngrow_ = ngrow;
    if (info.fabPerBox) {
        // AMReX's storage: each box grown by ngrow in the directions the level
        // extends in
        for (int b = 0; b < ba.size(); ++b) {
            Box g = ba[b];
            for (int d = 0; d < 3; ++d)
                if (dom_.length(d) > 1) {
                    g.lo[d] -= ngrow;
                    g.hi[d] += ngrow;
                }
            fabs_.push_back(new FArrayBox(g, ncomp));
        }
        return;
    }
    hip_check(hipMalloc(&data_, sizeof(Real) * (size_t)ncomp * (size_t)numCells()), "MultiFab: hipMalloc");
    hip_check(hipMemset(data_, 0, sizeof(Real) * (size_t)ncomp * (size_t)numCells()), "MultiFab: hipMemset");
}

MultiFab::~MultiFab() {
    for (FArrayBox* f : fabs_) delete f;
    if (data_) (void)hipFree(data_);
}

// the valid cells of box b between a FAB's host copy and a level-layout vector
static void fab_level_copy(const Box& vb, FArrayBox& f, Vector<Real>& lvl, const Box& dom, int ncomp, bool to_level) {
    const long long ljs = dom.length(0), lks = (long long)dom.length(0) * dom.length(1), lns = dom.numPts();
    const Box& g = f.box();
    for (int c = 0; c < ncomp; ++c)
        for (int k = vb.lo[2]; k <= vb.hi[2]; ++k)
            for (int j = vb.lo[1]; j <= vb.hi[1]; ++j) {
                Real* fr = f.host.data() + c * f.nstride() + (k - g.lo[2]) * f.kstride() + (j - g.lo[1]) * f.jstride() +
                           (vb.lo[0] - g.lo[0]);
                Real* lr = lvl.data() + c * lns + (k - dom.lo[2]) * lks + (j - dom.lo[1]) * ljs + (vb.lo[0] - dom.lo[0]);
                if (to_level) std::copy(fr, fr + vb.length(0), lr);
                else std::copy(lr, lr + vb.length(0), fr);
            }
}

void MultiFab::copyToHost(Vector<Real>& h) const {
    h.resize((size_t)ncomp_ * (size_t)numCells());
    if (fabPerBox()) {
        for (int b = 0; b < ba_.size(); ++b) {
            fabs_[b]->copyToHost();
            fab_level_copy(ba_[b], *fabs_[b], h, dom_, ncomp_, true);
        }
        return;
    }
    hip_check(hipDeviceSynchronize(), "MultiFab::copyToHost");
    hip_check(hipMemcpy(h.data(), data_, sizeof(Real) * h.size(), hipMemcpyDeviceToHost), "MultiFab::copyToHost");
}

void MultiFab::copyFromHost(const Vector<Real>& h) {
    if (h.size() != (size_t)ncomp_ * (size_t)numCells()) throw std::runtime_error("MultiFab::copyFromHost: size");
    if (fabPerBox()) {
        Vector<Real>& lvl = const_cast<Vector<Real>&>(h);
        for (int b = 0; b < ba_.size(); ++b) {
            fabs_[b]->copyToHost();  // keeps the ghost cells as they are
            fab_level_copy(ba_[b], *fabs_[b], lvl, dom_, ncomp_, false);
            fabs_[b]->copyFromHost();
        }
        return;
    }
    hip_check(hipDeviceSynchronize(), "MultiFab::copyFromHost");
    hip_check(hipMemcpy(data_, h.data(), sizeof(Real) * h.size(), hipMemcpyHostToDevice), "MultiFab::copyFromHost");
}

void MultiFab::Copy(MultiFab& dst, const MultiFab& src, int srccomp, int dstcomp, int numcomp, int) {
    if (dst.numCells() != src.numCells()) throw std::runtime_error("MultiFab::Copy: different levels");
    hip_check(hipDeviceSynchronize(), "MultiFab::Copy");
    if (dst.fabPerBox() != src.fabPerBox()) throw std::runtime_error("MultiFab::Copy: different storage forms");
    if (dst.fabPerBox()) {
        for (int b = 0; b < src.ba_.size(); ++b) {
            const FArrayBox &fs = *src.fabs_[b], &fd = *dst.fabs_[b];
            if (fs.nstride() != fd.nstride()) throw std::runtime_error("MultiFab::Copy: different ghost widths");
            hip_check(hipMemcpy(fd.dataPtr() + (size_t)dstcomp * fd.nstride(), fs.dataPtr() + (size_t)srccomp * fs.nstride(),
                                sizeof(Real) * (size_t)numcomp * (size_t)fs.nstride(), hipMemcpyDeviceToDevice),
                      "MultiFab::Copy");
        }
        return;
    }
    hip_check(hipMemcpy(dst.data_ + (size_t)dstcomp * dst.numCells(), src.data_ + (size_t)srccomp * src.numCells(),
                        sizeof(Real) * (size_t)numcomp * (size_t)src.numCells(), hipMemcpyDeviceToDevice),
              "MultiFab::Copy");
}

std::string Concatenate(const std::string& root, int num, int mindigits) {
    std::ostringstream os;
    os << root << std::setfill('0') << std::setw(mindigits) << num;
    return os.str();
}

namespace {

void make_dir(const std::string& d) {
    if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) throw std::runtime_error("cannot create directory " + d);
}

std::string ivec(const IntVect& v, int dim) {
    std::string s = "(";
    for (int d = 0; d < dim; ++d) s += (d ? "," : "") + std::to_string(v[d]);
    return s + ")";
}

// ((lo) (hi) (type)) of a cell-centred box
std::string box_str(const Box& b, int dim) {
    return "(" + ivec(b.lo, dim) + " " + ivec(b.hi, dim) + " " + ivec(IntVect(0, 0, 0), dim) + ")";
}

}  // namespace

void WriteSingleLevelPlotfile(const std::string& dir, const MultiFab& mf, const Vector<std::string>& varnames,
                              const Geometry& geom, Real time, int level_step) {
    const int dim = geom.dim();
    const int nc = mf.nComp();
    if ((int)varnames.size() < nc) throw std::runtime_error("WriteSingleLevelPlotfile: fewer names than components");
    Vector<Real> h;
    mf.copyToHost(h);
    const Box& dom = mf.domain();
    const long long ncell = mf.numCells();
    make_dir(dir);
    make_dir(dir + "/Level_0");
    std::ofstream hd(dir + "/Header");
    hd << std::setprecision(17);
    hd << "HyperCLaw-V1.1\n" << nc << "\n";
    for (int c = 0; c < nc; ++c) hd << varnames[c] << "\n";
    hd << dim << "\n" << time << "\n" << 0 << "\n";
    for (int d = 0; d < dim; ++d) hd << (d ? " " : "") << 0.0;
    hd << "\n";
    for (int d = 0; d < dim; ++d) hd << (d ? " " : "") << geom.ProbLength(d);
    hd << "\n\n" << box_str(geom.Domain(), dim) << "\n" << level_step << "\n";
    for (int d = 0; d < dim; ++d) hd << (d ? " " : "") << geom.CellSize(d);
    hd << "\n0\n0\n";
    const BoxArray& ba = mf.boxArray();
    hd << 0 << " " << ba.size() << " " << time << "\n" << level_step << "\n";
    for (int b = 0; b < ba.size(); ++b)
        for (int d = 0; d < dim; ++d)
            hd << ba[b].lo[d] * geom.CellSize(d) << " " << (ba[b].hi[d] + 1) * geom.CellSize(d) << "\n";
    hd << "Level_0/Cell\n";
    // FABs and the VisMF header
    std::ofstream fd(dir + "/Level_0/Cell_D_00000", std::ios::binary);
    std::ostringstream ch;
    ch << std::setprecision(17);
    ch << "1\n1\n" << nc << "\n0\n(" << ba.size() << " 0\n";
    for (int b = 0; b < ba.size(); ++b) ch << box_str(ba[b], dim) << "\n";
    ch << ")\n" << ba.size() << "\n";
    std::vector<std::vector<Real>> mins(ba.size()), maxs(ba.size());
    for (int b = 0; b < ba.size(); ++b) {
        const Box& bx = ba[b];
        ch << "FabOnDisk: Cell_D_00000 " << (long long)fd.tellp() << "\n";
        fd << "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))" << box_str(bx, dim) << " " << nc << "\n";
        std::vector<Real> buf;
        buf.reserve((size_t)bx.numPts());
        for (int c = 0; c < nc; ++c) {
            buf.clear();
            Real mn = std::numeric_limits<Real>::max(), mx = -std::numeric_limits<Real>::max();
            for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
                for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
                    for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
                        const long long cell = (long long)(k - dom.lo[2]) * mf.kstride() + (long long)(j - dom.lo[1]) * mf.jstride() +
                                               (i - dom.lo[0]);
                        const Real v = h[(size_t)c * (size_t)ncell + (size_t)cell];
                        buf.push_back(v);
                        mn = std::min(mn, v);
                        mx = std::max(mx, v);
                    }
            fd.write(reinterpret_cast<const char*>(buf.data()), sizeof(Real) * buf.size());
            mins[b].push_back(mn);
            maxs[b].push_back(mx);
        }
    }
    ch << "\n" << ba.size() << "," << nc << "\n";
    for (int b = 0; b < ba.size(); ++b) {
        for (Real v : mins[b]) ch << v << ",";
        ch << "\n";
    }
    ch << "\n" << ba.size() << "," << nc << "\n";
    for (int b = 0; b < ba.size(); ++b) {
        for (Real v : maxs[b]) ch << v << ",";
        ch << "\n";
    }
    std::ofstream(dir + "/Level_0/Cell_H") << ch.str();
}

}  // namespace amrex_lite

// ------------------------------------------------------------------ AccessPattern
// accesspattern.cpp:44-79: alpha, alphaRho, rho_k (per material), rho,
// rhoU(3), u(3), E, p, a, uStar, sigma(3x3) [mixtures: none on this path];
// solids add V(3x3), VStar(3x3), devH(3x3), HenckyJ2; plastic solids add
// epsilon and alphaRhoEpsilon (per material).  Names are nameBase + material
// + "_" + row + col (accesspattern.cpp:17).
AccessPattern::AccessPattern(ParameterStruct& parameters) {
    int n = 0;
    auto add = [&](Variable v, const std::string& name, int mats, int rows = 1, int cols = 1) {
        data[v] = n;
        for (int m = 0; m < mats; ++m)
            for (int r = 0; r < rows; ++r)
                for (int c = 0; c < cols; ++c)
                    variableNames.push_back(name + std::to_string(m) + "_" + std::to_string(r) + std::to_string(c));
        n += mats * rows * cols;
    };
    const int M = parameters.numberOfMaterials;
    add(ALPHA, "alpha", M);
    add(ALPHARHO, "alphaRho", M);
    add(RHO_K, "rho_k", M);
    add(RHO, "rho", 1);
    add(RHOU, "rhoU", 1, 3);
    add(VELOCITY, "u", 1, 3);
    add(TOTAL_E, "E", 1);
    add(P, "p", 1);
    add(SOUNDSPEED, "a", 1);
    add(USTAR, "uStar", 1);
    add(SIGMA, "sigma", 1, 3, 3);
    if (parameters.numberOfMixtures > 0) {
        add(RHO_MIX, "rhomix", parameters.numberOfMixtures, 2);
        add(LAMBDA, "lambda", parameters.numberOfMixtures);
        add(ALPHARHOLAMBDA, "alpharholambda", parameters.numberOfMixtures);
    }
    if (parameters.SOLID) {
        add(V_TENSOR, "V", 1, 3, 3);
        add(VSTAR, "VStar", 1, 3, 3);
        add(DEVH, "devH", 1, 3, 3);
        add(HJ2, "HenckyJ2", 1);
        if (parameters.PLASTIC) {
            add(EPSILON, "epsilon", M);
            add(ALPHARHOEPSILON, "alphaRhoEpsilon", M);
        }
    }
    ncomp = n;
    parameters.Ncomp = n;
}

// ------------------------------------------------------------------ CellArray
CellArray::CellArray(BoxArray& ba, DistributionMapping& dm, const int Ncomp, const int Nghost,
                     AccessPattern& _accessPattern, ParameterStruct& parameters_)
    : accessPattern(_accessPattern), numberOfMaterials(parameters_.numberOfMaterials), parameters(parameters_) {
    data.define(ba, dm, Ncomp, Nghost, MFInfo().SetFabPerBox(parameters_.fabPerBox != 0));
}

CellArray::~CellArray() {
    if (fabset_) tj_fabset_destroy(fabset_);
}

tj_mf CellArray::view() {
    tj_mf v{};
    if (data.fabPerBox()) {
        tj_ctx* ctx = dev(parameters).ctx;
        if (fabset_ && fabsetCtx_ != ctx) {  // the level was bound again
            tj_fabset_destroy(fabset_);
            fabset_ = nullptr;
        }
        if (!fabset_) {
            const BoxArray& ba = data.boxArray();
            std::vector<tj_fab> fabs(ba.size());
            for (int b = 0; b < ba.size(); ++b) {
                const FArrayBox& f = data.fab(b);
                tj_fab& r = fabs[b];
                r.data = f.dataPtr();
                for (int d = 0; d < 3; ++d) {
                    r.lo[d] = ba[b].lo[d];
                    r.hi[d] = ba[b].hi[d];
                    r.glo[d] = f.box().lo[d];
                }
                r.jstride = f.jstride();
                r.kstride = f.kstride();
                r.nstride = f.nstride();
            }
            tj_check(parameters, tj_fabset_create(ctx, ba.size(), fabs.data(), data.nComp(), &fabset_), "tj_fabset_create");
            fabsetCtx_ = ctx;
        }
        tj_fabset_view(fabset_, &v);
        return v;
    }
    v.data = data.devicePtr();
    v.jstride = data.jstride();
    v.kstride = data.kstride();
    v.nstride = data.nstride();
    v.ncomp = data.nComp();
    return v;
}

void CellArray::stage() {
    if (data.fabPerBox())
        for (int b = 0; b < data.boxArray().size(); ++b) data.fab(b).copyToHost();
    data.copyToHost(host);
}
void CellArray::commit() {
    if (data.fabPerBox()) {
        // the FABs' own staged copies carry the BoxAccessCellArray writes
        for (int b = 0; b < data.boxArray().size(); ++b) data.fab(b).copyFromHost();
        data.copyToHost(host);
        return;
    }
    data.copyFromHost(host);
}

// cellarray.cpp:5-18
void CellArray::conservativeToPrimitive() {
    tj_mf v = view();
    tj_check(parameters, tj_c2p(dev(parameters).ctx, &v), "conservativeToPrimitive");
}

// cellarray.cpp:20-33
void CellArray::primitiveToConservative() {
    tj_mf v = view();
    tj_check(parameters, tj_init_from_prims(dev(parameters).ctx, &v, 0), "primitiveToConservative");
}

// cellarray.cpp:35-48 (the CFL candidates it also feeds are re-armed by the
// next getTimeStep)
void CellArray::getSoundSpeed() {
    tj_mf v = view();
    tj_check(parameters, tj_post_step(dev(parameters).ctx, &v, TJ_OP_DT), "getSoundSpeed");
}

// cellarray.cpp:50-63
void CellArray::cleanUpV() {
    tj_mf v = view();
    tj_check(parameters, tj_post_step(dev(parameters).ctx, &v, TJ_OP_CLEAN_V), "cleanUpV");
}

// cellarray.cpp:65-78
void CellArray::cleanUpAlpha() {
    tj_mf v = view();
    tj_check(parameters, tj_post_step(dev(parameters).ctx, &v, TJ_OP_CLEAN_ALPHA), "cleanUpAlpha");
}

// cellarray.cpp:80-85: copy of every component
void CellArray::operator=(CellArray& U) {
    tj_mf a = U.view(), b = view();
    tj_check(parameters, tj_lincomb(dev(parameters).ctx, &a, 1.0, nullptr, 0.0, &b), "CellArray::operator=");
}

// cellarray.cpp:87-92: scales in place and returns *this
CellArray& CellArray::operator*(Real d) {
    tj_mf a = view();
    tj_check(parameters, tj_lincomb(dev(parameters).ctx, &a, d, nullptr, 0.0, &a), "CellArray::operator*");
    return *this;
}

// cellarray.cpp:94-99: adds U in place and returns *this
CellArray& CellArray::operator+(CellArray& U) {
    tj_mf a = view(), b = U.view();
    tj_check(parameters, tj_lincomb(dev(parameters).ctx, &a, 1.0, &b, 1.0, &a), "CellArray::operator+");
    return *this;
}

// cellarray.cpp:101-117
bool CellArray::contains_nan() {
    tj_mf v = view();
    tj_check(parameters, tj_post_step(dev(parameters).ctx, &v, TJ_OP_NANCHECK), "contains_nan");
    return tenjin_host::checkStatus(parameters);
}

// ------------------------------------------------------------------ BoxAccessCellArray
// the FAB whose valid box contains bx (fabPerBox)
static FArrayBox* owning_fab(CellArray& U, const Box& bx) {
    const BoxArray& ba = U.data.boxArray();
    for (int b = 0; b < ba.size(); ++b) {
        bool in = true;
        for (int d = 0; d < 3; ++d) in = in && ba[b].lo[d] <= bx.lo[d] && bx.hi[d] <= ba[b].hi[d];
        if (in) return &U.data.fab(b);
    }
    throw std::runtime_error("BoxAccessCellArray: the box lies in no single FAB of the CellArray");
}

BoxAccessCellArray::BoxAccessCellArray(MFIter& mfi, const Box& bx, CellArray& U_)
    : box(bx), U(U_), accessPattern(U_.accessPattern), numberOfMaterials(U_.numberOfMaterials) {
    if (U.data.fabPerBox()) fabp = &U.data.fab(mfi.index());
    if (fabp ? fabp->host.empty() : U.host.empty())
        throw std::runtime_error("BoxAccessCellArray: stage() the CellArray first");
}

BoxAccessCellArray::BoxAccessCellArray(const Box& bx, FArrayBox& fb, CellArray& U_)
    : box(bx), U(U_), fabp(&fb), accessPattern(U_.accessPattern), numberOfMaterials(U_.numberOfMaterials) {
    if (fb.host.empty()) throw std::runtime_error("BoxAccessCellArray: stage the FArrayBox first (copyToHost)");
}

BoxAccessCellArray::BoxAccessCellArray(const Box& bx, CellArray& U_)
    : box(bx), U(U_), accessPattern(U_.accessPattern), numberOfMaterials(U_.numberOfMaterials) {
    if (U.data.fabPerBox()) fabp = owning_fab(U, bx);
    if (fabp ? fabp->host.empty() : U.host.empty())
        throw std::runtime_error("BoxAccessCellArray: stage() the CellArray first");
}

// component of (var, mat, row, col) as in boxaccesscellarray.cpp:19-63:
// per-material variables add mat, vectors add row, 3x3 tensors add 3 row + col
Real& BoxAccessCellArray::operator()(int i, int j, int k, Variable var, int mat, int row, int col) {
    int comp = accessPattern[var];
    switch (var) {
        case ALPHA: case ALPHARHO: case RHO_K: case EPSILON: case ALPHARHOEPSILON: comp += mat; break;
        // + mixtureIndex, which is 0 for every mixture (initialise.cpp:187)
        case RHO_MIX: comp += row; break;
        case LAMBDA: case ALPHARHOLAMBDA: break;
        case RHOU: case VELOCITY: comp += row; break;
        case SIGMA: case V_TENSOR: case VSTAR: case DEVH: comp += row * 3 + col; break;
        case RHO: case TOTAL_E: case P: case SOUNDSPEED: case USTAR: case HJ2: break;
        default: throw std::runtime_error("Incorrect Access variable " + std::to_string((int)var));
    }
    if (fabp) {
        const Box& g = fabp->box();
        const long long cell = (long long)(k - g.lo[2]) * fabp->kstride() + (long long)(j - g.lo[1]) * fabp->jstride() +
                               (i - g.lo[0]);
        return fabp->host[(size_t)comp * (size_t)fabp->nstride() + (size_t)cell];
    }
    const Box& d = U.data.domain();
    const long long cell = ((long long)(k - d.lo[2]) * U.data.kstride()) + (long long)(j - d.lo[1]) * U.data.jstride() +
                           (i - d.lo[0]);
    return U.host[(size_t)comp * (size_t)U.data.nstride() + (size_t)cell];
}

Real& BoxAccessCellArray::operator()(int i, int j, int k, int var, int mat, int row, int col) {
    return (*this)(i, j, k, (Variable)var, mat, row, col);
}
Real& BoxAccessCellArray::operator()(int i, int j, int k, MaterialSpecifier& m) {
    return (*this)(i, j, k, m.var, m.mat, m.row, m.col);
}

// boxaccesscellarray.cpp:382-430: the cell one step down or up along d
Real& BoxAccessCellArray::left(Direction_enum d, int i, int j, int k, MaterialSpecifier& m) {
    switch (d) {
        case x: return (*this)(i - 1, j, k, m);
        case y: return (*this)(i, j - 1, k, m);
        case z: return (*this)(i, j, k - 1, m);
    }
    std::printf("Bad Direction in Left function\n");
    std::exit(1);
}
Real& BoxAccessCellArray::right(Direction_enum d, int i, int j, int k, MaterialSpecifier& m) {
    switch (d) {
        case x: return (*this)(i + 1, j, k, m);
        case y: return (*this)(i, j + 1, k, m);
        case z: return (*this)(i, j, k + 1, m);
    }
    std::printf("Bad Direction in Right function\n");
    std::exit(1);
}
Real& BoxAccessCellArray::left(Direction_enum d, int i, int j, int k, Variable var, int mat, int row, int col) {
    MaterialSpecifier t(var, mat, row, col);
    return left(d, i, j, k, t);
}
Real& BoxAccessCellArray::right(Direction_enum d, int i, int j, int k, Variable var, int mat, int row, int col) {
    MaterialSpecifier t(var, mat, row, col);
    return right(d, i, j, k, t);
}
Real& BoxAccessCellArray::neighbour(int di, int dj, int dk, int i, int j, int k, MaterialSpecifier& m) {
    return (*this)(i + di, j + dj, k + dk, m);
}
Real& BoxAccessCellArray::neighbour(int di, int dj, int dk, int i, int j, int k, Variable var, int mat, int row,
                                    int col) {
    MaterialSpecifier t(var, mat, row, col);
    return neighbour(di, dj, dk, i, j, k, t);
}
// boxaccesscellarray.cpp:432-444
void BoxAccessCellArray::amrexToArray(int i, int j, int k, Variable var, int m, double* copy, int nx, int ny) {
    for (int row = 0; row < nx; ++row)
        for (int col = 0; col < ny; ++col) copy[row * nx + col] = (*this)(i, j, k, var, m, row, col);
}

// The region's staged cells go to the device, the op runs on the region there,
// and the region's cells are staged again.  Staged copies hold the whole FAB
// (fabPerBox) or the whole level; only the region's rows move.
Real BoxAccessCellArray::deviceOp(const Box& region, unsigned ops, bool* nan) {
    ParameterStruct& prm = U.parameters;
    const int nc = U.data.nComp();
    Real* dptr;
    Real* hptr;
    Box st;  // the box the staged copy and the device buffer cover
    long long js, ks, ns;
    if (fabp) {
        st = fabp->box();
        dptr = fabp->dataPtr();
        hptr = fabp->host.data();
        js = fabp->jstride(), ks = fabp->kstride(), ns = fabp->nstride();
    } else {
        st = U.data.domain();
        dptr = U.data.devicePtr();
        hptr = U.host.data();
        js = U.data.jstride(), ks = U.data.kstride(), ns = U.data.nstride();
    }
    auto rows = [&](bool to_device) {
        for (int c = 0; c < nc; ++c)
            for (int k = region.lo[2]; k <= region.hi[2]; ++k) {
                const size_t off = (size_t)c * ns + (size_t)(k - st.lo[2]) * ks + (size_t)(region.lo[1] - st.lo[1]) * js +
                                   (size_t)(region.lo[0] - st.lo[0]);
                const size_t w = sizeof(Real) * region.length(0), pitch = sizeof(Real) * js;
                hip_check(to_device ? hipMemcpy2D(dptr + off, pitch, hptr + off, pitch, w, region.length(1),
                                                  hipMemcpyHostToDevice)
                                    : hipMemcpy2D(hptr + off, pitch, dptr + off, pitch, w, region.length(1),
                                                  hipMemcpyDeviceToHost),
                          "BoxAccessCellArray: region copy");
            }
    };
    hip_check(hipDeviceSynchronize(), "BoxAccessCellArray");
    rows(true);
    tj_mf v = U.view();
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = region.lo[d];
        hi[d] = region.hi[d];
    }
    double dt = 0.0;
    tj_check(prm, tj_box_op(dev(prm).ctx, &v, lo, hi, ops, &dt), "tj_box_op");
    hip_check(hipDeviceSynchronize(), "BoxAccessCellArray");
    rows(false);
    const bool n = tenjin_host::checkStatus(prm);
    if (nan) *nan = n;
    return dt;
}

void BoxAccessCellArray::conservativeToPrimitive() { deviceOp(box, TJ_BOX_C2P); }
void BoxAccessCellArray::primitiveToConservative() { deviceOp(box, TJ_BOX_P2C); }
void BoxAccessCellArray::getSoundSpeed() { deviceOp(box, TJ_OP_DT); }
void BoxAccessCellArray::cleanUpV() { deviceOp(box, TJ_OP_CLEAN_V); }
void BoxAccessCellArray::cleanUpAlpha() { deviceOp(box, TJ_OP_CLEAN_ALPHA); }
bool BoxAccessCellArray::contains_nan() {
    bool nan = false;
    deviceOp(box, TJ_OP_NANCHECK, &nan);
    return nan;
}
void BoxAccessCellArray::stressTensor(int i, int j, int k) { deviceOp(Box(IntVect(i, j, k), IntVect(i, j, k)), TJ_BOX_DERIVE); }
void BoxAccessCellArray::getHenckyJ2(int i, int j, int k) { deviceOp(Box(IntVect(i, j, k), IntVect(i, j, k)), TJ_BOX_DERIVE); }
void BoxAccessCellArray::getDeviatoricHenckyStrain(int i, int j, int k) {
    deviceOp(Box(IntVect(i, j, k), IntVect(i, j, k)), TJ_BOX_DERIVE);
}

// ------------------------------------------------------------------ TimeStep, PlasticEOS, advance
// timestep.cpp:7-34 over the staged cells of U.box
void TimeStep::boxTimeStepFunction(BoxAccessCellArray& U, Array4<Real> const& prop_arr_time, ParameterStruct& parameters) {
    const int dim = parameters.gpu ? dev(parameters).desc.dim : 3;
    for (int k = U.box.lo[2]; k <= U.box.hi[2]; ++k)
        for (int j = U.box.lo[1]; j <= U.box.hi[1]; ++j)
            for (int i = U.box.lo[0]; i <= U.box.hi[0]; ++i)
                for (int n = 0; n < dim; ++n)
                    prop_arr_time(i, j, k, n) = parameters.dx[n] / (U(i, j, k, SOUNDSPEED) + std::fabs(U(i, j, k, VELOCITY, 0, n)));
}

// plastic.cpp:83-225 on U.box
void PlasticEOS::boxPlasticUpdate(BoxAccessCellArray& U, ParameterStruct&) { U.deviceOp(U.box, TJ_OP_PLASTIC); }

// timestep.cpp:36-70: CFL * min over the valid cells of dx_d / (a + |u_d|)
Real TimeStep::getTimeStep(CellArray& U, ParameterStruct& parameters) {
    tj_mf v = U.view();
    double dt = 0.0;
    tj_check(parameters, tj_compute_dt(dev(parameters).ctx, &v, &dt), "getTimeStep");
    tenjin_host::checkStatus(parameters);
    return dt;
}

// plastic.cpp:227-244 (yield stresses come from the materials' EOS objects)
void PlasticEOS::plasticUpdate(CellArray& U, ParameterStruct& parameters) {
    tenjin_host::Device& d = dev(parameters);
    for (size_t m = 0; m < yieldStress.size() && m < (size_t)d.desc.nmat; ++m)
        if (yieldStress[m] != d.desc.mat[m].yield_stress)
            throw std::runtime_error("plasticUpdate: PlasticEOS::yieldStress differs from the bound materials");
    tj_mf v = U.view();
    tj_check(parameters, tj_post_step(d.ctx, &v, TJ_OP_PLASTIC), "plasticUpdate");
}

// reactive.cpp:60-103 with parameters.dt
void reactiveUpdate(CellArray& U, CellArray& U1, CellArray& U2, ParameterStruct& parameters) {
    (void)U2;
    tj_ctx* ctx = dev(parameters).ctx;
    if (parameters.numberOfMixtures <= 0) return;
    tj_mf u = U.view(), u1 = U1.view();
    tj_check(parameters, tj_lincomb(ctx, &u1, 1.0, nullptr, 0.0, &u), "reactiveUpdate: U = U1");
    tj_check(parameters, tj_set_dt(ctx, parameters.dt, 0), "reactiveUpdate: dt");
    tj_check(parameters, tj_post_step(ctx, &u1, TJ_OP_REACTIVE), "reactiveUpdate");
}

// radial.cpp:37-54 with parameters.dt
void geometricSourceTerm(CellArray& U, ParameterStruct& parameters) {
    tj_ctx* ctx = dev(parameters).ctx;
    tj_check(parameters, tj_set_dt(ctx, parameters.dt, 0), "geometricSourceTerm: dt");
    tj_mf v = U.view();
    tj_check(parameters, tj_post_step(ctx, &v, TJ_OP_RADIAL), "geometricSourceTerm");
}

// hllc.cpp:699-707: U1 = 1/2 U + 1/2 HLLCadvance(HLLCadvance(U)) with
// parameters.dt; as in the reference, U holds the averaged state afterwards.
void advance(CellArray& U, CellArray& U1, CellArray& U2, CellArray& MUSCLgrad, CellArray& UL, CellArray& UR,
             CellArray& ULStar, CellArray& URStar, CellArray& UStarStar, Array<MultiFab, AMREX_SPACEDIM>& flux_arr,
             Geometry const& geom, ParameterStruct& parameters, Vector<BCRec>& bc, THINCArray& THINC) {
    (void)U2, (void)MUSCLgrad, (void)UL, (void)UR, (void)ULStar, (void)URStar, (void)UStarStar, (void)flux_arr;
    (void)THINC;
    if (!parameters.gpu) tenjin_host::bindLevel(parameters, geom, bc);
    tj_ctx* ctx = parameters.gpu->ctx;
    tj_check(parameters, tj_set_dt(ctx, parameters.dt, 0), "advance: dt");
    tj_mf u = U.view(), u1 = U1.view();
    tj_check(parameters, tj_rk_stage(ctx, 1, &u, &u, &u1, TJ_PREC_FP64), "advance: stage 1");
    tj_check(parameters, tj_rk_stage(ctx, 2, &u1, &u, &u1, TJ_PREC_FP64), "advance: stage 2");
    tj_check(parameters, tj_lincomb(ctx, &u1, 1.0, nullptr, 0.0, &u), "advance: U = U1");
}

// ------------------------------------------------------------------ binding
namespace tenjin_host {

Device::~Device() {
    if (ctx) tj_destroy(ctx);
}

void bindLevel(ParameterStruct& parameters, const Geometry& geom, const Vector<BCRec>& bc, int device) {
    auto d = std::make_shared<Device>();
    tj_level_desc& L = d->desc;
    L.dim = geom.dim();
    for (int a = 0; a < 3; ++a) {
        L.n[a] = a < L.dim ? geom.Domain().length(a) : 1;
        L.L[a] = a < L.dim ? geom.ProbLength(a) : 1.0;
        const int lo = bc.empty() ? BCType::foextrap : bc[0].lo(a);
        const int hi = bc.empty() ? BCType::foextrap : bc[0].hi(a);
        L.bc_lo[a] = lo == BCType::reflect_even || lo == BCType::reflect_odd;
        L.bc_hi[a] = hi == BCType::reflect_even || hi == BCType::reflect_odd;
    }
    L.max_grid_size = parameters.max_grid_size;
    L.SOLID = parameters.SOLID;
    L.THINC = parameters.THINC;
    L.PLASTIC = parameters.PLASTIC;
    L.MUSCL = parameters.MUSCL;
    L.RADIAL = parameters.RADIAL;
    L.REACTIVE = parameters.REACTIVE;
    L.THINCbeta = parameters.THINCbeta;
    L.CFL = parameters.CFL;
    L.nmat = parameters.numberOfMaterials;
    if (L.nmat > TJ_MAX_MATERIALS) throw std::runtime_error("bindLevel: too many materials");
    for (int m = 0; m < L.nmat; ++m) {
        const MaterialDescriptor& md = parameters.materialInfo.at(m);
        if (!md.EOS) throw std::runtime_error("bindLevel: material without EOS");
        tj_material& t = L.mat[m];
        if (md.mixture) {
            const MixtureEOS* mx = dynamic_cast<const MixtureEOS*>(md.EOS);
            if (!mx) throw std::runtime_error("bindLevel: a mixture material needs a MixtureEOS");
            t.eos = 0;
            t.mixture = 1;
            t.gamma = mx->first.adiabaticIndex;
            t.pref = mx->first.pref;
            t.eref = mx->first.eref;
            t.CV = mx->first.CV;
            t.gamma_mix = mx->second.adiabaticIndex;
            t.pref_mix = mx->second.pref;
            t.eref_mix = mx->second.eref;
            t.CV_mix = mx->second.CV;
            t.solid = md.phase == solid;
            continue;
        }
        t.eos = md.EOS->kind();
        t.solid = md.phase == solid;
        t.plastic = md.plastic;
        t.gamma = md.EOS->adiabaticIndex;
        t.pref = md.EOS->pref;
        t.eref = md.EOS->eref;
        t.CV = md.EOS->CV;
        t.G0 = md.EOS->shearModulus;
        t.yield_stress = md.EOS->yieldStress;
        if (const UsUpSolidEOS* s = dynamic_cast<const UsUpSolidEOS*>(md.EOS)) {
            t.rho0 = s->rho0;
            t.c0 = s->c0;
            t.s = s->s;
        }
        if (const RomenskiiSolidEOS* r = dynamic_cast<const RomenskiiSolidEOS*>(md.EOS)) {
            t.rho0 = r->rho0;
            t.K0 = r->K0;
            t.eos_alpha = r->EOSalpha;
            t.eos_beta = r->EOSbeta;
        }
        if (const WilkinsSolidEOS* w = dynamic_cast<const WilkinsSolidEOS*>(md.EOS)) {
            t.rho0 = w->rho0;
            t.e1 = w->e1;
            t.e2 = w->e2;
            t.e3 = w->e3;
            t.e4 = w->e4;
            t.e5 = w->e5;
        }
    }
    if (tj_create(&d->ctx, device) != TJ_OK) throw std::runtime_error("bindLevel: tj_create failed");
    parameters.gpu = d;
    tj_check(parameters, tj_bind_level(d->ctx, &L), "tj_bind_level");
    for (int a = 0; a < L.dim; ++a) parameters.dx[a] = geom.CellSize(a);
}

void initialiseFromPrimitives(CellArray& U, bool shiftPressure) {
    tj_mf v = U.view();
    tj_check(U.parameters, tj_init_from_prims(dev(U.parameters).ctx, &v, shiftPressure ? 1 : 0),
             "initialiseFromPrimitives");
}

bool checkStatus(ParameterStruct& parameters) {
    uint32_t st = 0;
    tj_check(parameters, tj_status(dev(parameters).ctx, &st, 1), "tj_status");
    if (st & TJ_ST_PLASTIC_DIV) {
        std::printf("Divide by zero error\n");  // plastic.cpp:188-193
        std::exit(1);
    }
    if (st & TJ_ST_MIX_BRACKET) {
        std::printf("Error in root Bisection \n");  // equationofstate.cpp:336-343
        std::exit(1);
    }
    if (st & TJ_ST_MIX_ITER) {
        // the reference's bisection loop (equationofstate.cpp:347-373) has no
        // bound and never returns here; the device form stops after 2000 passes
        std::printf("Root finding is taking a while\n");
        std::exit(1);
    }
    if (st & TJ_ST_WAVESPEED) {
        std::printf("Nan in Wavespeeds \n");  // hllc.cpp:169-199, 367-372
        std::exit(1);
    }
    if (st & TJ_ST_ALPHA_NANS) {
        std::printf("Error in scaling volume fractions, too many Nans\n");  // boxaccesscellarray.cpp:559-561
        std::exit(1);
    }
    if (st & TJ_ST_ALPHA_SUM) {
        std::printf("Error in scaling volume fractions\n");  // boxaccesscellarray.cpp:585-587
        std::exit(1);
    }
    return (st & TJ_ST_NAN) != 0;
}

}  // namespace tenjin_host
