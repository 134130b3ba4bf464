// amrex_lite.h -- the few AMReX-shaped host types that the reference's hot-path
// signatures name (Box, BoxArray, DistributionMapping, Geometry, BCRec,
// MultiFab, MFIter, Array, Vector), for one level on one GPU.
//
// These are not AMReX.  They keep AMReX's names and meanings where the
// reference's code relies on them:
//   * BoxArray::maxSize cuts the domain by AMReX's rule (pieces of at most
//     max_grid_size cells, laid out as tj_bind_level lays them out);
//   * BCRec holds AMReX BCType codes (reflect_odd -1, reflect_even 1,
//     foextrap 2);
//   * a MultiFab's data are component-major with i fastest (FArrayBox order).
// They differ where a GPU-resident level wants something else: by default a
// MultiFab is one device array over the whole level's valid cells (its boxes
// are views into it) and stores no ghost cells -- ghost values are mapped on
// the fly by the kernels -- so nGrow() is recorded but not allocated.  With
// MFInfo().SetFabPerBox() it holds AMReX's own storage instead: one FArrayBox
// per box, grown by nGrow() cells in the active directions, and the library
// runs on those FABs in place (tj_fabset).
#pragma once

#include <array>
#include <cstddef>
#include <string>
#include <vector>

#ifndef AMREX_SPACEDIM
#define AMREX_SPACEDIM 3
#endif

namespace amrex_lite {

using Real = double;
template <class T>
using Vector = std::vector<T>;
template <class T, std::size_t N>
using Array = std::array<T, N>;

struct IntVect {
    int v[3] = {0, 0, 0};
    IntVect() = default;
    IntVect(int a, int b, int c) : v{a, b, c} {}
    int operator[](int d) const { return v[d]; }
    int& operator[](int d) { return v[d]; }
};

// Cell-centred index box [lo, hi] (inclusive, AMReX convention).
struct Box {
    IntVect lo, hi;
    Box() = default;
    Box(const IntVect& l, const IntVect& h) : lo(l), hi(h) {}
    const IntVect& smallEnd() const { return lo; }
    const IntVect& bigEnd() const { return hi; }
    int smallEnd(int d) const { return lo[d]; }
    int bigEnd(int d) const { return hi[d]; }
    int length(int d) const { return hi[d] - lo[d] + 1; }
    long long numPts() const { return (long long)length(0) * length(1) * length(2); }
};

class BoxArray {
public:
    BoxArray() = default;
    explicit BoxArray(const Box& domain) : boxes_{domain} {}
    // Chops every box into pieces of at most `chunk` cells per direction.
    void maxSize(int chunk);
    int size() const { return (int)boxes_.size(); }
    const Box& operator[](int b) const { return boxes_[b]; }
    // the union of the boxes (the level's domain)
    Box minimalBox() const;

private:
    Vector<Box> boxes_;
};

// Single process, single GPU: every box is local.
class DistributionMapping {
public:
    DistributionMapping() = default;
    explicit DistributionMapping(const BoxArray&) {}
};

struct RealBox {
    Real lo[3] = {0, 0, 0};
    Real hi[3] = {1, 1, 1};
    RealBox() = default;
    RealBox(const Array<Real, 3>& l, const Array<Real, 3>& h) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = l[d];
            hi[d] = h[d];
        }
    }
};

class Geometry {
public:
    Geometry() = default;
    // dim: number of active directions (the reference's AMREX_SPACEDIM)
    Geometry(const Box& domain, const RealBox& rb, int dim) : domain_(domain), rb_(rb), dim_(dim) {}
    const Box& Domain() const { return domain_; }
    Real CellSize(int d) const { return (rb_.hi[d] - rb_.lo[d]) / domain_.length(d); }
    Real ProbLength(int d) const { return rb_.hi[d] - rb_.lo[d]; }
    int dim() const { return dim_; }

private:
    Box domain_;
    RealBox rb_;
    int dim_ = AMREX_SPACEDIM;
};

namespace BCType {
enum { reflect_odd = -1, int_dir = 0, reflect_even = 1, foextrap = 2, ext_dir = 3 };
}

class BCRec {
public:
    BCRec() = default;
    void setLo(int d, int v) { lo_[d] = v; }
    void setHi(int d, int v) { hi_[d] = v; }
    int lo(int d) const { return lo_[d]; }
    int hi(int d) const { return hi_[d]; }

private:
    int lo_[3] = {BCType::foextrap, BCType::foextrap, BCType::foextrap};
    int hi_[3] = {BCType::foextrap, BCType::foextrap, BCType::foextrap};
};

// Host view of one FAB's cells (AMReX Array4): p at cell begin, strides in
// elements; (i, j, k, n) in level indices.
template <class T>
struct Array4 {
    T* p = nullptr;
    long long jstride = 0, kstride = 0, nstride = 0;
    IntVect begin, end;  // first cell, one past the last
    int ncomp = 0;
    T& operator()(int i, int j, int k, int n = 0) const {
        return p[(i - begin[0]) + (j - begin[1]) * jstride + (k - begin[2]) * kstride + n * nstride];
    }
};

// Storage options of a MultiFab (AMReX's MFInfo carries allocation options;
// this one names the storage form).
struct MFInfo {
    bool fabPerBox = false;
    MFInfo& SetFabPerBox(bool v = true) {
        fabPerBox = v;
        return *this;
    }
};

// One box's cell data on the GPU (AMReX FArrayBox): ncomp components over the
// cells of box() -- the valid box grown by the ghost width --, component-major,
// i fastest.  host is a staging copy of the whole FAB (copyToHost/FromHost).
class FArrayBox {
public:
    FArrayBox(const Box& grown, int ncomp);
    ~FArrayBox();
    FArrayBox(const FArrayBox&) = delete;
    FArrayBox& operator=(const FArrayBox&) = delete;
    const Box& box() const { return box_; }
    int nComp() const { return ncomp_; }
    Real* dataPtr() const { return data_; }
    long long jstride() const { return box_.length(0); }
    long long kstride() const { return (long long)box_.length(0) * box_.length(1); }
    long long nstride() const { return box_.numPts(); }
    void copyToHost();
    void copyFromHost();
    Vector<Real> host;
    // Array4 over the host staging copy (after copyToHost)
    Array4<Real> hostArray();

private:
    Box box_;
    int ncomp_ = 0;
    Real* data_ = nullptr;
};

// One level's cell data on the GPU: ncomp components over the valid cells of
// the BoxArray's minimal box, component-major, i fastest -- or, with
// MFInfo().SetFabPerBox(), one FArrayBox per box.
class MultiFab {
public:
    MultiFab() = default;
    MultiFab(const BoxArray& ba, const DistributionMapping& dm, int ncomp, int ngrow, const MFInfo& info = MFInfo()) {
        define(ba, dm, ncomp, ngrow, info);
    }
    ~MultiFab();
    MultiFab(const MultiFab&) = delete;
    MultiFab& operator=(const MultiFab&) = delete;

    void define(const BoxArray& ba, const DistributionMapping& dm, int ncomp, int ngrow, const MFInfo& info = MFInfo());
    bool fabPerBox() const { return !fabs_.empty(); }
    // the FAB of box b (fabPerBox() only)
    FArrayBox& fab(int b) { return *fabs_.at(b); }
    const FArrayBox& fab(int b) const { return *fabs_.at(b); }
    int nComp() const { return ncomp_; }
    int nGrow() const { return ngrow_; }
    const BoxArray& boxArray() const { return ba_; }
    const Box& domain() const { return dom_; }
    long long numCells() const { return dom_.numPts(); }
    // device pointer to component 0 of cell (lo, lo, lo); strides in doubles
    // (the level array; null with fabPerBox())
    Real* devicePtr() const { return data_; }
    long long jstride() const { return dom_.length(0); }
    long long kstride() const { return (long long)dom_.length(0) * dom_.length(1); }
    long long nstride() const { return numCells(); }

    // host staging (synchronous): the whole level's valid cells, level-array
    // layout, whatever the storage
    void copyToHost(Vector<Real>& h) const;
    void copyFromHost(const Vector<Real>& h);

    static void Copy(MultiFab& dst, const MultiFab& src, int srccomp, int dstcomp, int numcomp, int nghost);

private:
    BoxArray ba_;
    Box dom_;
    int ncomp_ = 0, ngrow_ = 0;
    Real* data_ = nullptr;
    Vector<FArrayBox*> fabs_;
};

// Iterates over the boxes of a MultiFab's BoxArray.
class MFIter {
public:
    explicit MFIter(const MultiFab& mf) : ba_(&mf.boxArray()) {}
    bool isValid() const { return b_ < ba_->size(); }
    void operator++() { ++b_; }
    int index() const { return b_; }
    const Box& validbox() const { return (*ba_)[b_]; }

private:
    const BoxArray* ba_;
    int b_ = 0;
};

// Plotfile name of main.cpp:104, 176, 196: root followed by the step number
// zero-padded to `mindigits` (AMReX Concatenate).
std::string Concatenate(const std::string& root, int num, int mindigits = 5);

// One-level AMReX plotfile (the "HyperCLaw-V1.1" format AMReX tools and yt
// read): <dir>/Header, <dir>/Level_0/Cell_H and <dir>/Level_0/Cell_D_00000
// holding each box of mf's BoxArray as an IEEE-754 little-endian FAB.
// Signature of AMReX's WriteSingleLevelPlotfile as called at main.cpp:106,
// 178, 198.  Synchronises with the device (copies mf to the host).
void WriteSingleLevelPlotfile(const std::string& dir, const MultiFab& mf, const Vector<std::string>& varnames,
                              const Geometry& geom, Real time, int level_step);

}  // namespace amrex_lite
