// Drives the reference's time loop (main.cpp:121-191) through the C++ host
// interface (tenjin_host.h), on a problem read from a file written by the
// tests: the level description, the primitive initial state, finalT and a step
// cap.  Writes (t, n) and the final state of U1.
//
//   main_loop [--fab-per-box] [--box-checks] <input> <output> [plotfile root]
//
// With a plotfile root, the initial and the final state of U1 are written as
// AMReX plotfiles <root>00000 and <root><n> (main.cpp:104-107, 195-199).
// --fab-per-box: every CellArray holds one ghost-grown FArrayBox per box, as an
// AMReX MultiFab does, and the library runs on those FABs in place.
// --box-checks: after the loop, the per-box tier against the level-wide calls
// on copies of U1 (U2 box by box, U level-wide): boxPlasticUpdate / cleanUpV /
// cleanUpAlpha / conservativeToPrimitive / getSoundSpeed of each
// BoxAccessCellArray and boxTimeStepFunction's minimum must give the same bits
// as plasticUpdate / cleanUpV / cleanUpAlpha / conservativeToPrimitive and
// getTimeStep; exit status 3 if not.
//
// input : "TJIN" | int32 1 | tj_level_desc | double finalT | int32 max_steps
//         | int32 shift_pressure | int64 ncomp | int64 ncells | ncomp*ncells doubles
// output: "TJOU" | double t | int32 n | int32 nan_stop | int64 count | count doubles
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "tenjin_host.h"

namespace {

template <class T>
void get(std::ifstream& f, T& v) {
    f.read(reinterpret_cast<char*>(&v), sizeof(T));
    if (!f) throw std::runtime_error("main_loop: short input");
}

}  // namespace

int main(int argc, char** argv) {
    bool fab_per_box = false, box_checks = false;
    std::vector<std::string> pos;
    for (int a = 1; a < argc; ++a) {
        const std::string s = argv[a];
        if (s == "--fab-per-box") fab_per_box = true;
        else if (s == "--box-checks") box_checks = true;
        else pos.push_back(s);
    }
    if (pos.size() != 2 && pos.size() != 3) {
        std::fprintf(stderr, "usage: main_loop [--fab-per-box] [--box-checks] <input> <output> [plotfile root]\n");
        return 2;
    }
    const std::string plotroot = pos.size() == 3 ? pos[2] : "";
    std::ifstream in(pos[0], std::ios::binary);
    char magic[4];
    in.read(magic, 4);
    int32_t version = 0;
    get(in, version);
    if (std::memcmp(magic, "TJIN", 4) != 0 || version != 1) {
        std::fprintf(stderr, "main_loop: bad input header\n");
        return 2;
    }
    tj_level_desc desc;
    get(in, desc);
    double finalT;
    int32_t max_steps, shift;
    int64_t ncomp, ncells;
    get(in, finalT);
    get(in, max_steps);
    get(in, shift);
    get(in, ncomp);
    get(in, ncells);
    Vector<Real> prims((size_t)(ncomp * ncells));
    in.read(reinterpret_cast<char*>(prims.data()), sizeof(Real) * prims.size());
    if (!in) throw std::runtime_error("main_loop: short state");

    // ---- set-up (main.cpp:29-119 with the settings of the input)
    ParameterStruct parameters;
    for (int d = 0; d < AMREX_SPACEDIM; ++d) {
        parameters.n_cells[d] = desc.n[d];
        parameters.dimL[d] = desc.L[d];
    }
    parameters.SOLID = desc.SOLID;
    parameters.THINC = desc.THINC;
    parameters.PLASTIC = desc.PLASTIC;
    parameters.RADIAL = desc.RADIAL;
    parameters.MUSCL = desc.MUSCL;
    parameters.CFL = desc.CFL;
    parameters.THINCbeta = desc.THINCbeta;
    parameters.max_grid_size = desc.max_grid_size;
    parameters.numberOfMaterials = desc.nmat;
    parameters.REACTIVE = desc.REACTIVE;
    parameters.fabPerBox = fab_per_box ? 1 : 0;
    std::vector<std::unique_ptr<MieGruneisenEOS>> eos;
    for (int m = 0; m < desc.nmat; ++m) {
        const tj_material& t = desc.mat[m];
        if (t.eos == 1) {
            eos.emplace_back(new UsUpSolidEOS(t.rho0, t.gamma - 1.0, t.c0, t.s, t.G0, t.yield_stress));
        } else if (t.eos == 2) {  // initialise.cpp:149-160 then G0 (:209)
            eos.emplace_back(new RomenskiiSolidEOS());
            eos.back()->define({t.gamma, t.pref, t.eref, t.CV, t.rho0, t.K0, t.eos_alpha, t.eos_beta, t.G0});
            eos.back()->yieldStress = t.yield_stress;
        } else if (t.eos == 3) {  // initialise.cpp:161-172 then G0
            eos.emplace_back(new WilkinsSolidEOS());
            eos.back()->define({t.gamma, t.pref, t.eref, t.CV, t.rho0, t.e1, t.e2, t.e3, t.e4, t.e5, t.G0});
            eos.back()->yieldStress = t.yield_stress;
        } else if (t.mixture) {  // initialise.cpp:183-193: MixtureEOS::define with the eight numbers
            auto* mx = new MixtureEOS();
            mx->define({t.gamma, t.pref, t.eref, t.CV, t.gamma_mix, t.pref_mix, t.eref_mix, t.CV_mix});
            eos.emplace_back(mx);
            parameters.numberOfMixtures = 1;
        } else {
            eos.emplace_back(new MieGruneisenEOS(t.gamma, t.pref, t.eref, t.CV));
            eos.back()->shearModulus = t.G0;
            eos.back()->yieldStress = t.yield_stress;
        }
        MaterialDescriptor md;
        md.phase = t.solid ? solid : fluid;
        md.plastic = t.plastic;
        md.mixture = t.mixture != 0;
        md.EOS = eos.back().get();
        parameters.materialInfo.push_back(md);
    }

    Box domain(IntVect(0, 0, 0), IntVect(desc.n[0] - 1, desc.n[1] - 1, desc.n[2] - 1));
    BoxArray ba(domain);
    ba.maxSize(parameters.max_grid_size);
    DistributionMapping dm(ba);
    Geometry geom(domain, RealBox({0.0, 0.0, 0.0}, {desc.L[0], desc.L[1], desc.L[2]}), desc.dim);
    Vector<BCRec> bc(1);
    for (int d = 0; d < 3; ++d) {
        bc[0].setLo(d, desc.bc_lo[d] ? BCType::reflect_even : BCType::foextrap);
        bc[0].setHi(d, desc.bc_hi[d] ? BCType::reflect_even : BCType::foextrap);
    }

    AccessPattern accessPattern(parameters);
    const int Ncomp = parameters.Ncomp;
    const int Nghost = parameters.Nghost;
    if (Ncomp != ncomp) {
        std::fprintf(stderr, "main_loop: input has %lld components, layout %d\n", (long long)ncomp, Ncomp);
        return 2;
    }
    CellArray U(ba, dm, Ncomp, Nghost, accessPattern, parameters);
    CellArray U1(ba, dm, Ncomp, Nghost, accessPattern, parameters);
    CellArray U2(ba, dm, Ncomp, Nghost, accessPattern, parameters);
    CellArray UL(ba, dm, Ncomp, Nghost, accessPattern, parameters);
    CellArray UR(ba, dm, Ncomp, Nghost, accessPattern, parameters);
    CellArray MUSCLgrad(ba, dm, Ncomp, Nghost, accessPattern, parameters);
    CellArray ULStar(ba, dm, Ncomp, Nghost, accessPattern, parameters);
    CellArray URStar(ba, dm, Ncomp, Nghost, accessPattern, parameters);
    CellArray UStarStar(ba, dm, Ncomp, Nghost, accessPattern, parameters);
    Array<MultiFab, AMREX_SPACEDIM> flux_arr;
    THINCArray THINCArr(ba, dm, Nghost, parameters);
    TimeStep timeStep(ba, dm, AMREX_SPACEDIM, Nghost);
    PlasticEOS plastic;

    tenjin_host::bindLevel(parameters, geom, bc);

    // initial conditions: the primitive variables cell by cell through the
    // per-box accessor, then the pressure shift and primitiveToConservative
    U1.stage();
    const Box& dom = U1.data.domain();
    const long long nc = U1.data.nstride();
    for (MFIter mfi(U1.data); mfi.isValid(); ++mfi) {
        const Box& bx = mfi.validbox();
        BoxAccessCellArray baca(mfi, bx, U1);
        for (int k = bx.smallEnd(2); k <= bx.bigEnd(2); ++k)
            for (int j = bx.smallEnd(1); j <= bx.bigEnd(1); ++j)
                for (int i = bx.smallEnd(0); i <= bx.bigEnd(0); ++i) {
                    const long long cell = ((long long)(k - dom.lo[2]) * U1.data.kstride()) +
                                           (long long)(j - dom.lo[1]) * U1.data.jstride() + (i - dom.lo[0]);
                    for (int c = 0; c < Ncomp; ++c) {
                        // the variable and (mat | row, col) of component c
                        Variable var = ALPHA;
                        int first = -1;
                        for (auto& kv : accessPattern.data)
                            if (kv.second <= c && kv.second > first) {
                                first = kv.second;
                                var = kv.first;
                            }
                        const int off = c - first;
                        const bool tensor = var == SIGMA || var == V_TENSOR || var == VSTAR || var == DEVH;
                        const bool vec = var == RHOU || var == VELOCITY || var == RHO_MIX;
                        Real& ref = tensor ? baca(i, j, k, var, 0, off / 3, off % 3)
                                           : (vec ? baca(i, j, k, var, 0, off) : baca(i, j, k, var, off));
                        ref = prims[(size_t)c * (size_t)nc + (size_t)cell];
                    }
                }
    }
    U1.commit();
    tenjin_host::initialiseFromPrimitives(U1, shift != 0);
    if (!plotroot.empty())
        WriteSingleLevelPlotfile(Concatenate(plotroot, 0, 5), U1.data, U1.accessPattern.variableNames, geom, 0.0, 0);

    // ---- the time loop of main.cpp:121-191
    int n = 0;
    Real t = 0.0;
    int nan_stop = 0;
    for (t = 0.0; t < finalT && n < max_steps; t += parameters.dt, n++) {
        parameters.dt = timeStep.getTimeStep(U1, parameters);

        if (t + parameters.dt > finalT) {
            parameters.dt = finalT - t;
        }
        if (n < 5) {
            parameters.dt *= 0.2;
        }

        U = U1;

        advance(U, U1, U2, MUSCLgrad, UL, UR, ULStar, URStar, UStarStar, flux_arr, geom, parameters, bc, THINCArr);

        if (parameters.REACTIVE) {
            reactiveUpdate(U, U1, U2, parameters);
        }

        if (parameters.PLASTIC) {
            plastic.plasticUpdate(U1, parameters);
        }
        if (parameters.RADIAL) {
            geometricSourceTerm(U1, parameters);
        }
        if (parameters.SOLID) {
            U1.cleanUpV();
        }
        {
            U1.cleanUpAlpha();
        }
        if (U1.contains_nan()) {
            std::cout << "Nan found in U1" << std::endl;
            nan_stop = 1;
            break;
        }
    }

    if (!plotroot.empty())
        WriteSingleLevelPlotfile(Concatenate(plotroot, n, 5), U1.data, U1.accessPattern.variableNames, geom, t, n);

    int status = 0;
    if (box_checks) {
        // level-wide on U
        U = U1;
        if (parameters.PLASTIC) plastic.plasticUpdate(U, parameters);
        if (parameters.SOLID) U.cleanUpV();
        U.cleanUpAlpha();
        U.conservativeToPrimitive();
        const Real dt_level = timeStep.getTimeStep(U, parameters);
        // box by box on U2, through the per-box tier
        U2 = U1;
        U2.stage();
        Real tmin = std::numeric_limits<Real>::max();
        const int dim = desc.dim;
        for (MFIter mfi(U2.data); mfi.isValid(); ++mfi) {
            const Box& bx = mfi.validbox();
            BoxAccessCellArray baca(mfi, bx, U2);
            if (parameters.PLASTIC) plastic.boxPlasticUpdate(baca, parameters);
            if (parameters.SOLID) baca.cleanUpV();
            baca.cleanUpAlpha();
            baca.conservativeToPrimitive();
            baca.getSoundSpeed();
            // prop_arr_time over the box, as TimeStep's own MultiFab holds it
            std::vector<Real> dtb((size_t)bx.numPts() * dim);
            Array4<Real> arr;
            arr.p = dtb.data();
            arr.jstride = bx.length(0);
            arr.kstride = (long long)bx.length(0) * bx.length(1);
            arr.nstride = bx.numPts();
            arr.begin = bx.lo;
            arr.end = IntVect(bx.hi[0] + 1, bx.hi[1] + 1, bx.hi[2] + 1);
            arr.ncomp = dim;
            timeStep.boxTimeStepFunction(baca, arr, parameters);
            for (Real v : dtb) tmin = v < tmin ? v : tmin;
            if (baca.contains_nan()) status = 3;
        }
        U2.commit();
        U.stage();
        U2.stage();
        bool same = U.host.size() == U2.host.size();
        for (size_t q = 0; same && q < U.host.size(); ++q)
            same = std::memcmp(&U.host[q], &U2.host[q], sizeof(Real)) == 0;
        const Real dt_box = parameters.CFL * tmin;
        std::cout << "box checks: state " << (same ? "equal" : "DIFFERENT") << ", dt " << dt_box << " vs "
                  << dt_level << std::endl;
        if (!same || std::memcmp(&dt_box, &dt_level, sizeof(Real)) != 0) status = 3;
    }

    U1.stage();
    std::ofstream out(pos[1], std::ios::binary);
    out.write("TJOU", 4);
    out.write(reinterpret_cast<const char*>(&t), sizeof t);
    int32_t n32 = n, ns = nan_stop;
    out.write(reinterpret_cast<const char*>(&n32), sizeof n32);
    out.write(reinterpret_cast<const char*>(&ns), sizeof ns);
    int64_t count = (int64_t)U1.host.size();
    out.write(reinterpret_cast<const char*>(&count), sizeof count);
    out.write(reinterpret_cast<const char*>(U1.host.data()), sizeof(Real) * U1.host.size());
    std::printf("main_loop: n=%d t=%.17g\n", n, t);
    return status;
}
