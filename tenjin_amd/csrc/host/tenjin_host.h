// tenjin_host.h -- the reference's hot-path C++ interface, forwarding to the
// gfx950 library through its C ABI (include/tenjin.h).
//
// The names, argument lists and meanings follow the reference:
//   advance                         hllc.cpp:699-707 (declared simulationheader.h:32)
//   TimeStep::getTimeStep           timestep.cpp:36-70 (timestep.h:15-16)
//   PlasticEOS::plasticUpdate       plastic.cpp:227-244 (plastic.h:18-19)
//   geometricSourceTerm             radial.cpp:37-54 (simulationheader.h:40)
//   reactiveUpdate                  reactive.cpp:60-103 (simulationheader.h:42)
//   CellArray::{primitiveToConservative, conservativeToPrimitive,
//     getSoundSpeed, cleanUpV, cleanUpAlpha, operator=, operator*,
//     operator+, contains_nan}     cellarray.cpp:5-117 (cellarray.h:24-33)
//   ParameterStruct, MaterialDescriptor, Variable, Material_type
//                                   structdefinitions.h:10-144
//   AccessPattern                   accesspattern.cpp:44-79
// Error behaviour is the reference's: the conditions on which it prints and
// calls exit(1) print the same message and exit(1) here (checked whenever a
// call synchronises with the GPU); a NaN makes contains_nan() return true.
//
// What differs, and why:
//   * all cell data live on the GPU (amrex_lite::MultiFab); host access goes
//     through BoxAccessCellArray over a host staging copy (stage/commit);
//   * the scratch CellArrays (U2, UL, UR, MUSCLgrad, ULStar, URStar,
//     UStarStar), the flux MultiFabs and the THINCArray are accepted for
//     signature compatibility; the library keeps its own scratch, including
//     the THINC history the reference keeps in ULStar/URStar;
//   * the level must be bound once (tenjin_host::bindLevel) with the geometry
//     and BC table before the first call; advance() binds on first use.
#pragma once

#include <map>
#include <memory>
#include <string>

#include "amrex_lite.h"
#include "tenjin.h"

using namespace amrex_lite;

// structdefinitions.h:10-32 (numbering shared with tj_comp_start)
enum Variable {
    ALPHA, ALPHARHO, RHO_K, RHO, RHOU, TOTAL_E, VELOCITY, P, SOUNDSPEED, USTAR,
    RHO_MIX, ALPHARHOLAMBDA, LAMBDA, SIGMA, V_TENSOR, VSTAR, DEVH, HJ2, EPSILON,
    ALPHARHOEPSILON
};
constexpr int NUMBER_OF_VARIABLES = ALPHARHOEPSILON + 1;

enum Material_type { solid, fluid };
enum Direction_enum { x, y, z };

// MieGruneisenEOS (equationofstate.h:31-60): stiffened gas / constant-Gamma
// Mie-Gruneisen with reference pressure and energy.
// MieGruneisenEOS (equationofstate.h:33-71): the constructor's argument order
// and defaults are the reference's (equationofstate.cpp:5-12).
class MieGruneisenEOS {
public:
    MieGruneisenEOS() = default;
    MieGruneisenEOS(Real gamma, Real _pref = 0.0, Real _eref = 0.0, Real _CV = 0.0)
        : adiabaticIndex(gamma), GruneisenGamma(gamma - 1.0), pref(_pref), eref(_eref), CV(_CV) {}
    virtual ~MieGruneisenEOS() = default;
    // define(params) of equationofstate.cpp:14-26: adiabaticIndex, pref, eref, CV
    virtual void define(const Vector<Real>& params) {
        adiabaticIndex = params.at(0);
        GruneisenGamma = adiabaticIndex - 1.0;
        pref = params.at(1);
        eref = params.at(2);
        CV = params.at(3);
    }
    virtual int kind() const { return 0; }
    Real adiabaticIndex = 1.4, GruneisenGamma = 0.4, pref = 0.0, eref = 0.0, CV = 0.0;
    Real shearModulus = 0.0;
    Real yieldStress = 0.0;  // the reference keeps it in PlasticEOS::yieldStress
};

// MixtureEOS (equationofstate.h:73-106): a reactant (first) and its products
// (second), each a MieGruneisenEOS, mixed by the mass fraction lambda with
// the phases in thermal equilibrium.
class MixtureEOS : public MieGruneisenEOS {
public:
    MixtureEOS() = default;
    MixtureEOS(const MieGruneisenEOS& a, const MieGruneisenEOS& b) : first(a), second(b) {}
    // equationofstate.cpp:177-200: the first four entries the reactant's, the rest the products'
    void define(const Vector<Real>& params) override {
        first.define(Vector<Real>(params.begin(), params.begin() + 4));
        second.define(Vector<Real>(params.begin() + 4, params.end()));
    }
    int kind() const override { return 0; }  // a fluid of the MieGruneisen family (tj_material::mixture)
    MieGruneisenEOS first, second;
};

// RomenskiiSolidEOS (equationofstate.h:108-143): define(params) of
// equationofstate.cpp:403-417 reads adiabaticIndex, pref, eref, CV, rho0, K0,
// EOSalpha, EOSbeta, G0 (initialise.cpp:149-160).
class RomenskiiSolidEOS : public MieGruneisenEOS {
public:
    RomenskiiSolidEOS() = default;
    void define(const Vector<Real>& params) override {
        MieGruneisenEOS::define(params);
        rho0 = params.at(4);
        K0 = params.at(5);
        EOSalpha = params.at(6);
        EOSbeta = params.at(7);
        G0 = params.at(8);
        shearModulus = G0;
    }
    int kind() const override { return 2; }
    Real rho0 = 0.0, K0 = 0.0, EOSalpha = 0.0, EOSbeta = 0.0, G0 = 0.0;
};

// WilkinsSolidEOS (equationofstate.h:145-178): define(params) of
// equationofstate.cpp:502-518 reads adiabaticIndex, pref, eref, CV, rho0,
// e1 .. e5, G0 (initialise.cpp:161-172).
class WilkinsSolidEOS : public MieGruneisenEOS {
public:
    WilkinsSolidEOS() = default;
    void define(const Vector<Real>& params) override {
        MieGruneisenEOS::define(params);
        rho0 = params.at(4);
        e1 = params.at(5);
        e2 = params.at(6);
        e3 = params.at(7);
        e4 = params.at(8);
        e5 = params.at(9);
        G0 = params.at(10);
        shearModulus = G0;
    }
    int kind() const override { return 3; }
    Real rho0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0, e5 = 0.0, G0 = 0.0;
};

// Us-up Mie-Gruneisen solid (no reference class; the solid-EOS interface of
// equationofstate.cpp:403-596 with the principal Hugoniot as the cold curve).
class UsUpSolidEOS : public MieGruneisenEOS {
public:
    UsUpSolidEOS(Real rho0_, Real Gamma0, Real c0_, Real s_, Real G, Real Y)
        : MieGruneisenEOS(Gamma0 + 1.0, 0.0, 0.0, 1.0), rho0(rho0_), c0(c0_), s(s_) {
        shearModulus = G;
        yieldStress = Y;
    }
    int kind() const override { return 1; }
    Real rho0, c0, s;
};

struct MaterialDescriptor {
    Material_type phase = fluid;
    bool mixture = 0;
    int mixtureIndex = 0;
    int plastic = false;
    MieGruneisenEOS* EOS = nullptr;
};

namespace tenjin_host {
struct Device;
}

struct ParameterStruct {
    Vector<Real> dimL = Vector<Real>(AMREX_SPACEDIM);
    Vector<Real> dx = Vector<Real>(AMREX_SPACEDIM);
    Vector<int> n_cells = Vector<int>(AMREX_SPACEDIM);

    int SOLID = 0;
    int THINC = 0;
    int RADIAL = 0;
    int PLASTIC = 0;
    int REACTIVE = 0;
    int MUSCL = 1;

    int Ncomp = 0;
    int Nghost = 2;
    int numberOfMaterials = 0;
    int numberOfMixtures = 0;

    int max_grid_size = 64;

    // host layer: CellArrays hold AMReX's storage, one ghost-grown FArrayBox
    // per box (MFInfo().SetFabPerBox()), and the library runs on those FABs
    // in place; 0: one level array per CellArray
    int fabPerBox = 0;

    Real CFL = 0.5;
    Real x0 = 0.0;
    Real dt = 0.0;
    Real THINCbeta = 1.6;

    Vector<MaterialDescriptor> materialInfo;

    // the GPU context bound to this level (created by tenjin_host::bindLevel)
    std::shared_ptr<tenjin_host::Device> gpu;
};

// AccessPattern (accesspattern.cpp:44-79): first component of each variable.
class AccessPattern {
public:
    explicit AccessPattern(ParameterStruct& parameters);
    int& operator[](Variable var) { return data[var]; }
    std::map<Variable, int> data;
    Vector<std::string> variableNames;
    int ncomp = 0;
};

class CellArray {
public:
    CellArray(BoxArray& ba, DistributionMapping& dm, const int Ncomp, const int Nghost, AccessPattern& _accessPattern,
              ParameterStruct& parameters);
    ~CellArray();
    CellArray(const CellArray&) = delete;
    CellArray& operator=(const CellArray&) = delete;

    void primitiveToConservative();
    void conservativeToPrimitive();
    void getSoundSpeed();
    void cleanUpV();
    void cleanUpAlpha();

    void operator=(CellArray& U);
    CellArray& operator*(Real d);
    CellArray& operator+(CellArray& U);
    bool contains_nan();

    MultiFab data;
    AccessPattern& accessPattern;
    int numberOfMaterials;

    // host staging for BoxAccessCellArray: stage() copies the level to the
    // host (with fabPerBox: each FAB to its own host copy), commit() copies it
    // back
    void stage();
    void commit();
    Vector<Real> host;

    // the library's view: the level array, or (fabPerBox) a tj_fabset over the
    // FABs, made on first use after bindLevel
    tj_mf view();
    ParameterStruct& parameters;

private:
    tj_fabset* fabset_ = nullptr;
    tj_ctx* fabsetCtx_ = nullptr;
};

// structdefinitions.h:63-80: a component named by variable, material, row, col
struct MaterialSpecifier {
    Variable var;
    int mat, row, col;
    MaterialSpecifier(Variable v = RHO, int m = 0, int r = 0, int c = 0) : var(v), mat(m), row(r), col(c) {}
};

// Per-box accessor (cellarray.h:46-107, boxaccesscellarray.cpp:3-23, 382-430)
// over a host staging copy: the CellArray's (stage() first), or with fabPerBox
// the box's FArrayBox's; i, j, k are level indices.  The per-box methods run on
// the GPU for this box alone (tj_box_op): they copy the box's staged cells to
// the device, run, and stage the box's cells again, so the staged copy and the
// device agree for the box afterwards.
class BoxAccessCellArray {
public:
    BoxAccessCellArray(MFIter& mfi, const Box& bx, CellArray& U);
    BoxAccessCellArray(const Box& bx, FArrayBox& fb, CellArray& U);
    BoxAccessCellArray(const Box& bx, CellArray& U);
    const Box& box;
    CellArray& U;
    FArrayBox* fabp = nullptr;  // the FAB the accessor reads (fabPerBox), else null
    AccessPattern& accessPattern;
    int numberOfMaterials;
    static const int numberOfComponents = 3;

    Real& operator()(int i, int j, int k, MaterialSpecifier& m);
    Real& operator()(int i, int j, int k, Variable var, int mat = 0, int row = 0, int col = 0);
    Real& operator()(int i, int j, int k, int var, int mat = 0, int row = 0, int col = 0);
    Real& left(Direction_enum d, int i, int j, int k, MaterialSpecifier& m);
    Real& right(Direction_enum d, int i, int j, int k, MaterialSpecifier& m);
    Real& left(Direction_enum d, int i, int j, int k, Variable var, int mat = 0, int row = 0, int col = 0);
    Real& right(Direction_enum d, int i, int j, int k, Variable var, int mat = 0, int row = 0, int col = 0);
    Real& neighbour(int di, int dj, int dk, int i, int j, int k, MaterialSpecifier& m);
    Real& neighbour(int di, int dj, int dk, int i, int j, int k, Variable var, int mat = 0, int row = 0, int col = 0);
    void amrexToArray(int i, int j, int k, Variable var, int m, double* copy, int nx = 3, int ny = 3);

    // per-box methods on the GPU (boxaccesscellarray.cpp:65-599)
    void conservativeToPrimitive();
    void primitiveToConservative();
    void getSoundSpeed();
    void cleanUpV();
    void cleanUpAlpha();
    bool contains_nan();
    // per-cell: stress, HenckyJ2 and devH of cell (i, j, k) (stressTensor,
    // getHenckyJ2, getDeviatoricHenckyStrain; boxaccesscellarray.cpp:176-216, 326-380)
    void stressTensor(int i, int j, int k);
    void getHenckyJ2(int i, int j, int k);
    void getDeviatoricHenckyStrain(int i, int j, int k);

    // copy [lo, hi] of the staged cells to the device, run tj_box_op with ops
    // there, stage [lo, hi] again; returns the op's dt with TJ_OP_DT, and in
    // *nan whether the NaN bit was raised (the reference's fatal bits exit)
    Real deviceOp(const Box& region, unsigned ops, bool* nan = nullptr);
};

// The reference's THINC flag array; the library keeps the THINC state itself.
class THINCArray {
public:
    THINCArray() = default;
    THINCArray(BoxArray&, DistributionMapping&, const int, ParameterStruct&) {}
};

class TimeStep {
public:
    TimeStep() = default;
    TimeStep(BoxArray&, DistributionMapping&, const int, const int) {}
    // timestep.cpp:7-34: prop_arr_time(i, j, k, n) = dx[n] / (a + |u_n|) over
    // U.box, from U's staged cells (host arithmetic over the staged copy; the
    // time loop's getTimeStep runs on the GPU)
    void boxTimeStepFunction(BoxAccessCellArray& U, Array4<Real> const& prop_arr_time, ParameterStruct& parameters);
    Real getTimeStep(CellArray& U, ParameterStruct& parameters);
};

class PlasticEOS {
public:
    PlasticEOS() = default;
    void plasticUpdate(CellArray& U, ParameterStruct& parameters);
    // plastic.cpp:83-225 on U.box, on the GPU (the box's staged cells in and out)
    void boxPlasticUpdate(BoxAccessCellArray& U, ParameterStruct& parameters);
    Vector<Real> yieldStress;
};

void advance(CellArray& U, CellArray& U1, CellArray& U2, CellArray& MUSCLgrad, CellArray& UL, CellArray& UR,
             CellArray& ULStar, CellArray& URStar, CellArray& UStarStar, Array<MultiFab, AMREX_SPACEDIM>& flux_arr,
             Geometry const& geom, ParameterStruct& parameters, Vector<BCRec>& bc, THINCArray& THINC);

// reactive.cpp:60-103: Heun's RK2 of the pressure-based burn of the mixture
// material's alpha rho lambda with parameters.dt, then conservativeToPrimitive
// of U1; U is left a copy of U1's state before the burn, as in the reference.
// U2 is not used (the library keeps no scratch state for it).
void reactiveUpdate(CellArray& U, CellArray& U1, CellArray& U2, ParameterStruct& parameters);

// The cylindrical (2D, 3D) or spherical (1D) source about x = 0 with
// parameters.dt, then conservativeToPrimitive.  On solid levels the stress is
// that of the state, as plasticUpdate leaves it; RADIAL solid levels must be
// PLASTIC (bindLevel checks).
void geometricSourceTerm(CellArray& U, ParameterStruct& parameters);

namespace tenjin_host {

// One GPU context bound to one level.
struct Device {
    tj_ctx* ctx = nullptr;
    tj_level_desc desc{};
    ~Device();
};

// Binds the level described by parameters/geom/bc to a new GPU context on
// HIP device `device` and stores it in parameters.gpu (once per run, after
// the set-up of main.cpp:29-95 and setBoundaryConditions).
void bindLevel(ParameterStruct& parameters, const Geometry& geom, const Vector<BCRec>& bc, int device = 0);

// setInitialConditions' last steps (initialise.cpp:104-123): primitive
// variables already in U -> thermal-to-total pressure shift (when
// shiftPressure) and primitiveToConservative.
void initialiseFromPrimitives(CellArray& U, bool shiftPressure);

// The reference's "print and exit(1)" for fatal status bits; returns whether
// the NaN bit was set (and clears the status word).
bool checkStatus(ParameterStruct& parameters);

}  // namespace tenjin_host
