// Test infrastructure only (the oracle): a CPU restatement of the reference's
// per-cell explicit update, used by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker.  Nothing in the product links it.
//
// It follows the reference step by step, with the reference's data flow:
//   * every CellArray is a set of per-box FABs over the box grown by Nghost,
//     component-major with i fastest (AMReX FArrayBox layout);
//   * the nine CellArrays of main.cpp:64-72 and the nodal flux MultiFabs of
//     main.cpp:83-92 exist as such, and the scratch arrays keep their history
//     between passes, steps and stages exactly as in the reference;
//   * every arithmetic expression keeps the reference's operation order, and the
//     file is compiled with -ffp-contract=off, so no FMA is formed.
// Deviations (each is a place where the reference's result is undefined, racy
// or depends on an unvendored library) are listed where they occur:
//   D1 ghost fill: one fill per array that maps every ghost cell to its source
//      cell (foextrap / reflect per dimension, then the neighbouring box); the
//      reference's FillDomainBoundary-before-FillBoundary order can leave
//      out-of-domain corner ghosts of interior box corners one fill stale.
//   D2 THINC runs clean UL/UR alpha once per box right after that box's THINC
//      pass instead of once per box over the whole MultiFab (hllc.cpp:628-629).
//   D3 3x3 products sum k = 0,1,2 left to right without FMA (reference: CBLAS
//      dgemm, order unspecified); the SPD square root of cleanUpV uses a cyclic
//      Jacobi eigen-solver (reference: LAPACK dgesdd, tensor.cpp:452-475).
//   D4 scratch arrays start zero-filled (reference: allocation contents).
//   D5 exp/cosh/tanh/pow(.,-1/3) come from tjo_math.hpp (math_mode 0).
#pragma once
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "tjo_math.hpp"

namespace tjo {

// structdefinitions.h:10-32
enum Var {
    ALPHA, ALPHARHO, RHO_K, RHO, RHOU, TOTAL_E, VELOCITY, P, SOUNDSPEED, USTAR,
    RHO_MIX, ALPHARHOLAMBDA, LAMBDA, SIGMA, V_TENSOR, VSTAR, DEVH, HJ2, EPSILON,
    ALPHARHOEPSILON, NVAR
};

struct Spec { Var var; int mat, row, col; };

// EOS kinds. 0: MieGruneisenEOS of equationofstate.cpp:5-115 (stiffened gas /
// constant-Gamma Mie-Gruneisen with pref, eref).  1: Us-up Mie-Gruneisen solid
// (no reference class; it follows the solid-EOS interface pattern of
// equationofstate.cpp:502-596 with the Hugoniot as the cold curve and a
// thermodynamically consistent sound speed).
struct Material {
    int eos = 0;
    int solid = 0;
    int plastic = 0;
    double adiabaticIndex = 1.4;  // gamma; GruneisenGamma = gamma - 1
    double pref = 0.0, eref = 0.0, CV = 1.0;
    double rho0 = 1.0, c0 = 0.0, s = 0.0, G0 = 0.0, yield = 0.0;
    // MixtureEOS (equationofstate.cpp:116-376): `first` is the MieGruneisen
    // EOS above (reactant a), `second` the one below (products b)
    int mixture = 0;
    double adiabaticIndex2 = 1.4, pref2 = 0.0, eref2 = 0.0, CV2 = 1.0;
    // eos 2: RomenskiiSolidEOS (equationofstate.cpp:403-495), eos 3:
    // WilkinsSolidEOS (:502-596), both with rho0 and G0 above
    double K0 = 0.0, EOSalpha = 0.0, EOSbeta = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0, e5 = 0.0;
    double GG() const { return adiabaticIndex - 1.0; }
    double GG2() const { return adiabaticIndex2 - 1.0; }
};

struct Params {
    int dim = 1;
    int n[3] = {1, 1, 1};
    double L[3] = {1.0, 1.0, 1.0};
    int max_grid_size = 1 << 30;
    int nghost = 2;
    int SOLID = 0, THINC = 0, PLASTIC = 0, MUSCL = 1;
    int RADIAL = 0;  // geometric source of radial.cpp (structdefinitions.h:113)
    int REACTIVE = 0;  // reactiveUpdate of reactive.cpp (structdefinitions.h:115)
    int nmix = 0;      // parameters.numberOfMixtures (0 or 1: every mixture uses mixtureIndex 0)
    double THINCbeta = 1.6;
    double CFL = 0.5;
    int nmat = 2;
    Material mat[4];
    int bc_lo[3] = {0, 0, 0}, bc_hi[3] = {0, 0, 0};  // 0 transmissive, 1 reflective
    int math_mode = 0;
};

// ---------------------------------------------------------------- layout
// accesspattern.cpp:44-79 (numberOfMixtures 0 or 1)
struct Layout {
    int start[NVAR];
    int ncomp = 0;
    std::vector<Spec> cons, prim;
    std::vector<std::string> names;
    void add(int& pos, const char* base, int type /*0 prim 1 cons 2 both 3 neither*/, Var v,
             int nm, int nr, int nc) {
        if (nm * nr * nc <= 0) return;
        start[v] = pos;
        pos += nm * nr * nc;
        for (int m = 0; m < nm; ++m)
            for (int r = 0; r < nr; ++r)
                for (int c = 0; c < nc; ++c) {
                    names.push_back(std::string(base) + std::to_string(m) + "_" +
                                    std::to_string(r) + std::to_string(c));
                    Spec s{v, m, r, c};
                    if (type == 1 || type == 2) cons.push_back(s);
                    if (type == 0 || type == 2) prim.push_back(s);
                }
    }
    void build(const Params& p) {
        for (int& s : start) s = -1;
        int n = 0;
        add(n, "alpha", 2, ALPHA, p.nmat, 1, 1);
        add(n, "alphaRho", 1, ALPHARHO, p.nmat, 1, 1);
        add(n, "rho_k", 0, RHO_K, p.nmat, 1, 1);
        add(n, "rho", 3, RHO, 1, 1, 1);
        add(n, "rhoU", 1, RHOU, 1, 3, 1);
        add(n, "u", 0, VELOCITY, 1, 3, 1);
        add(n, "E", 1, TOTAL_E, 1, 1, 1);
        add(n, "p", 0, P, 1, 1, 1);
        add(n, "a", 3, SOUNDSPEED, 1, 1, 1);
        add(n, "uStar", 3, USTAR, 1, 1, 1);
        add(n, "sigma", 3, SIGMA, 1, 3, 3);
        add(n, "rhomix", 3, RHO_MIX, p.nmix, 2, 1);
        add(n, "lambda", 0, LAMBDA, p.nmix, 1, 1);
        add(n, "alpharholambda", 1, ALPHARHOLAMBDA, p.nmix, 1, 1);
        if (p.SOLID) {
            add(n, "V", 2, V_TENSOR, 1, 3, 3);
            add(n, "VStar", 3, VSTAR, 1, 3, 3);
            add(n, "devH", 3, DEVH, 1, 3, 3);
            add(n, "HenckyJ2", 3, HJ2, 1, 1, 1);
            if (p.PLASTIC) {
                add(n, "epsilon", 0, EPSILON, p.nmat, 1, 1);
                add(n, "alphaRhoEpsilon", 1, ALPHARHOEPSILON, p.nmat, 1, 1);
            }
        }
        ncomp = n;
    }
    // boxaccesscellarray.cpp:19-65
    int comp(Var v, int m, int r, int c) const {
        switch (v) {
            case ALPHA: case ALPHARHO: case RHO_K: case EPSILON: case ALPHARHOEPSILON:
                return start[v] + m;
            case RHOU: case VELOCITY:
                return start[v] + r;
            case SIGMA: case V_TENSOR: case VSTAR: case DEVH:
                return start[v] + r * 3 + c;
            case RHO_MIX:  // + mixtureIndex (0 for every mixture, initialise.cpp:187) + row
                return start[v] + r;
            default:  // LAMBDA, ALPHARHOLAMBDA: + mixtureIndex (0)
                return start[v];
        }
    }
    int comp(const Spec& s) const { return comp(s.var, s.mat, s.row, s.col); }
};

// ---------------------------------------------------------------- boxes
struct IBox { int lo[3], hi[3]; };

// BoxArray::maxSize for one domain box: per dimension, reduce by powers of two,
// cut into numblk nearly equal blocks, larger blocks first from the high end.
inline std::vector<int> chop_dim(int n, int chunk) {
    std::vector<int> cuts;  // block start positions, ascending
    if (n <= chunk) { cuts.push_back(0); return cuts; }
    int ratio = 1, bs = chunk, nlen = n;
    while (bs % 2 == 0 && nlen % 2 == 0) { ratio *= 2; bs /= 2; nlen /= 2; }
    int numblk = nlen / bs + (nlen % bs ? 1 : 0);
    int sz = nlen / numblk, extra = nlen % numblk;
    int hi = n - 1;
    std::vector<int> starts;
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

// ---------------------------------------------------------------- FAB
struct Fab {
    int glo[3], ghi[3];  // grown box
    int nx, ny, nz, nc;
    std::vector<double> d;
    void define(const int lo[3], const int hi[3], int ncomp) {
        for (int a = 0; a < 3; ++a) { glo[a] = lo[a]; ghi[a] = hi[a]; }
        nx = ghi[0] - glo[0] + 1; ny = ghi[1] - glo[1] + 1; nz = ghi[2] - glo[2] + 1;
        nc = ncomp;
        d.assign((size_t)nx * ny * nz * nc, 0.0);
    }
    double& operator()(int i, int j, int k, int n) {
        return d[(((size_t)n * nz + (k - glo[2])) * ny + (j - glo[1])) * nx + (i - glo[0])];
    }
};

template <class T> inline T delta(int i, int j) { return (T)(i == j); }
template <class T> inline int sgn(T v) { return (T(0) < v) - (v < T(0)); }
inline double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min
inline double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max

// ---------------------------------------------------------------- 3x3 algebra
// tensor.cpp:81-115 (products; D3), 140-150 (trace), 187-233 (invert), 296-333 (det)
inline void mm(const double* A, const double* B, double* C) {  // C = A B
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            C[r * 3 + c] = (A[r * 3] * B[c] + A[r * 3 + 1] * B[3 + c]) + A[r * 3 + 2] * B[6 + c];
}
inline void mmT(const double* A, const double* B, double* C) {  // C = A B^T
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            C[r * 3 + c] = (A[r * 3] * B[c * 3] + A[r * 3 + 1] * B[c * 3 + 1]) + A[r * 3 + 2] * B[c * 3 + 2];
}
inline double trace3(const double* M) { double t = 0.0; t += M[0]; t += M[4]; t += M[8]; return t; }
inline double det2(const double* M) { return M[0] * M[3] - M[1] * M[2]; }
inline double det3(const double* M) {
    double det = 0.0;
    for (int x = 0; x < 3; ++x) {
        double sub[4];
        int c = 0;
        for (int i = 1; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (j != x) sub[c++] = M[i * 3 + j];
        double sign = (x % 2) ? -1.0 : 1.0;
        det += sign * M[x] * det2(sub);
    }
    return det;
}
inline void inv3(const double* M, double* Minv) {
    double D = det3(M);
    double cof[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double sub[4];
            int c = 0;
            for (int k = 0; k < 3; ++k)
                for (int m = 0; m < 3; ++m)
                    if (k != i && m != j) sub[c++] = M[k * 3 + m];
            double sign = ((i + j) % 2) ? -1.0 : 1.0;
            cof[i * 3 + j] = sign * det2(sub);
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Minv[i * 3 + j] = cof[j * 3 + i] / D;
}
// Principal square root of a symmetric positive definite matrix (D3): cyclic
// Jacobi rotations (Golub & Van Loan 8.5) to diagonalise A = Q L Q^T, then
// Q sqrt(L) Q^T.  Fixed sweep count so the device evaluates the same sequence.
inline void sqrtm_spd(const double* Ain, double* S) {
    double a[9], q[9];
    for (int i = 0; i < 9; ++i) { a[i] = Ain[i]; q[i] = delta<double>(i / 3, i % 3); }
    const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int sweep = 0; sweep < 6; ++sweep) {
        for (int r = 0; r < 3; ++r) {
            int p = P[r], qq = Q[r];
            double apq = a[p * 3 + qq];
            if (apq == 0.0) continue;
            double app = a[p * 3 + p], aqq = a[qq * 3 + qq];
            double theta = (aqq - app) / (2.0 * apq);
            double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
            double c = 1.0 / std::sqrt(t * t + 1.0);
            double s = t * c;
            for (int k = 0; k < 3; ++k) {  // A <- A J (columns p, q)
                double akp = a[k * 3 + p], akq = a[k * 3 + qq];
                a[k * 3 + p] = c * akp - s * akq;
                a[k * 3 + qq] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) {  // A <- J^T A (rows p, q)
                double apk = a[p * 3 + k], aqk = a[qq * 3 + k];
                a[p * 3 + k] = c * apk - s * aqk;
                a[qq * 3 + k] = s * apk + c * aqk;
            }
            for (int k = 0; k < 3; ++k) {  // Q <- Q J
                double qkp = q[k * 3 + p], qkq = q[k * 3 + qq];
                q[k * 3 + p] = c * qkp - s * qkq;
                q[k * 3 + qq] = s * qkp + c * qkq;
            }
        }
    }
    double sl[3] = {std::sqrt(a[0]), std::sqrt(a[4]), std::sqrt(a[8])};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            S[r * 3 + c] = (q[r * 3] * sl[0] * q[c * 3] + q[r * 3 + 1] * sl[1] * q[c * 3 + 1]) +
                           q[r * 3 + 2] * sl[2] * q[c * 3 + 2];
}

// ---------------------------------------------------------------- solver
struct CellArray {
    std::vector<Fab> f;
};

class Solver {
  public:
    Params p;
    Layout L;
    Math mth;
    std::vector<IBox> boxes;
    std::vector<int> cuts[3];
    double dx[3];
    double dt = 0.0;
    CellArray U, U1, U2, UL, UR, ULStar, URStar, UStarStar, MUSCLgrad;
    std::vector<Fab> flux[3];     // nodal, 0 ghosts
    std::vector<std::vector<int>> thincFlags;  // per box, (nmat+1) ints per grown cell
    std::vector<Fab> tsfab;       // TimeStep MultiFab (SPACEDIM comps)
    int ghost_lo[3], ghost_hi[3];

    explicit Solver(const Params& prm) : p(prm) {
        L.build(p);
        mth.mode = p.math_mode;
        for (int a = 0; a < 3; ++a) {
            if (a < p.dim) {
                cuts[a] = chop_dim(p.n[a], p.max_grid_size);
                dx[a] = p.L[a] / p.n[a];
            } else {
                cuts[a] = {0};
                dx[a] = 1.0;
            }
        }
        for (size_t kz = 0; kz < cuts[2].size(); ++kz)
            for (size_t jy = 0; jy < cuts[1].size(); ++jy)
                for (size_t ix = 0; ix < cuts[0].size(); ++ix) {
                    IBox b;
                    size_t id[3] = {ix, jy, kz};
                    for (int a = 0; a < 3; ++a) {
                        int na = a < p.dim ? p.n[a] : 1;
                        b.lo[a] = cuts[a][id[a]];
                        b.hi[a] = (id[a] + 1 < cuts[a].size() ? cuts[a][id[a] + 1] : na) - 1;
                    }
                    boxes.push_back(b);
                }
        CellArray* all[] = {&U, &U1, &U2, &UL, &UR, &ULStar, &URStar, &UStarStar, &MUSCLgrad};
        for (CellArray* ca : all) {
            ca->f.resize(boxes.size());
            for (size_t b = 0; b < boxes.size(); ++b) {
                int glo[3], ghi[3];
                grown(boxes[b], glo, ghi);
                ca->f[b].define(glo, ghi, L.ncomp);
            }
        }
        for (int d = 0; d < p.dim; ++d) {
            flux[d].resize(boxes.size());
            for (size_t b = 0; b < boxes.size(); ++b) {
                int lo[3], hi[3];
                for (int a = 0; a < 3; ++a) { lo[a] = boxes[b].lo[a]; hi[a] = boxes[b].hi[a]; }
                hi[d] += 1;
                flux[d][b].define(lo, hi, L.ncomp);
            }
        }
        thincFlags.resize(boxes.size());
        tsfab.resize(boxes.size());
        for (size_t b = 0; b < boxes.size(); ++b) {
            const Fab& g = U.f[b];
            thincFlags[b].assign((size_t)g.nx * g.ny * g.nz * (p.nmat + 1), 0);
            tsfab[b].define(boxes[b].lo, boxes[b].hi, p.dim);
        }
    }

    void grown(const IBox& b, int glo[3], int ghi[3]) const {
        for (int a = 0; a < 3; ++a) {
            int g = a < p.dim ? p.nghost : 0;
            glo[a] = b.lo[a] - g;
            ghi[a] = b.hi[a] + g;
        }
    }

    // ------------------------------------------------------------ accessors
    double& A(Fab& f, int i, int j, int k, Var v, int m = 0, int r = 0, int c = 0) {
        return f(i, j, k, L.comp(v, m, r, c));
    }
    double& A(Fab& f, int i, int j, int k, const Spec& s) { return f(i, j, k, L.comp(s)); }
    static void shift(int d, int s, int& i, int& j, int& k) {
        if (d == 0) i += s; else if (d == 1) j += s; else k += s;
    }

    // ------------------------------------------------------------ EOS
    // equationofstate.cpp:37-115 (kind 0); Us-up solid (kind 1), see Material.
    double pH(const Material& M, double rho) const {
        double eta = 1.0 - M.rho0 / rho;
        double den = 1.0 - M.s * eta;
        return M.rho0 * M.c0 * M.c0 * eta / (den * den);
    }
    double eH(const Material& M, double rho) const {
        double eta = 1.0 - M.rho0 / rho;
        return pH(M, rho) * eta / (2.0 * M.rho0);
    }
    double dpHdrho(const Material& M, double rho) const {
        double eta = 1.0 - M.rho0 / rho;
        double den = 1.0 - M.s * eta;
        double detadrho = M.rho0 / (rho * rho);
        return M.rho0 * M.c0 * M.c0 * (1.0 + M.s * eta) / (den * den * den) * detadrho;
    }
    double deHdrho(const Material& M, double rho) const {
        double eta = 1.0 - M.rho0 / rho;
        double detadrho = M.rho0 / (rho * rho);
        return (dpHdrho(M, rho) * eta + pH(M, rho) * detadrho) / (2.0 * M.rho0);
    }
    // RomenskiiSolidEOS / WilkinsSolidEOS terms (equationofstate.cpp:420-495, 519-596);
    // std::pow and std::log as the project's routines (D5)
    double romG(const Material& M, double rk) const { return M.G0 * mth.pow((rk / M.rho0), M.EOSbeta + 1.0); }
    double solidDGdrho(const Material& M, double rk) const {
        if (M.eos == 2) return M.G0 * (M.EOSbeta + 1.0) * mth.pow((rk / M.rho0), M.EOSbeta) / M.rho0;
        return 0.0;
    }
    double solidDG2drho2(const Material& M, double rk) const {
        if (M.eos == 2) return M.G0 * (M.EOSbeta + 1.0) * (M.EOSbeta) * mth.pow((rk / M.rho0), M.EOSbeta - 1.0) / (M.rho0 * M.rho0);
        return 0.0;
    }
    double solidDpcdrho(const Material& M, double rk) const {
        const double x = rk / M.rho0;
        if (M.eos == 2)
            return (M.K0 / (M.EOSalpha * M.rho0)) * mth.pow(x, M.EOSalpha) *
                   ((2.0 * M.EOSalpha + 1.0) * mth.pow(x, M.EOSalpha) - (M.EOSalpha + 1.0));
        return (6.0 * M.e2 * x * x + 2.0 * M.e3 * x - M.e5) / M.rho0;
    }
    double coldE(Fab& f, int i, int j, int k, int m) {
        const Material& M = p.mat[m];
        if (M.eos == 1) return A(f, i, j, k, ALPHARHO, m) * eH(M, A(f, i, j, k, RHO_K, m));
        if (M.eos == 2) {
            const double x = A(f, i, j, k, RHO_K, m) / M.rho0;
            return A(f, i, j, k, ALPHARHO, m) * ((M.K0 / (2.0 * M.rho0 * M.EOSalpha * M.EOSalpha)) *
                                                 (mth.pow(x, M.EOSalpha) - 1.0) * (mth.pow(x, M.EOSalpha) - 1.0));
        }
        if (M.eos == 3) {
            const double x = A(f, i, j, k, RHO_K, m) / M.rho0;
            return A(f, i, j, k, ALPHARHO, m) * ((M.e1 + M.e2 * x * x + M.e3 * x + M.e4 / x - M.e5 * mth.log(x)) / M.rho0);
        }
        if (M.mixture) {  // equationofstate.cpp:158-162
            const double lam = A(f, i, j, k, LAMBDA, m);
            return lam * (M.eref * A(f, i, j, k, ALPHARHO, m)) + (1.0 - lam) * (M.eref2 * A(f, i, j, k, ALPHARHO, m));
        }
        return M.eref * A(f, i, j, k, ALPHARHO, m);
    }
    double coldP(Fab& f, int i, int j, int k, int m) {
        const Material& M = p.mat[m];
        if (M.eos == 1) return (A(f, i, j, k, ALPHA, m) / M.GG()) * pH(M, A(f, i, j, k, RHO_K, m));
        if (M.eos == 2) {
            const double x = A(f, i, j, k, RHO_K, m) / M.rho0;
            return (A(f, i, j, k, ALPHA, m) / M.GG()) *
                   ((M.K0 / M.EOSalpha) * mth.pow(x, M.EOSalpha + 1.0) * (mth.pow(x, M.EOSalpha) - 1.0));
        }
        if (M.eos == 3) {
            const double x = A(f, i, j, k, RHO_K, m) / M.rho0;
            return (A(f, i, j, k, ALPHA, m) / M.GG()) * ((2.0 * M.e2 * x * x * x + M.e3 * x * x - M.e4 - M.e5 * x));
        }
        if (M.mixture) {  // equationofstate.cpp:164-168
            const double lam = A(f, i, j, k, LAMBDA, m), rk = A(f, i, j, k, RHO_K, m);
            const double pa = M.pref * A(f, i, j, k, ALPHA, m) / M.GG();
            const double pb = M.pref2 * A(f, i, j, k, ALPHA, m) / M.GG2();
            return (pa)*lam * rk / (A(f, i, j, k, RHO_MIX, m, 0)) + (pb) * (1.0 - lam) * rk / (A(f, i, j, k, RHO_MIX, m, 1));
        }
        return M.pref * A(f, i, j, k, ALPHA, m) / M.GG();
    }
    // equationofstate.cpp:170-175, 233-236: the mixture's Gruneisen coefficient
    // from the CV-weighted adiabatic index
    double mixGamma(Fab& f, int i, int j, int k, int m) {
        const Material& M = p.mat[m];
        const double lam = A(f, i, j, k, LAMBDA, m);
        return (lam * M.adiabaticIndex * M.CV + (1.0 - lam) * M.adiabaticIndex2 * M.CV2) / (lam * M.CV + (1.0 - lam) * M.CV2) - 1.0;
    }
    // componentShearModulus (equationofstate.cpp:77-80, 447-450, 548-551)
    double shearG(Fab& f, int i, int j, int k, int m) {
        const Material& M = p.mat[m];
        if (M.eos == 1 || M.eos == 3) return M.G0;
        if (M.eos == 2) return romG(M, A(f, i, j, k, RHO_K, m));
        return 0.0;
    }
    double shearE(Fab& f, int i, int j, int k, int m) {
        if (p.mat[m].eos >= 1) return A(f, i, j, k, ALPHA, m) * (shearG(f, i, j, k, m) * A(f, i, j, k, HJ2));
        return 0.0 * A(f, i, j, k, ALPHA, m);
    }
    double shearP(Fab& f, int i, int j, int k, int m) {
        const Material& M = p.mat[m];
        if (M.eos == 1)
            return (A(f, i, j, k, ALPHA, m) / M.GG()) * ((A(f, i, j, k, RHO_K, m) * 0.0 - shearG(f, i, j, k, m)) * A(f, i, j, k, HJ2));
        if (M.eos >= 2)
            return (A(f, i, j, k, ALPHA, m) / M.GG()) *
                   ((A(f, i, j, k, RHO_K, m) * solidDGdrho(M, A(f, i, j, k, RHO_K, m)) - shearG(f, i, j, k, m)) * A(f, i, j, k, HJ2));
        return 0.0 * A(f, i, j, k, ALPHA, m);
    }
    double invGr(Fab& f, int i, int j, int k, int m) {
        if (p.mat[m].mixture) return A(f, i, j, k, ALPHA, m) / (mixGamma(f, i, j, k, m));
        return A(f, i, j, k, ALPHA, m) / p.mat[m].GG();
    }
    double xi(Fab& f, int i, int j, int k, int m) {
        if (p.mat[m].mixture) return 1.0 / (mixGamma(f, i, j, k, m));
        return 1.0 / p.mat[m].GG();
    }
    double ssContrib(Fab& f, int i, int j, int k, int m) {
        const Material& M = p.mat[m];
        double rk = A(f, i, j, k, RHO_K, m);
        double pp = A(f, i, j, k, P);
        if (M.mixture) {  // mixtureSoundSpeed, equationofstate.cpp:212-226 (constant pref terms)
            const double lam = A(f, i, j, k, LAMBDA, m);
            double dedrho = (lam) * ((M.pref - pp) / (M.GG() * rk * rk));
            dedrho += (1.0 - lam) * ((M.pref2 - pp) / (M.GG2() * rk * rk));
            return (pp / (rk * rk) - dedrho) /
                   (lam / (A(f, i, j, k, RHO_MIX, m, 0) * M.GG()) + (1.0 - lam) / (A(f, i, j, k, RHO_MIX, m, 1) * M.GG2()));
        }
        if (M.eos >= 2) {  // equationofstate.cpp:435-438, 536-539
            return (pp * (M.GG() + 1.0) - (coldP(f, i, j, k, m) + shearP(f, i, j, k, m)) * (M.GG() + 1.0)) / rk +
                   solidDpcdrho(M, rk) + rk * solidDG2drho2(M, rk) * A(f, i, j, k, HJ2) +
                   (4.0 / 3.0) * shearG(f, i, j, k, m) / rk;
        }
        if (M.eos == 1) {
            double gp1 = M.GG() + 1.0;
            double hj2 = A(f, i, j, k, HJ2);
            return (pp * gp1 - pH(M, rk) + gp1 * M.G0 * hj2) / rk + dpHdrho(M, rk) -
                   M.GG() * rk * deHdrho(M, rk) + (4.0 / 3.0) * M.G0 / rk;
        }
        return pp * (M.GG() + 1.0) / rk - M.pref / rk;
    }
    // boxaccesscellarray.cpp:218-252
    double effInvG(Fab& f, int i, int j, int k) {
        double t = 0.0;
        for (int m = 0; m < p.nmat; ++m) t += invGr(f, i, j, k, m);
        return t;
    }
    double effNTE(Fab& f, int i, int j, int k) {
        double t = 0.0;
        for (int m = 0; m < p.nmat; ++m) t += (coldE(f, i, j, k, m) + shearE(f, i, j, k, m));
        return t;
    }
    double effNTP(Fab& f, int i, int j, int k) {
        double t = 0.0;
        for (int m = 0; m < p.nmat; ++m) t += coldP(f, i, j, k, m) + shearP(f, i, j, k, m);
        return t;
    }

    // ------------------------------------------------------------ per-cell thermo
    // boxaccesscellarray.cpp:326-380 (A12: single shared V; devH = (G - G^-1)/4)
    void henckyJ2(Fab& f, int i, int j, int k) {
        bool anySolid = false;
        for (int m = 0; m < p.nmat; ++m) anySolid |= p.mat[m].solid != 0;
        if (!anySolid) return;
        double V[9], G[9], Gi[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) V[r * 3 + c] = A(f, i, j, k, V_TENSOR, 0, r, c);
        mmT(V, V, G);
        inv3(G, Gi);
        double dH[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                dH[r * 3 + c] = 0.5 * 0.5 * (G[r * 3 + c] - Gi[r * 3 + c]);
                A(f, i, j, k, DEVH, 0, r, c) = dH[r * 3 + c];
            }
        double prod[9];
        mmT(dH, dH, prod);
        A(f, i, j, k, HJ2) = trace3(prod);
    }
    // boxaccesscellarray.cpp:176-216
    void stress(Fab& f, int i, int j, int k) {
        double Gt = 0.0;
        int solid = 0;
        for (int m = 0; m < p.nmat; ++m)
            if (p.mat[m].solid) {
                solid = 1;
                Gt += shearG(f, i, j, k, m) * invGr(f, i, j, k, m);
            }
        Gt = Gt / effInvG(f, i, j, k);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                if (solid)
                    A(f, i, j, k, SIGMA, 0, r, c) = -A(f, i, j, k, P) * delta<double>(r, c) + 2.0 * Gt * A(f, i, j, k, DEVH, 0, r, c);
                else
                    A(f, i, j, k, SIGMA, 0, r, c) = -A(f, i, j, k, P) * delta<double>(r, c);
            }
    }
    // ------------------------------------------------------------ mixture EOS
    // equationofstate.cpp:239-247 rhobFunc: the products' density from the
    // reactants' at the cell's mixture density and mass fraction
    double rhob_of(Fab& f, int i, int j, int k, int m, double rhoa) {
        return (1.0 - A(f, i, j, k, LAMBDA, m)) / (1.0 / A(f, i, j, k, RHO_K, m) - A(f, i, j, k, LAMBDA, m) / rhoa);
    }
    // :249-262 pressureFunction: difference of the two phases' temperatures
    double pressure_fn(Fab& f, int i, int j, int k, int m, double pp) {
        const Material& M = p.mat[m];
        return ((pp - M.pref) / (A(f, i, j, k, RHO_MIX, m, 0) * M.GG() * M.CV)) -
               ((pp - M.pref2) / (A(f, i, j, k, RHO_MIX, m, 1) * M.GG2() * M.CV2));
    }
    // :264-273 bisectionFunction (stores the trial densities in RHO_MIX)
    double bisection_fn(Fab& f, int i, int j, int k, int m, double rhoTry, double ke) {
        A(f, i, j, k, RHO_MIX, m, 0) = rhoTry;
        A(f, i, j, k, RHO_MIX, m, 1) = rhob_of(f, i, j, k, m, rhoTry);
        const double pp = (A(f, i, j, k, TOTAL_E) - ke - effNTE(f, i, j, k) + effNTP(f, i, j, k)) / (effInvG(f, i, j, k));
        return pressure_fn(f, i, j, k, m, pp);
    }
    static int sgn(double v) { return (0.0 < v) - (v < 0.0); }
    // Bisection bound: the reference loops until convergence with no bound
    // (equationofstate.cpp:324-353); the interval halves each pass, so a
    // finite bracket converges in under 100 passes.
    static constexpr int kMixMaxIter = 2000;
    std::atomic<int> mix_fatal{0};  // 1: no sign change ("Error in root Bisection", exit); 2: bound hit
    // :275-354 rootFind: thermal equilibrium of the two phases by bisection on
    // the reactants' density
    void root_find(Fab& f, int i, int j, int k, int m, double ke) {
        const Material& M = p.mat[m];
        const double toleranceForSinglePhaseTreatment = 1E-3;
        const double toleranceForConvergence = 1E-7;
        const double toleranceForBeingNearRoot = 1E-4;
        double& ra = A(f, i, j, k, RHO_MIX, m, 0);
        double& rb = A(f, i, j, k, RHO_MIX, m, 1);
        const double lam = A(f, i, j, k, LAMBDA, m), rk = A(f, i, j, k, RHO_K, m);
        if (M.pref == 0.0 && M.pref2 == 0.0) {
            ra = rk * (lam + (1.0 - lam) * (M.GG2() * M.CV2) / (M.GG() * M.CV));
            rb = rhob_of(f, i, j, k, m, ra);
            return;
        }
        if (lam > 1.0 - toleranceForSinglePhaseTreatment) {
            ra = rk;
            rb = rk;
            return;
        } else if (lam < toleranceForSinglePhaseTreatment) {
            rb = rk;
            ra = rk;
            return;
        }
        double a = lam * rk + 1E-10;
        double b = 1000000.0;
        double mid = 0.5 * (a + b);
        if (std::fabs(bisection_fn(f, i, j, k, m, a, ke)) <= toleranceForBeingNearRoot) {
            ra = a;
            rb = rhob_of(f, i, j, k, m, ra);
            return;
        } else if (std::fabs(bisection_fn(f, i, j, k, m, b, ke)) <= toleranceForBeingNearRoot) {
            ra = b;
            rb = rhob_of(f, i, j, k, m, ra);
            return;
        }
        const int sa = sgn(bisection_fn(f, i, j, k, m, a, ke));
        const int sb = sgn(bisection_fn(f, i, j, k, m, b, ke));
        if (sa == sb) {  // fatal in the reference; the trial densities of b stay
            mix_fatal |= 1;
            return;
        }
        for (int n = 0;; n++) {
            const int smid = sgn(bisection_fn(f, i, j, k, m, mid, ke));
            if (smid == sgn(bisection_fn(f, i, j, k, m, a, ke))) a = mid;
            else b = mid;
            mid = 0.5 * (a + b);
            if (std::fabs(a - b) / (a + b) < toleranceForConvergence || n >= kMixMaxIter) {
                if (n >= kMixMaxIter) mix_fatal |= 2;
                ra = mid;
                rb = rhob_of(f, i, j, k, m, ra);
                return;
            }
        }
    }
    // :364-370 defineMixtureDensities: the phase densities of a primitive state
    void define_mix(Fab& f, int i, int j, int k, int m) {
        const Material& M = p.mat[m];
        const double lam = A(f, i, j, k, LAMBDA, m), pp = A(f, i, j, k, P);
        A(f, i, j, k, RHO_MIX, m, 0) =
            A(f, i, j, k, RHO_K, m) * (lam + (1.0 - lam) * ((pp - M.pref) / (pp - M.pref2)) * ((M.CV2 * M.GG2()) / (M.CV * M.GG())));
        A(f, i, j, k, RHO_MIX, m, 1) = rhob_of(f, i, j, k, m, A(f, i, j, k, RHO_MIX, m, 0));
    }

    // boxaccesscellarray.cpp:67-122
    void c2p_cell(Fab& f, int i, int j, int k) {
        A(f, i, j, k, RHO) = 0;
        for (int m = 0; m < p.nmat; ++m) {
            A(f, i, j, k, RHO_K, m) = A(f, i, j, k, ALPHARHO, m) / A(f, i, j, k, ALPHA, m);
            A(f, i, j, k, RHO) += A(f, i, j, k, ALPHARHO, m);
            if (p.mat[m].plastic)
                A(f, i, j, k, EPSILON, m) = A(f, i, j, k, ALPHARHOEPSILON, m) / A(f, i, j, k, ALPHARHO, m);
        }
        henckyJ2(f, i, j, k);
        double ke = 0.0;
        for (int r = 0; r < 3; ++r) {
            A(f, i, j, k, VELOCITY, 0, r) = A(f, i, j, k, RHOU, 0, r) / A(f, i, j, k, RHO);
            ke += 0.5 * A(f, i, j, k, RHO) * A(f, i, j, k, VELOCITY, 0, r) * A(f, i, j, k, VELOCITY, 0, r);
        }
        for (int m = 0; m < p.nmat; ++m)
            if (p.mat[m].mixture) {
                A(f, i, j, k, LAMBDA, m) = A(f, i, j, k, ALPHARHOLAMBDA, m) / A(f, i, j, k, ALPHARHO, m);
                root_find(f, i, j, k, m, ke);
            }
        A(f, i, j, k, P) = (A(f, i, j, k, TOTAL_E) - ke - effNTE(f, i, j, k) + effNTP(f, i, j, k)) / effInvG(f, i, j, k);
        stress(f, i, j, k);
    }
    // boxaccesscellarray.cpp:124-174
    void p2c_cell(Fab& f, int i, int j, int k) {
        A(f, i, j, k, RHO) = 0;
        for (int m = 0; m < p.nmat; ++m) {
            A(f, i, j, k, ALPHARHO, m) = A(f, i, j, k, ALPHA, m) * A(f, i, j, k, RHO_K, m);
            A(f, i, j, k, RHO) += A(f, i, j, k, ALPHARHO, m);
            if (p.mat[m].mixture) {
                A(f, i, j, k, ALPHARHOLAMBDA, m) = A(f, i, j, k, LAMBDA, m) * A(f, i, j, k, ALPHA, m) * A(f, i, j, k, RHO_K, m);
                define_mix(f, i, j, k, m);
            }
            if (p.mat[m].plastic)
                A(f, i, j, k, ALPHARHOEPSILON, m) = A(f, i, j, k, EPSILON, m) * A(f, i, j, k, ALPHARHO, m);
        }
        henckyJ2(f, i, j, k);
        double ke = 0.0;
        for (int r = 0; r < 3; ++r) {
            A(f, i, j, k, RHOU, 0, r) = A(f, i, j, k, VELOCITY, 0, r) * A(f, i, j, k, RHO);
            ke += 0.5 * A(f, i, j, k, RHO) * A(f, i, j, k, VELOCITY, 0, r) * A(f, i, j, k, VELOCITY, 0, r);
        }
        A(f, i, j, k, TOTAL_E) = A(f, i, j, k, P) * effInvG(f, i, j, k) + effNTE(f, i, j, k) - effNTP(f, i, j, k) + ke;
        stress(f, i, j, k);
    }
    // boxaccesscellarray.cpp:254-285
    void sound_cell(Fab& f, int i, int j, int k) {
        double a = 0.0, xt = 0.0;
        for (int m = 0; m < p.nmat; ++m) {
            a += dmax(0.0, xi(f, i, j, k, m) * ssContrib(f, i, j, k, m) * A(f, i, j, k, ALPHARHO, m) / A(f, i, j, k, RHO));
            xt += xi(f, i, j, k, m) * A(f, i, j, k, ALPHA, m);
        }
        if (a <= 0.0) a = 1E-6;
        A(f, i, j, k, SOUNDSPEED) = std::sqrt(a / xt);
    }
    // boxaccesscellarray.cpp:287-324
    double transverse(Fab& f, int i, int j, int k) {
        double b = 0.0;
        henckyJ2(f, i, j, k);
        for (int m = 0; m < p.nmat; ++m)
            if (p.mat[m].solid) b += invGr(f, i, j, k, m) * shearG(f, i, j, k, m) / A(f, i, j, k, RHO_K, m);
        if (std::isnan(b) || b < 0.0) b = 0.0;
        double t = std::sqrt(b / effInvG(f, i, j, k));
        if (std::isnan(t)) t = 0.0;
        return t;
    }
    // boxaccesscellarray.cpp:530-599 (the Nan index quirk of :553-556 kept)
    int cleanup_alpha_cell(Fab& f, int i, int j, int k) {
        double tot = 0.0;
        int nan = 0;
        for (int m = 0; m < p.nmat; ++m) {
            double& a = A(f, i, j, k, ALPHA, m);
            if (std::isnan(a)) {
                a = 0.0;
                if (nan == 0) nan = m;
                else return 1;  // "too many Nans": fatal in the reference
            } else if (a < 0.0) {
                a = 1E-6;
            } else if (a > 1.0) {
                a = 1.0;
            }
            tot += a;
        }
        if (nan > 0) {
            A(f, i, j, k, ALPHA, nan) = 1.0 - tot;
            tot += A(f, i, j, k, ALPHA, nan);
        }
        if (tot <= 0.0) return 2;  // fatal in the reference
        for (int m = 0; m < p.nmat; ++m) A(f, i, j, k, ALPHA, m) *= 1.0 / tot;
        return 0;
    }

    template <class F>
    void for_valid(int b, F&& fn) {
        const IBox& bx = boxes[b];
        for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
            for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
                for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) fn(i, j, k);
    }

    // ------------------------------------------------------------ ghost fill (D1)
    // boundary.cpp:3-42 BC types; FillDomainBoundary + FillBoundary semantics.
    int box_of(const int c[3]) const {
        size_t id[3];
        for (int a = 0; a < 3; ++a) {
            const std::vector<int>& cu = cuts[a];
            id[a] = (size_t)(std::upper_bound(cu.begin(), cu.end(), c[a]) - cu.begin()) - 1;
        }
        return (int)((id[2] * cuts[1].size() + id[1]) * cuts[0].size() + id[0]);
    }
    void fill(CellArray& ca) {
        const int nc = L.ncomp;
        const int uo = L.start[VELOCITY], mo = L.start[RHOU];
#pragma omp parallel for schedule(dynamic)
        for (size_t b = 0; b < boxes.size(); ++b) {
            Fab& f = ca.f[b];
            const IBox& bx = boxes[b];
            for (int k = f.glo[2]; k <= f.ghi[2]; ++k)
                for (int j = f.glo[1]; j <= f.ghi[1]; ++j)
                    for (int i = f.glo[0]; i <= f.ghi[0]; ++i) {
                        int c[3] = {i, j, k};
                        bool valid = true;
                        for (int a = 0; a < 3; ++a) valid &= (c[a] >= bx.lo[a] && c[a] <= bx.hi[a]);
                        if (valid) continue;
                        int s[3];
                        bool flip[3] = {false, false, false};
                        for (int a = 0; a < 3; ++a) {
                            int n = a < p.dim ? p.n[a] : 1;
                            s[a] = c[a];
                            if (c[a] < 0) {
                                if (p.bc_lo[a] == 1) { s[a] = -1 - c[a]; flip[a] = true; }
                                else s[a] = 0;
                            } else if (c[a] > n - 1) {
                                if (p.bc_hi[a] == 1) { s[a] = 2 * n - 1 - c[a]; flip[a] = true; }
                                else s[a] = n - 1;
                            }
                        }
                        int sb = box_of(s);
                        Fab& g = ca.f[sb];
                        for (int n = 0; n < nc; ++n) {
                            double v = g(s[0], s[1], s[2], n);
                            for (int a = 0; a < 3; ++a)
                                if (flip[a] && (n == uo + a || n == mo + a)) v = -v;
                            f(i, j, k, n) = v;
                        }
                    }
        }
    }

    // ------------------------------------------------------------ MultiFab algebra
    static void copy(CellArray& dst, CellArray& src) {  // cellarray.cpp:80-84
#pragma omp parallel for schedule(static)
        for (size_t b = 0; b < dst.f.size(); ++b) dst.f[b].d = src.f[b].d;
    }

    // ------------------------------------------------------------ MUSCL
    // hllc.cpp:527-575
    static double vanleer(double r) {
        if (r <= 0.0) return 0.0;
        return dmin(2.0 / (1.0 + r), 2.0 * r / (1.0 + r));
    }
    void muscl(int b, int d) {
        Fab& u = U_in->f[b];
        Fab& ul = UL.f[b];
        Fab& ur = UR.f[b];
        Fab& gr = MUSCLgrad.f[b];
        for (const Spec& s : L.prim) {
            int n = L.comp(s);
            for_valid(b, [&](int i, int j, int k) {
                int il = i, jl = j, kl = k, ir = i, jr = j, kr = k;
                shift(d, -1, il, jl, kl);
                shift(d, 1, ir, jr, kr);
                double uc = u(i, j, k, n), um = u(il, jl, kl, n), up = u(ir, jr, kr, n);
                double r = (uc - um) / (up - uc);
                if (std::isinf(r) || std::isnan(r)) r = 0.0;
                gr(i, j, k, n) = vanleer(r) * 0.5 * (up - um);
                ul(i, j, k, n) = uc - 0.5 * gr(i, j, k, n);
                ur(i, j, k, n) = uc + 0.5 * gr(i, j, k, n);
            });
        }
    }

    // ------------------------------------------------------------ THINC / BVD
    // thinc.cpp:19-208 (A6, A7)
    int& mixedFlag(int b, int i, int j, int k) {
        const Fab& g = U.f[b];
        size_t cell = ((size_t)(k - g.glo[2]) * g.ny + (j - g.glo[1])) * g.nx + (i - g.glo[0]);
        return thincFlags[b][cell];
    }
    int& tbvFlag(int b, int i, int j, int k, int m) {
        const Fab& g = U.f[b];
        size_t ncell = (size_t)g.nx * g.ny * g.nz;
        size_t cell = ((size_t)(k - g.glo[2]) * g.ny + (j - g.glo[1])) * g.nx + (i - g.glo[0]);
        return thincFlags[b][(1 + m) * ncell + cell];
    }
    double youngs(Fab& u, int d, int m, int i, int j, int k) {
        auto a = [&](int di, int dj) { return A(u, i + di, j + dj, k, ALPHA, m); };
        double nx = -1.0 / (8.0 * dx[0]) * (a(1, 1) + 2.0 * a(1, 0) + a(1, -1) - a(-1, 1) - 2.0 * a(-1, 0) - a(-1, -1));
        double ny = -1.0 / (8.0 * dx[1]) * (a(1, 1) + 2.0 * a(0, 1) + a(-1, 1) - a(-1, -1) - 2.0 * a(0, -1) - a(1, -1));
        double norm = std::sqrt(nx * nx + ny * ny);
        if (norm < 1E-20) return 0;
        return d == 0 ? std::fabs(nx / norm) : std::fabs(ny / norm);
    }
    void thinc(int b, int d) {
        Fab& u = U_in->f[b];
        Fab& ul = UL.f[b];
        Fab& ur = UR.f[b];
        Fab& tl = ULStar.f[b];
        Fab& tr = URStar.f[b];
        const double beta0 = p.THINCbeta, eps = 1E-20;
        for_valid(b, [&](int i, int j, int k) {
            int il = i, jl = j, kl = k, ir = i, jr = j, kr = k;
            shift(d, -1, il, jl, kl);
            shift(d, 1, ir, jr, kr);
            for (int m = 0; m < p.nmat; ++m) {
                double aL = A(u, il, jl, kl, ALPHA, m), aR = A(u, ir, jr, kr, ALPHA, m), aC = A(u, i, j, k, ALPHA, m);
                double mn = dmin(aL, aR);
                double mx = dmax(aL, aR) - mn;
                double theta = (double)sgn(aR - aL);
                double C = (mx < eps) ? 0.0 : (aC - mn + eps) / (mx + eps);
                mixedFlag(b, i, j, k) = 0;
                const double dl = 1E-4;
                bool mixed = (C > dl) && (C < (0.999999 - dl)) && ((aR - aC) * (aC - aL) >= 0.0);
                if (mixed) {
                    mixedFlag(b, i, j, k) = 1;
                    double nvc = p.dim == 2 ? youngs(u, d, m, i, j, k) : 1.0;
                    double beta = nvc * beta0 + 0.01;
                    double coshB = mth.cosh(beta), tanhB = mth.tanh(beta);
                    double B = mth.exp(theta * beta * ((2.0 * C) - 1.0));
                    double Acoef = ((B / coshB) - 1.0) / tanhB;
                    A(tr, i, j, k, ALPHA, m) = mn + (mx / 2.0) * (1.0 + theta * ((tanhB + Acoef) / (1.0 + Acoef * tanhB)));
                    A(tl, i, j, k, ALPHA, m) = mn + (mx / 2.0) * (1.0 + theta * Acoef);
                    double rL = A(u, il, jl, kl, ALPHARHO, m), rR = A(u, ir, jr, kr, ALPHARHO, m);
                    double mn2 = dmin(rL, rR);
                    double mx2 = dmax(rL, rR) - mn2;
                    double th2 = (double)sgn(rR - rL);
                    A(tr, i, j, k, ALPHARHO, m) = mn2 + (mx2 / 2.0) * (1.0 + th2 * ((tanhB + Acoef) / (1.0 + Acoef * tanhB)));
                    A(tl, i, j, k, ALPHARHO, m) = mn2 + (mx2 / 2.0) * (1.0 + th2 * Acoef);
                }
            }
        });
        auto tbv = [&](Fab& pl, Fab& pr, int i, int j, int k, int m) {
            int il = i, jl = j, kl = k, ir = i, jr = j, kr = k;
            shift(d, -1, il, jl, kl);
            shift(d, 1, ir, jr, kr);
            double URl = A(ur, il, jl, kl, ALPHA, m), ULr = A(ul, ir, jr, kr, ALPHA, m);
            double TRl = A(tr, il, jl, kl, ALPHA, m), TLr = A(tl, ir, jr, kr, ALPHA, m);
            double PL = A(pl, i, j, k, ALPHA, m), PR = A(pr, i, j, k, ALPHA, m);
            return dmin(dmin(std::fabs(URl - PL) + std::fabs(PR - ULr), std::fabs(TRl - PL) + std::fabs(PR - TLr)),
                        dmin(std::fabs(URl - PL) + std::fabs(PR - TLr), std::fabs(TRl - PL) + std::fabs(PR - ULr)));
        };
        for_valid(b, [&](int i, int j, int k) {
            for (int m = 0; m < p.nmat; ++m) {
                double tm = tbv(ul, ur, i, j, k, m);
                double tt = tbv(tl, tr, i, j, k, m);
                tbvFlag(b, i, j, k, m) = (mixedFlag(b, i, j, k) && tt < tm) ? 1 : 0;
            }
        });
        for_valid(b, [&](int i, int j, int k) {
            for (int m = 0; m < p.nmat; ++m)
                if (tbvFlag(b, i, j, k, m)) {
                    A(ul, i, j, k, ALPHA, m) = A(tl, i, j, k, ALPHA, m);
                    A(ur, i, j, k, ALPHA, m) = A(tr, i, j, k, ALPHA, m);
                }
        });
    }

    // ------------------------------------------------------------ fluxes
    // flux.cpp:5-47 (physical flux of a cell view)
    struct CellV {
        Solver* s;
        Fab* f;
        int i, j, k;
        double& operator()(Var v, int m = 0, int r = 0, int c = 0) { return s->A(*f, i, j, k, v, m, r, c); }
        double& operator()(const Spec& sp) { return s->A(*f, i, j, k, sp); }
    };
    double vdots(CellV& c, int d) {
        double r = 0.0;
        for (int a = 0; a < 3; ++a) r += c(VELOCITY, 0, a) * c(SIGMA, 0, a, d);
        return r;
    }
    double pflux(const Spec& n, CellV& c, int d) {
        switch (n.var) {
            case ALPHA: return c(ALPHA, n.mat) * c(VELOCITY, 0, d);
            case ALPHARHO: return c(ALPHARHO, n.mat) * c(VELOCITY, 0, d);
            case ALPHARHOEPSILON: return c(ALPHARHOEPSILON, n.mat) * c(VELOCITY, 0, d);
            case ALPHARHOLAMBDA: return c(ALPHARHOLAMBDA, n.mat) * c(VELOCITY, 0, d);
            case RHOU: return c(RHOU, 0, n.row) * c(VELOCITY, 0, d) - c(SIGMA, 0, n.row, d);
            case TOTAL_E: return c(TOTAL_E) * c(VELOCITY, 0, d) - vdots(c, d);
            case V_TENSOR:
                return c(V_TENSOR, 0, n.row, n.col) * c(VELOCITY, 0, d) - c(V_TENSOR, 0, d, n.col) * c(VELOCITY, 0, n.row);
            default: return 0.0;
        }
    }
    // hllc.cpp:5-8
    static double sstar(CellV& l, CellV& r, double SL, double SR, int d) {
        return (-r(SIGMA, 0, d, d) + l(SIGMA, 0, d, d) + l(RHOU, 0, d) * (SL - l(VELOCITY, 0, d)) - r(RHOU, 0, d) * (SR - r(VELOCITY, 0, d))) /
               (l(RHO) * (SL - l(VELOCITY, 0, d)) - r(RHO) * (SR - r(VELOCITY, 0, d)));
    }
    // hllc.cpp:12-73; cell.cpp:114-120 semantics (writes through the view)
    void star_state(CellV& u, CellV& us, double SK, double Ss, int d) {
        const double tol = 1E-20;
        double mult = (std::fabs((Ss - u(VELOCITY, 0, d)) / Ss) < tol ? 1.0 : (SK - u(VELOCITY, 0, d)) / (SK - Ss));
        us(RHO) = 0.0;
        for (int m = 0; m < p.nmat; ++m) {
            us(ALPHA, m) = u(ALPHA, m);
            us(ALPHARHO, m) = mult * u(ALPHARHO, m);
            us(RHO_K, m) = us(ALPHARHO, m) / us(ALPHA, m);
            if (p.mat[m].mixture) us(ALPHARHOLAMBDA, m) = mult * u(ALPHARHOLAMBDA, m);
            if (p.mat[m].plastic) us(ALPHARHOEPSILON, m) = mult * u(ALPHARHOEPSILON, m);
            us(RHO) += us(ALPHARHO, m);
        }
        for (int r = 0; r < 3; ++r) {
            if (r == d) us(RHOU, 0, r) = mult * u(RHO) * Ss;
            else us(RHOU, 0, r) = mult * u(RHOU, 0, r);
            us(VELOCITY, 0, r) = us(RHOU, 0, r) / us(RHO);
            if (p.SOLID)
                for (int c = 0; c < 3; ++c) {
                    if (r == d) us(V_TENSOR, 0, r, c) = u(V_TENSOR, 0, r, c);
                    else us(V_TENSOR, 0, r, c) = mult * u(V_TENSOR, 0, r, c);
                }
        }
        us(TOTAL_E) = mult * u(RHO) * (u(TOTAL_E) / u(RHO) + (Ss - u(VELOCITY, 0, d)) * (Ss - (u(SIGMA, 0, d, d)) / (u(RHO) * (SK - u(VELOCITY, 0, d)))));
    }
    // cell.cpp:114-120: copies every cell variable (accesspattern CELL list)
    void copy_cell(CellV& dst, CellV& src) {
        for (int n = 0; n < L.ncomp; ++n) {
            bool cellvar = !(n >= L.start[DEVH] && L.start[DEVH] >= 0 && n < L.start[DEVH] + 9) &&
                           !(L.start[HJ2] >= 0 && n == L.start[HJ2]) &&
                           !(L.start[RHO_MIX] >= 0 && n >= L.start[RHO_MIX] && n < L.start[RHO_MIX] + 2);
            if (cellvar) (*dst.f)(dst.i, dst.j, dst.k, n) = (*src.f)(src.i, src.j, src.k, n);
        }
    }
    // hllc.cpp:75-83
    void sigma_star(CellV& uks, double Ss, double SLT, double SRT, CellV& ul, CellV& ur, CellV& uls, CellV& urs, int d) {
        for (int i = 0; i < 3; ++i)
            uks(SIGMA, 0, i, d) = (uls(RHO) * (Ss - SLT) * urs(RHO) * (Ss - SRT) * (ul(VELOCITY, 0, i) - ur(VELOCITY, 0, i)) +
                                   uls(RHO) * (Ss - SLT) * ur(SIGMA, 0, i, d) - urs(RHO) * (Ss - SRT) * ul(SIGMA, 0, i, d)) /
                                  (uls(RHO) * (Ss - SLT) - urs(RHO) * (Ss - SRT));
    }
    // hllc.cpp:87-129
    void starstar(CellV& ul, CellV& ur, CellV& uls, CellV& urs, CellV& uss, double Ss, double SLT, double SRT, int d,
                  CellV& uk, CellV& uks, double SKT) {
        copy_cell(uss, uks);
        sigma_star(uks, Ss, SLT, SRT, ul, ur, uls, urs, d);
        for (int r = 0; r < 3; ++r) {
            if (r == d) continue;
            uss(RHOU, 0, r) += (uks(SIGMA, 0, r, d) - uk(SIGMA, 0, r, d)) / (Ss - SKT);
            uss(VELOCITY, 0, r) = uss(RHOU, 0, r) / uss(RHO);
        }
        for (int r = 0; r < 3; ++r) {
            if (r == d) continue;
            uss(TOTAL_E) += (uss(VELOCITY, 0, r) * uks(SIGMA, 0, r, d) - uk(VELOCITY, 0, r) * uk(SIGMA, 0, r, d)) / (Ss - SKT);
            for (int c = 0; c < 3; ++c)
                uss(V_TENSOR, 0, r, c) += (uss(V_TENSOR, 0, d, c) * (uss(VELOCITY, 0, r) - uk(VELOCITY, 0, r))) / (Ss - SKT);
        }
    }
    std::atomic<int> fatal{0};  // 1: NaN wave speeds (hllc.cpp:169-200, 367-373)

    // hllc.cpp:384-487: one face (i,j,k) of box b; the box loop is calc_fluxes
    void hllc_face(int b, int d, int i, int j, int k) {
        Fab& fl = flux[d][b];
        int ext[3] = {0, 0, 0};
        ext[d] = 1;
        CellV l{this, &UR.f[b], i - ext[0], j - ext[1], k - ext[2]};
        CellV r{this, &UL.f[b], i, j, k};
        CellV st{this, &ULStar.f[b], i, j, k};
        double SR = dmax(std::fabs(l(VELOCITY, 0, d)) + l(SOUNDSPEED), std::fabs(r(VELOCITY, 0, d)) + r(SOUNDSPEED));
        double SL = -SR;
        double Ss = sstar(l, r, SL, SR, d);
        A(fl, i, j, k, USTAR) = Ss;
        if (SL >= 0.0) {
            star_state(l, st, SL, Ss, d);
            for (const Spec& n : L.cons) A(fl, i, j, k, n) = n.var == ALPHA ? pflux(n, st, d) : pflux(n, l, d);
        } else if (Ss >= 0.0) {
            star_state(l, st, SL, Ss, d);
            for (const Spec& n : L.cons)
                A(fl, i, j, k, n) = n.var == ALPHA ? pflux(n, st, d) : pflux(n, l, d) + SL * (st(n) - l(n));
        } else if (SR >= 0.0) {
            star_state(r, st, SR, Ss, d);
            for (const Spec& n : L.cons)
                A(fl, i, j, k, n) = n.var == ALPHA ? pflux(n, st, d) : pflux(n, r, d) + SR * (st(n) - r(n));
        } else {
            star_state(r, st, SR, Ss, d);
            for (const Spec& n : L.cons) A(fl, i, j, k, n) = n.var == ALPHA ? pflux(n, st, d) : pflux(n, r, d);
        }
    }
    void calc_fluxes(int b, int d) {
        const IBox& bx = boxes[b];
        int ext[3] = {0, 0, 0};
        ext[d] = 1;
        for (int k = bx.lo[2]; k <= bx.hi[2] + ext[2]; ++k)
            for (int j = bx.lo[1]; j <= bx.hi[1] + ext[1]; ++j)
                for (int i = bx.lo[0]; i <= bx.hi[0] + ext[0]; ++i) hllc_face(b, d, i, j, k);
    }

    // hllc.cpp:133-380: one face (i,j,k) of box b; returns false on NaN wave
    // speeds (the reference prints and exits).  SLT/SRT persist across the
    // faces of a box as in the reference (function-scope variables).
    bool wave5_face(int b, int d, int i, int j, int k, double& SLT, double& SRT) {
        Fab& fl = flux[d][b];
        int ext[3] = {0, 0, 0};
        ext[d] = 1;
        {
            {
                {
                    CellV l{this, &UR.f[b], i - ext[0], j - ext[1], k - ext[2]};
                    CellV r{this, &UL.f[b], i, j, k};
                    CellV ls{this, &ULStar.f[b], i, j, k};
                    CellV rs{this, &URStar.f[b], i, j, k};
                    CellV ss{this, &UStarStar.f[b], i, j, k};
                    double SR = dmax(std::fabs(l(VELOCITY, 0, d)) + l(SOUNDSPEED), std::fabs(r(VELOCITY, 0, d)) + r(SOUNDSPEED));
                    double SL = -SR;
                    double Ss = sstar(l, r, SL, SR, d);
                    if (std::isnan(SL) || std::isnan(SR) || std::isnan(Ss)) { fatal |= 1; return false; }
                    A(fl, i, j, k, USTAR) = Ss;
                    auto vstar_from = [&](CellV& c) {
                        for (int rr = 0; rr < 3; ++rr)
                            for (int cc = 0; cc < 3; ++cc) A(fl, i, j, k, VSTAR, 0, rr, cc) = c(V_TENSOR, 0, rr, cc);
                    };
                    if (SL >= 0.0) {
                        star_state(l, ls, SL, Ss, d);
                        for (const Spec& n : L.cons) A(fl, i, j, k, n) = n.var == ALPHA ? pflux(n, ls, d) : pflux(n, l, d);
                        vstar_from(l);
                    } else if (Ss >= 0.0) {
                        star_state(l, ls, SL, Ss, d);
                        SLT = Ss - transverse(ULStar.f[b], i, j, k);
                        if (SLT >= 0.0) {
                            for (const Spec& n : L.cons)
                                A(fl, i, j, k, n) = n.var == ALPHA ? pflux(n, ls, d) : pflux(n, l, d) + SL * (ls(n) - l(n));
                            vstar_from(ls);
                        } else {
                            star_state(r, rs, SR, Ss, d);
                            SRT = Ss + transverse(URStar.f[b], i, j, k);
                            starstar(l, r, ls, rs, ss, Ss, SLT, SRT, d, l, ls, SLT);
                            for (const Spec& n : L.cons)
                                A(fl, i, j, k, n) = n.var == ALPHA ? pflux(n, ls, d)
                                                                   : pflux(n, l, d) + SL * (ls(n) - l(n)) + SLT * (ss(n) - ls(n));
                            vstar_from(ss);
                        }
                    } else if (SR >= 0.0) {
                        star_state(r, rs, SR, Ss, d);
                        SRT = Ss + transverse(URStar.f[b], i, j, k);
                        if (SRT >= 0.0) {
                            star_state(l, ls, SL, Ss, d);
                            SLT = Ss - transverse(ULStar.f[b], i, j, k);
                            starstar(l, r, ls, rs, ss, Ss, SLT, SRT, d, r, rs, SRT);
                            for (const Spec& n : L.cons)
                                A(fl, i, j, k, n) = n.var == ALPHA ? pflux(n, rs, d)
                                                                   : pflux(n, r, d) + SR * (rs(n) - r(n)) + SRT * (ss(n) - rs(n));
                            vstar_from(ss);
                        } else {
                            for (const Spec& n : L.cons)
                                A(fl, i, j, k, n) = n.var == ALPHA ? pflux(n, rs, d) : pflux(n, r, d) + SR * (rs(n) - r(n));
                            vstar_from(rs);
                        }
                    } else {
                        star_state(r, rs, SR, Ss, d);
                        for (const Spec& n : L.cons) A(fl, i, j, k, n) = n.var == ALPHA ? pflux(n, rs, d) : pflux(n, r, d);
                        vstar_from(r);
                    }
                    if (std::isnan(SL) || std::isnan(SR) || std::isnan(Ss) || std::isnan(SLT) || std::isnan(SRT)) {
                        fatal |= 1;
                        return false;
                    }
                }
            }
        }
        return true;
    }
    void calc_5wave(int b, int d) {
        const IBox& bx = boxes[b];
        int ext[3] = {0, 0, 0};
        ext[d] = 1;
        double SLT = 0.0, SRT = 0.0;
        for (int k = bx.lo[2]; k <= bx.hi[2] + ext[2]; ++k)
            for (int j = bx.lo[1]; j <= bx.hi[1] + ext[1]; ++j)
                for (int i = bx.lo[0]; i <= bx.hi[0] + ext[0]; ++i)
                    if (!wave5_face(b, d, i, j, k, SLT, SRT)) return;
    }

    // hllc.cpp:493-523
    void update(int b, int d, double dtv) {
        Fab& fl = flux[d][b];
        Fab& u = U_in->f[b];
        Fab& u1 = U_out->f[b];
        const int ns = L.start[USTAR];
        for (const Spec& n : L.cons) {
            int c = L.comp(n);
            for_valid(b, [&](int i, int j, int k) {
                int ir = i, jr = j, kr = k;
                shift(d, 1, ir, jr, kr);
                double f0 = fl(i, j, k, c), f1 = fl(ir, jr, kr, c);
                double s0 = fl(i, j, k, ns), s1 = fl(ir, jr, kr, ns);
                if (n.var == ALPHA) {
                    u1(i, j, k, c) += (dtv / dx[d]) * (f0 - f1 - u(i, j, k, c) * (s0 - s1));
                } else if (n.var == V_TENSOR) {
                    int vs0 = L.comp(VSTAR, 0, d, n.col);
                    u1(i, j, k, c) += (dtv / dx[d]) * (f0 - f1 - (2.0 / 3.0) * u(i, j, k, c) * (s0 - s1) +
                                                       A(u, i, j, k, VELOCITY, 0, n.row) * (fl(i, j, k, vs0) - fl(ir, jr, kr, vs0)));
                } else {
                    u1(i, j, k, c) += (dtv / dx[d]) * (f0 - f1);
                }
            });
        }
    }

    // ------------------------------------------------------------ stage / advance
    CellArray* U_in = nullptr;
    CellArray* U_out = nullptr;
    std::atomic<int> alpha_fatal{0};

    // hllc.cpp:579-695
    void hllc_advance(CellArray& Uin, CellArray& Uout, double dtv) {
        U_in = &Uin;
        U_out = &Uout;
        copy(Uout, Uin);
        for (int d = 0; d < p.dim; ++d) {
            copy(UL, Uin);
            copy(UR, Uin);
#pragma omp parallel for schedule(dynamic)
            for (size_t b = 0; b < boxes.size(); ++b) {
                if (p.MUSCL) muscl((int)b, d);
                if (p.THINC) {
                    thinc((int)b, d);
                    for_valid((int)b, [&](int i, int j, int k) {  // D2
                        alpha_fatal |= cleanup_alpha_cell(UL.f[b], i, j, k);
                        alpha_fatal |= cleanup_alpha_cell(UR.f[b], i, j, k);
                    });
                }
                for_valid((int)b, [&](int i, int j, int k) {
                    p2c_cell(UL.f[b], i, j, k);
                    p2c_cell(UR.f[b], i, j, k);
                });
                for_valid((int)b, [&](int i, int j, int k) {
                    sound_cell(UL.f[b], i, j, k);
                    sound_cell(UR.f[b], i, j, k);
                });
            }
            fill(UL);
            fill(UR);
#pragma omp parallel for schedule(dynamic)
            for (size_t b = 0; b < boxes.size(); ++b) {
                if (p.SOLID) calc_5wave((int)b, d);
                else calc_fluxes((int)b, d);
                update((int)b, d, dtv);
            }
        }
#pragma omp parallel for schedule(dynamic)
        for (size_t b = 0; b < boxes.size(); ++b)
            for_valid((int)b, [&](int i, int j, int k) { c2p_cell(Uout.f[b], i, j, k); });
        fill(Uout);
    }

    // hllc.cpp:699-707 (U and U2 are scaled in place, cellarray.cpp:87-99)
    void advance(double dtv) {
        hllc_advance(U, U1, dtv);
        hllc_advance(U1, U2, dtv);
#pragma omp parallel for schedule(static)
        for (size_t b = 0; b < boxes.size(); ++b) {
            std::vector<double>& a = U.f[b].d;
            std::vector<double>& c = U2.f[b].d;
            std::vector<double>& o = U1.f[b].d;
            for (size_t q = 0; q < a.size(); ++q) {
                a[q] *= (1.0 / 2.0);
                c[q] *= (1.0 / 2.0);
                a[q] += c[q];
                o[q] = a[q];
            }
        }
    }

    // ------------------------------------------------------------ post-step
    // plastic.cpp:76-79, 83-244
    void plastic_update() {
        const double s3 = std::sqrt(3.0), s23 = std::sqrt(2.0 / 3.0);
#pragma omp parallel for schedule(dynamic)
        for (size_t b = 0; b < boxes.size(); ++b) {
            Fab& f = U1.f[b];
            for_valid((int)b, [&](int i, int j, int k) {
                double ts = 0.0;
                for (int m = 0; m < p.nmat; ++m)
                    if (p.mat[m].solid) ts += A(f, i, j, k, ALPHA, m);
                if (ts <= 0.5) {
                    for (int m = 0; m < p.nmat; ++m) {
                        A(f, i, j, k, EPSILON, m) = 0.0;
                        A(f, i, j, k, ALPHARHOEPSILON, m) = 0.0;
                    }
                    return;
                }
                henckyJ2(f, i, j, k);
                double JT = 0.0, JB = 0.0, norm = 0.0;
                for (int m = 0; m < p.nmat; ++m) {
                    double J, Jn;
                    if (p.mat[m].solid) {
                        J = std::sqrt(A(f, i, j, k, HJ2));
                        if (shearG(f, i, j, k, m) * J * s3 > p.mat[m].yield) Jn = p.mat[m].yield / (s3 * shearG(f, i, j, k, m));
                        else Jn = J;
                        A(f, i, j, k, EPSILON, m) += s23 * (J - Jn);
                        A(f, i, j, k, ALPHARHOEPSILON, m) = A(f, i, j, k, EPSILON, m) * A(f, i, j, k, ALPHARHO, m);
                    } else {
                        J = 0.0;
                        Jn = 0.0;
                        A(f, i, j, k, EPSILON, m) = 0.0;
                        A(f, i, j, k, ALPHARHOEPSILON, m) = 0.0;
                    }
                    JB += invGr(f, i, j, k, m) * J;
                    JT += invGr(f, i, j, k, m) * Jn;
                    norm += invGr(f, i, j, k, m);
                }
                JB *= 1.0 / norm;
                JT *= 1.0 / norm;
                if (JB == 0.0 && JT == 0.0) { JB = 1.0; JT = 1.0; }
                double t1[9], t2[9], t2i[9], Vn[9];
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        double& dh = A(f, i, j, k, DEVH, 0, r, c);
                        if (dh != 0.0 && JB == 0.0 && JT != 0.0) { alpha_fatal |= 4; }
                        else dh *= JT / JB;
                        t1[r * 3 + c] = delta<double>(r, c) + 0.5 * dh;
                        t2[r * 3 + c] = delta<double>(r, c) - 0.5 * dh;
                    }
                inv3(t2, t2i);
                mm(t2i, t1, Vn);
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) A(f, i, j, k, V_TENSOR, 0, r, c) = Vn[r * 3 + c];
            });
            for_valid((int)b, [&](int i, int j, int k) { c2p_cell(f, i, j, k); });
        }
    }
    // boxaccesscellarray.cpp:476-528
    void cleanup_v() {
#pragma omp parallel for schedule(dynamic)
        for (size_t b = 0; b < boxes.size(); ++b) {
            Fab& f = U1.f[b];
            for_valid((int)b, [&](int i, int j, int k) {
                double ta = 0.0, tsd = 0.0;
                for (int m = 0; m < p.nmat; ++m) {
                    ta += A(f, i, j, k, ALPHA, m);
                    if (p.mat[m].solid) tsd += A(f, i, j, k, ALPHA, m);
                }
                double V[9], G[9], S[9];
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) V[r * 3 + c] = A(f, i, j, k, V_TENSOR, 0, r, c);
                mmT(V, V, G);
                sqrtm_spd(G, S);
                double norm = mth.pow_m13(det3(S));
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c)
                        A(f, i, j, k, V_TENSOR, 0, r, c) = ((S[r * 3 + c] * norm * tsd) + delta<double>(r, c) * (ta - tsd)) / ta;
            });
        }
    }
    void cleanup_alpha_all(CellArray& ca) {
#pragma omp parallel for schedule(dynamic)
        for (size_t b = 0; b < boxes.size(); ++b)
            for_valid((int)b, [&](int i, int j, int k) { alpha_fatal |= cleanup_alpha_cell(ca.f[b], i, j, k); });
    }
    // cellarray.cpp:101-117, boxaccesscellarray.cpp:601-645 (over every box)
    bool contains_nan(CellArray& ca) {
        for (size_t b = 0; b < boxes.size(); ++b) {
            bool bad = false;
            for_valid((int)b, [&](int i, int j, int k) {
                for (const Spec& s : L.cons) bad |= std::isnan(A(ca.f[b], i, j, k, s));
                for (const Spec& s : L.prim) bad |= std::isnan(A(ca.f[b], i, j, k, s));
            });
            if (bad) return true;
        }
        return false;
    }

    // timestep.cpp:7-70
    double get_timestep(CellArray& ca) {
        double dtm = std::numeric_limits<double>::max();
#pragma omp parallel for schedule(dynamic)
        for (size_t b = 0; b < boxes.size(); ++b) {
            Fab& f = ca.f[b];
            Fab& t = tsfab[b];
            for_valid((int)b, [&](int i, int j, int k) { sound_cell(f, i, j, k); });
            for_valid((int)b, [&](int i, int j, int k) {
                for (int n = 0; n < p.dim; ++n) t(i, j, k, n) = dx[n] / (A(f, i, j, k, SOUNDSPEED) + std::fabs(A(f, i, j, k, VELOCITY, 0, n)));
            });
        }
        for (int n = 0; n < p.dim; ++n) {
            double mn = std::numeric_limits<double>::max();
            for (size_t b = 0; b < boxes.size(); ++b)
                for_valid((int)b, [&](int i, int j, int k) { mn = std::min(mn, tsfab[b](i, j, k, n)); });
            dtm = (mn < dtm ? mn : dtm);
        }
        return p.CFL * dtm;
    }

    // flux.cpp:51-77 geometricFlux: the radial flux of conservative variable
    // n in cell c (the stress is the cell's stored sigma)
    double gflux(const Spec& n, CellV& c) {
        switch (n.var) {
            case ALPHA: return 0.0;
            case ALPHARHO: case ALPHARHOLAMBDA: case ALPHARHOEPSILON:
                return c(n) * c(VELOCITY, 0, 0);
            case RHOU:
                if (n.row == 0) return c(n) * c(VELOCITY, 0, 0) - c(SIGMA, 0, 0, 0) + c(SIGMA, 0, 2, 2);
                if (n.row == 1) return c(n) * c(VELOCITY, 0, 0) - c(SIGMA, 0, 0, 1);
                return -2.0 * c(SIGMA, 0, 0, 2);
            case TOTAL_E: return c(n) * c(VELOCITY, 0, 0) - vdots(c, 0);
            case V_TENSOR:
                if (n.row == 0 || n.row == 1) return (1.0 / 3.0) * c(n) * c(VELOCITY, 0, 0);
                return -(2.0 / 3.0) * c(n) * c(VELOCITY, 0, 0);
            default: throw std::runtime_error("oracle: bad radial flux variable");
        }
    }
    // radial.cpp:5-54 geometricSourceTerm on U1: per conservative variable,
    // U += dt * (multiplier / ((i + 1/2) dx)) * geometricFlux, multiplier -2
    // in one dimension (spherical) and -1 otherwise (cylindrical); then
    // conservativeToPrimitive box by box.
    void geometric_source(double dtv) {
        const double mult = (p.dim == 1 ? -2.0 : -1.0);
#pragma omp parallel for schedule(dynamic)
        for (size_t b = 0; b < boxes.size(); ++b) {
            Fab& f = U1.f[b];
            for (const Spec& n : L.cons)
                for_valid((int)b, [&](int i, int j, int k) {
                    CellV c{this, &f, i, j, k};
                    const double inverse_r = mult / (((double)i + 0.5) * dx[0]);
                    A(f, i, j, k, n) += dtv * inverse_r * gflux(n, c);
                });
            for_valid((int)b, [&](int i, int j, int k) { c2p_cell(f, i, j, k); });
        }
    }

    // reactive.cpp:3-24 pressureBased_UpdateMassFraction: one explicit step of
    // the pressure-dependent burn rate from state u into o (cell, material m)
    void burn(Fab& fu, Fab& fo, int i, int j, int k, int m, double dtv) {
        const double sigma = 20.0, nu = 0.5, n = 1.2;
        if (A(fu, i, j, k, ALPHA, m) < 0.5) {
            A(fo, i, j, k, ALPHARHOLAMBDA, m) = A(fu, i, j, k, ALPHARHOLAMBDA, m);
            return;
        }
        A(fo, i, j, k, ALPHARHOLAMBDA, m) =
            A(fu, i, j, k, ALPHARHOLAMBDA, m) - (dtv) * (sigma * mth.pow(A(fu, i, j, k, ALPHARHO, m), 1.0 - nu) *
                                                         mth.pow(A(fu, i, j, k, ALPHARHOLAMBDA, m), nu) *
                                                         mth.pow((A(fu, i, j, k, P) / 1E6), n));
        if (A(fo, i, j, k, ALPHARHOLAMBDA, m) < 0.0) A(fo, i, j, k, ALPHARHOLAMBDA, m) = 0.0;
    }
    // reactive.cpp:60-103 reactiveUpdate: Heun's RK2 of the burn on the mixture
    // materials of U1, then conservativeToPrimitive box by box
    void reactive_update(double dtv) {
        if (p.nmix <= 0) return;
        copy(U, U1);
#pragma omp parallel for schedule(dynamic)
        for (size_t b = 0; b < boxes.size(); ++b) {
            Fab &fu = U.f[b], &f1 = U1.f[b], &f2 = U2.f[b];
            for (int m = 0; m < p.nmat; ++m)
                if (p.mat[m].mixture)
                    for_valid((int)b, [&](int i, int j, int k) {
                        burn(fu, f1, i, j, k, m, dtv);
                        burn(f1, f2, i, j, k, m, dtv);
                        A(f1, i, j, k, ALPHARHOLAMBDA, m) =
                            ((A(fu, i, j, k, ALPHARHOLAMBDA, m) * (1.0 / 2.0)) + (A(f2, i, j, k, ALPHARHOLAMBDA, m) * (1.0 / 2.0)));
                    });
            for_valid((int)b, [&](int i, int j, int k) { c2p_cell(f1, i, j, k); });
        }
    }

    // main.cpp:124-171: one time step starting from U1 (ghosts filled here)
    // with a caller-chosen dt (main.cpp:128-143 chooses it).
    void step(double dtv) {
        fill(U1);
        copy(U, U1);
        advance(dtv);
        if (p.REACTIVE) reactive_update(dtv);
        if (p.PLASTIC) plastic_update();
        if (p.RADIAL) geometric_source(dtv);
        if (p.SOLID) cleanup_v();
        cleanup_alpha_all(U1);
    }
    // main.cpp:121-191 time loop; returns steps taken
    int run(double finalT, int max_steps, double* t_out) {
        int n = 0;
        double t = 0.0;
        for (; t < finalT && n < max_steps; t += dt, n++) {
            fill(U1);
            dt = get_timestep(U1);
            if (t + dt > finalT) dt = finalT - t;
            if (n < 5) dt *= 0.2;
            step(dt);
            if (contains_nan(U1) || fatal || alpha_fatal || mix_fatal) break;  // main.cpp:183-187
        }
        if (t_out) *t_out = t;
        return n;
    }
};

}  // namespace tjo
