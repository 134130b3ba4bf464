// Test infrastructure only (the oracle): C entry points over the CPU
// restatement in tjo_oracle.hpp, loaded by tests/ and bench.py through ctypes.
#include <cstring>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <vector>

#include "tjo_oracle.hpp"

extern "C" {

struct tjo_params {
    int dim;
    int n[3];
    double L[3];
    int max_grid_size;
    int nghost;
    int SOLID, THINC, PLASTIC, MUSCL;
    double THINCbeta;
    double CFL;
    int nmat;
    int bc_lo[3];
    int bc_hi[3];
    int math_mode;
    int eos[4];
    int solid[4];
    int plastic[4];
    double gamma[4], pref[4], eref[4], CV[4], rho0[4], c0[4], s[4], G0[4], yield[4];
    int RADIAL;
    int REACTIVE;
    int mixture[4];
    double gamma2[4], pref2[4], eref2[4], CV2[4];
    double K0[4], ealpha[4], ebeta[4], e1[4], e2[4], e3[4], e4[4], e5[4];
};

static tjo::Params to_params(const tjo_params* q) {
    tjo::Params p;
    p.dim = q->dim;
    for (int a = 0; a < 3; ++a) {
        p.n[a] = q->n[a];
        p.L[a] = q->L[a];
        p.bc_lo[a] = q->bc_lo[a];
        p.bc_hi[a] = q->bc_hi[a];
    }
    p.max_grid_size = q->max_grid_size;
    p.nghost = q->nghost;
    p.SOLID = q->SOLID;
    p.THINC = q->THINC;
    p.PLASTIC = q->PLASTIC;
    p.RADIAL = q->RADIAL;
    p.REACTIVE = q->REACTIVE;
    p.MUSCL = q->MUSCL;
    p.THINCbeta = q->THINCbeta;
    p.CFL = q->CFL;
    p.nmat = q->nmat;
    p.math_mode = q->math_mode;
    for (int m = 0; m < 4; ++m) {
        tjo::Material& M = p.mat[m];
        M.eos = q->eos[m];
        M.solid = q->solid[m];
        M.plastic = q->plastic[m];
        M.adiabaticIndex = q->gamma[m];
        M.pref = q->pref[m];
        M.eref = q->eref[m];
        M.CV = q->CV[m];
        M.rho0 = q->rho0[m];
        M.c0 = q->c0[m];
        M.s = q->s[m];
        M.G0 = q->G0[m];
        M.yield = q->yield[m];
        M.mixture = q->mixture[m];
        M.adiabaticIndex2 = q->gamma2[m];
        M.pref2 = q->pref2[m];
        M.eref2 = q->eref2[m];
        M.CV2 = q->CV2[m];
        M.K0 = q->K0[m];
        M.EOSalpha = q->ealpha[m];
        M.EOSbeta = q->ebeta[m];
        M.e1 = q->e1[m];
        M.e2 = q->e2[m];
        M.e3 = q->e3[m];
        M.e4 = q->e4[m];
        M.e5 = q->e5[m];
        if (m < p.nmat && M.mixture) p.nmix = 1;
    }
    return p;
}

void* tjo_create(const tjo_params* q) { return new tjo::Solver(to_params(q)); }
void tjo_destroy(void* h) { delete static_cast<tjo::Solver*>(h); }
int tjo_ncomp(void* h) { return static_cast<tjo::Solver*>(h)->L.ncomp; }
int tjo_comp_start(void* h, int var) { return static_cast<tjo::Solver*>(h)->L.start[var]; }
int tjo_nboxes(void* h) { return (int)static_cast<tjo::Solver*>(h)->boxes.size(); }
void tjo_box(void* h, int b, int* lo, int* hi) {
    auto* s = static_cast<tjo::Solver*>(h);
    for (int a = 0; a < 3; ++a) { lo[a] = s->boxes[b].lo[a]; hi[a] = s->boxes[b].hi[a]; }
}

static tjo::CellArray* pick(tjo::Solver* s, int which) {
    switch (which) {
        case 0: return &s->U1;
        case 1: return &s->U;
        case 2: return &s->U2;
        case 3: return &s->UL;
        case 4: return &s->UR;
        case 5: return &s->ULStar;
        case 6: return &s->URStar;
        case 7: return &s->UStarStar;
        default: return &s->MUSCLgrad;
    }
}

// Valid cells of the whole domain, [comp][k][j][i].
void tjo_get_valid(void* h, int which, double* out) {
    auto* s = static_cast<tjo::Solver*>(h);
    tjo::CellArray* ca = pick(s, which);
    const int nx = s->p.n[0], ny = s->p.dim > 1 ? s->p.n[1] : 1, nz = s->p.dim > 2 ? s->p.n[2] : 1;
    for (size_t b = 0; b < s->boxes.size(); ++b)
        s->for_valid((int)b, [&](int i, int j, int k) {
            for (int n = 0; n < s->L.ncomp; ++n)
                out[(((size_t)n * nz + k) * ny + j) * nx + i] = ca->f[b](i, j, k, n);
        });
}
void tjo_set_valid(void* h, int which, const double* in) {
    auto* s = static_cast<tjo::Solver*>(h);
    tjo::CellArray* ca = pick(s, which);
    const int nx = s->p.n[0], ny = s->p.dim > 1 ? s->p.n[1] : 1, nz = s->p.dim > 2 ? s->p.n[2] : 1;
    for (size_t b = 0; b < s->boxes.size(); ++b)
        s->for_valid((int)b, [&](int i, int j, int k) {
            for (int n = 0; n < s->L.ncomp; ++n)
                ca->f[b](i, j, k, n) = in[(((size_t)n * nz + k) * ny + j) * nx + i];
        });
}
// One grown FAB of one array (for scratch-history checks): [comp][k][j][i] over the grown box.
void tjo_get_fab(void* h, int which, int b, double* out) {
    auto* s = static_cast<tjo::Solver*>(h);
    const tjo::Fab& f = pick(s, which)->f[b];
    std::memcpy(out, f.d.data(), f.d.size() * sizeof(double));
}

// setInitialConditions (initialise.cpp:104, 111-123): optional thermal-to-total
// pressure shift, then primitiveToConservative, then the ghost fill.
void tjo_init_from_prims(void* h, int shift_pressure) {
    auto* s = static_cast<tjo::Solver*>(h);
    for (size_t b = 0; b < s->boxes.size(); ++b) {
        tjo::Fab& f = s->U1.f[b];
        s->for_valid((int)b, [&](int i, int j, int k) {
            if (shift_pressure)
                s->A(f, i, j, k, tjo::P) += s->effNTP(f, i, j, k) / s->effInvG(f, i, j, k);
            s->p2c_cell(f, i, j, k);
        });
    }
    s->fill(s->U1);
}

double tjo_timestep(void* h) {
    auto* s = static_cast<tjo::Solver*>(h);
    s->fill(s->U1);
    return s->get_timestep(s->U1);
}
void tjo_step(void* h, double dt) {
    auto* s = static_cast<tjo::Solver*>(h);
    s->dt = dt;
    s->step(dt);
}
int tjo_run(void* h, double finalT, int max_steps, double* t_out) {
    return static_cast<tjo::Solver*>(h)->run(finalT, max_steps, t_out);
}
// bit 0: NaN in U1 (contains_nan); bit 1: NaN wave speed (5-wave); bits 2-4: cleanup fatal codes
int tjo_status(void* h) {
    auto* s = static_cast<tjo::Solver*>(h);
    int st = 0;
    if (s->contains_nan(s->U1)) st |= 1;
    if (s->fatal) st |= 2;
    st |= (s->alpha_fatal & 7) << 2;
    st |= (s->mix_fatal & 3) << 5;
    return st;
}
// Stand-alone cleanUpAlpha on U1 (boxaccesscellarray.cpp:530-599)
int tjo_cleanup_alpha(void* h) {
    auto* s = static_cast<tjo::Solver*>(h);
    s->alpha_fatal = 0;
    s->cleanup_alpha_all(s->U1);
    return s->alpha_fatal;
}
void tjo_cleanup_v(void* h) { static_cast<tjo::Solver*>(h)->cleanup_v(); }
void tjo_plastic(void* h) { static_cast<tjo::Solver*>(h)->plastic_update(); }
void tjo_geometric_source(void* h, double dt) { static_cast<tjo::Solver*>(h)->geometric_source(dt); }
void tjo_reactive(void* h, double dt) { static_cast<tjo::Solver*>(h)->reactive_update(dt); }
void tjo_c2p(void* h) {
    auto* s = static_cast<tjo::Solver*>(h);
    for (size_t b = 0; b < s->boxes.size(); ++b)
        s->for_valid((int)b, [&](int i, int j, int k) { s->c2p_cell(s->U1.f[b], i, j, k); });
}
void tjo_p2c(void* h) {
    auto* s = static_cast<tjo::Solver*>(h);
    for (size_t b = 0; b < s->boxes.size(); ++b)
        s->for_valid((int)b, [&](int i, int j, int k) { s->p2c_cell(s->U1.f[b], i, j, k); });
}
void tjo_sound(void* h) {
    auto* s = static_cast<tjo::Solver*>(h);
    for (size_t b = 0; b < s->boxes.size(); ++b)
        s->for_valid((int)b, [&](int i, int j, int k) { s->sound_cell(s->U1.f[b], i, j, k); });
}

// Riemann flux between two full cell vectors given as primitives: both are
// converted by primitiveToConservative and getSoundSpeed, placed as UR(i-1)
// and UL(i) of box 0, and the face flux of face 1 along d is returned
// (all Ncomp components of the nodal flux FAB, USTAR and VSTAR included).
void tjo_face_flux(void* h, int d, const double* left, const double* right, double* out) {
    auto* s = static_cast<tjo::Solver*>(h);
    tjo::Fab& ur = s->UR.f[0];
    tjo::Fab& ul = s->UL.f[0];
    int i0[3] = {0, 0, 0}, i1[3] = {0, 0, 0};
    i1[d] = 1;
    for (int n = 0; n < s->L.ncomp; ++n) {
        ur(i0[0], i0[1], i0[2], n) = left[n];
        ul(i1[0], i1[1], i1[2], n) = right[n];
    }
    s->p2c_cell(ur, i0[0], i0[1], i0[2]);
    s->sound_cell(ur, i0[0], i0[1], i0[2]);
    s->p2c_cell(ul, i1[0], i1[1], i1[2]);
    s->sound_cell(ul, i1[0], i1[1], i1[2]);
    if (s->p.SOLID) {
        double slt = 0.0, srt = 0.0;
        s->wave5_face(0, d, i1[0], i1[1], i1[2], slt, srt);
    } else {
        s->hllc_face(0, d, i1[0], i1[1], i1[2]);
    }
    for (int n = 0; n < s->L.ncomp; ++n) out[n] = s->flux[d][0](i1[0], i1[1], i1[2], n);
}

// Reconstruction pass alone (hllc.cpp:589-638 without the flux): UL, UR of
// direction d from U1, MUSCL then THINC/BVD and cleanUpAlpha when enabled,
// then primitiveToConservative and getSoundSpeed.  Read back with which = 3/4.
void tjo_reconstruct(void* h, int d) {
    auto* s = static_cast<tjo::Solver*>(h);
    s->fill(s->U1);
    s->U_in = &s->U1;
    tjo::Solver::copy(s->UL, s->U1);
    tjo::Solver::copy(s->UR, s->U1);
    for (size_t b = 0; b < s->boxes.size(); ++b) {
        if (s->p.MUSCL) s->muscl((int)b, d);
        if (s->p.THINC) {
            s->thinc((int)b, d);
            s->for_valid((int)b, [&](int i, int j, int k) {
                s->alpha_fatal |= s->cleanup_alpha_cell(s->UL.f[b], i, j, k);
                s->alpha_fatal |= s->cleanup_alpha_cell(s->UR.f[b], i, j, k);
            });
        }
        s->for_valid((int)b, [&](int i, int j, int k) {
            s->p2c_cell(s->UL.f[b], i, j, k);
            s->p2c_cell(s->UR.f[b], i, j, k);
            s->sound_cell(s->UL.f[b], i, j, k);
            s->sound_cell(s->UR.f[b], i, j, k);
        });
    }
}

double tjo_vanleer(double r) { return tjo::Solver::vanleer(r); }
double tjo_math(int kind, int mode, double x) {
    tjo::Math m;
    m.mode = mode;
    switch (kind) {
        case 0: return m.exp(x);
        case 1: return m.cosh(x);
        case 2: return m.tanh(x);
        case 4: return m.log(x);
        case 5: return m.pow(x, 0.5);
        case 6: return m.pow(x, 1.2);
        default: return m.pow_m13(x);
    }
}
void tjo_sqrtm(const double* A, double* S) { tjo::sqrtm_spd(A, S); }
void tjo_inv3(const double* A, double* B) { tjo::inv3(A, B); }
double tjo_det3(const double* A) { return tjo::det3(A); }
int tjo_chop(int n, int chunk, int* starts) {
    std::vector<int> c = tjo::chop_dim(n, chunk);
    for (size_t i = 0; i < c.size(); ++i) starts[i] = c[i];
    return (int)c.size();
}

// transverseWaveSpeed (boxaccesscellarray.cpp:287-324) of valid cell (i,j,k) of U1
// (it recomputes that cell's HJ2 first, as the reference does).
double tjo_transverse(void* h, int i, int j, int k) {
    auto* s = static_cast<tjo::Solver*>(h);
    for (size_t b = 0; b < s->boxes.size(); ++b) {
        const tjo::IBox& bx = s->boxes[b];
        if (i >= bx.lo[0] && i <= bx.hi[0] && j >= bx.lo[1] && j <= bx.hi[1] && k >= bx.lo[2] && k <= bx.hi[2])
            return s->transverse(s->U1.f[b], i, j, k);
    }
    return -1.0;
}
// OpenMP threads of the box loops (the CPU baseline times several counts in one process).
void tjo_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
}  // extern "C"
