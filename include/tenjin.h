/*
 * tenjin.h -- C ABI of the MI355X (gfx950) per-cell explicit update.
 *
 * The reference has no plugin or FFI layer: its hot path is a set of C++ free
 * functions and methods called from main_main().  Each entry point below is
 * the device-side replacement of one of them; the C++ host layer
 * (tenjin_amd/csrc/host/) keeps the reference's signatures and forwards here.
 *
 * Conventions
 *   - Return 0 on success, a negative TJ_E* code on API misuse; a positive
 *     status word (TJ_ST_*) is reported by tj_status() for the conditions on
 *     which the reference prints and calls exit(1) or stops its time loop.
 *   - All state buffers are caller-owned device memory described by tj_mf;
 *     the library owns only its scratch (flux planes, THINC history planes,
 *     the device-resident time-step words).
 *   - One host thread per context; every call is asynchronous on the
 *     context's HIP stream unless it returns a host value.
 */
#ifndef TENJIN_H
#define TENJIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TJ_MAX_MATERIALS 4

/* EOS kinds: 0 = MieGruneisenEOS of equationofstate.cpp:5-115 (stiffened gas,
 * constant-Gamma Mie-Gruneisen with pref/eref); 1 = Us-up Mie-Gruneisen solid
 * (no reference class; solid-EOS interface of equationofstate.cpp:502-596);
 * 2 = RomenskiiSolidEOS (equationofstate.cpp:403-495: rho0, K0, eos_alpha,
 * eos_beta, G0); 3 = WilkinsSolidEOS (equationofstate.cpp:502-596: rho0,
 * e1 .. e5, G0).  Kinds 1-3 are solids (solid = 1) on SOLID levels. */
typedef struct tj_material {
    int eos;
    int solid;    /* MaterialDescriptor::phase == solid (structdefinitions.h:90-98) */
    int plastic;  /* MaterialDescriptor::plastic */
    int mixture;  /* MaterialDescriptor::mixture: MixtureEOS (equationofstate.cpp:116-376) of the
                     MieGruneisen reactant (gamma, pref, eref, CV) and products (*_mix below);
                     at most one mixture material, on fluid levels */
    double gamma; /* adiabaticIndex; GruneisenGamma = gamma - 1 */
    double pref, eref, CV;
    double rho0, c0, s, G0;
    double yield_stress; /* PlasticEOS::yieldStress[m] (plastic.h:21) */
    double gamma_mix, pref_mix, eref_mix, CV_mix; /* the products' EOS (adiabaticIndexmix, prefmix,
                                                     erefmix, CVmix of initialise.cpp:189-192) */
    double K0, eos_alpha, eos_beta;               /* RomenskiiSolidEOS (K0, EOSalpha, EOSbeta) */
    double e1, e2, e3, e4, e5;                    /* WilkinsSolidEOS */
} tj_material;

/* Level description: what ParameterStruct (structdefinitions.h:105-144),
 * the BoxArray of main.cpp:34-36 and the BCRec table of boundary.cpp:3-42
 * carry for the hot path. */
typedef struct tj_level_desc {
    int dim;                 /* AMREX_SPACEDIM: 1, 2 or 3 */
    int n[3];                /* parameters.n_cells */
    double L[3];             /* parameters.dimL */
    int max_grid_size;       /* BoxArray::maxSize chunk (main.cpp:36) */
    int SOLID, THINC, PLASTIC, MUSCL;
    double THINCbeta, CFL;
    int nmat;
    int bc_lo[3], bc_hi[3];  /* 0 transmissive (foextrap), 1 reflective */
    tj_material mat[TJ_MAX_MATERIALS];
    int RADIAL;              /* parameters.RADIAL: geometricSourceTerm each step (main.cpp:159-162);
                                x is the radius, spherical in 1D, cylindrical otherwise.  On solid
                                levels it needs PLASTIC (the source reads the stress of the state
                                that plasticUpdate's conservativeToPrimitive leaves). */
    int REACTIVE;            /* parameters.REACTIVE: reactiveUpdate each step (main.cpp:149-152) on
                                the mixture material */
} tj_level_desc;

typedef struct tj_fabset tj_fabset;

/* Non-owning view of a CellArray on the device, in one of two forms:
 *   - fabs == NULL: one array over the level, component-major, i fastest.
 *     data points at component 0 of valid cell (0,0,0); strides in doubles,
 *     so a single ghost-grown FAB is described by its own strides;
 *   - fabs != NULL: a MultiFab of one FAB per box (tj_fabset_create); data
 *     and the strides are not read.
 * Every entry point taking tj_mf accepts either form; the views of one call
 * must all be of the same form. */
typedef struct tj_mf {
    double* data;
    long long jstride, kstride, nstride;
    int ncomp;
    int pad_;
    const tj_fabset* fabs;
} tj_mf;

/* One FAB of a MultiFab (an AMReX FArrayBox as its Array4 sees it: base
 * pointer, first stored index and strides): device memory holding the cells glo ..
 * glo + extent - 1 of the level's index space, component-major.  lo/hi is the
 * box it owns (a box of the level's BoxArray::maxSize decomposition, inclusive);
 * the stored cells beyond it are its ghost cells (glo = lo - nGrow for an
 * AMReX FAB). */
typedef struct tj_fab {
    double* data;                         /* component 0 of stored cell glo */
    int lo[3], hi[3];                     /* the valid box (level indices) */
    int glo[3];                           /* first stored index per dimension */
    int pad_;
    long long jstride, kstride, nstride;  /* in doubles */
} tj_fab;

typedef struct tj_ctx tj_ctx;

enum {
    TJ_OK = 0,
    TJ_EINVAL = -1,
    TJ_EUNSUPPORTED = -2,
    TJ_EHIP = -3,
    TJ_ENOLEVEL = -4,
    TJ_ESCRATCH = -5  /* tj_run_static: the level's buffers exceed TJ_STATIC_SCRATCH_BYTES */
};

/* status bits (tj_status) */
enum {
    TJ_ST_NAN = 1,          /* CellArray::contains_nan (cellarray.cpp:101-117) -> time loop stops */
    TJ_ST_WAVESPEED = 2,    /* NaN wave speeds in the 5-wave solver (hllc.cpp:169-200, 367-373) */
    TJ_ST_ALPHA_NANS = 4,   /* "too many Nans" in cleanUpAlpha (boxaccesscellarray.cpp:559-561) */
    TJ_ST_ALPHA_SUM = 8,    /* total alpha <= 0 in cleanUpAlpha (boxaccesscellarray.cpp:583-588) */
    TJ_ST_PLASTIC_DIV = 16, /* "Divide by zero error" in boxPlasticUpdate (plastic.cpp:188-193) */
    TJ_ST_MIX_BRACKET = 32, /* "Error in root Bisection" in MixtureEOS::rootFind (equationofstate.cpp:313-321) */
    TJ_ST_MIX_ITER = 64,    /* rootFind's bisection reached its bound of 2000 passes (the reference
                               loops without one, equationofstate.cpp:324-353) */
    TJ_ST_TANGENT_NAN = 128, /* a NaN derivative part in the tangent-linear form (no reference
                               counterpart; the run goes on) */
    /* The device time loop stops at the step that raises any of these (t and n
     * are not advanced), where the reference stops its loop (NaN, main.cpp:
     * 183-187) or calls exit(1) (the others; bit 64: its loop never ends). */
    TJ_ST_STOP = 0x7f
};

/* precision forms */
enum { TJ_PREC_FP64 = 0, TJ_PREC_MIXED = 1 };

/* post-step operation bits (tj_post_step) */
enum {
    TJ_OP_PLASTIC = 1,      /* PlasticEOS::plasticUpdate (plastic.cpp:227-244) */
    TJ_OP_CLEAN_V = 2,      /* CellArray::cleanUpV (boxaccesscellarray.cpp:476-528) */
    TJ_OP_CLEAN_ALPHA = 4,  /* CellArray::cleanUpAlpha (boxaccesscellarray.cpp:530-599) */
    TJ_OP_DT = 8,           /* TimeStep::getTimeStep (timestep.cpp:36-70) */
    TJ_OP_NANCHECK = 16,    /* CellArray::contains_nan (cellarray.cpp:101-117) */
    TJ_OP_RADIAL = 32,      /* geometricSourceTerm (radial.cpp:37-54) with the dt of tj_set_dt;
                               applied after PLASTIC and before CLEAN_V */
    TJ_OP_REACTIVE = 64     /* reactiveUpdate (reactive.cpp:60-103) with the dt of tj_set_dt;
                               applied first, before PLASTIC */
};

/* Context on HIP device `device` with its own stream (default priority). */
int tj_create(tj_ctx** ctx, int device);
void tj_destroy(tj_ctx* ctx);
const char* tj_last_error(const tj_ctx* ctx);
/* The context's stream as a void* (hipStream_t), for callers that order their own work on it. */
void* tj_stream(tj_ctx* ctx);

/* Binds a level: box table (BoxArray::maxSize cuts), component map of
 * AccessPattern (accesspattern.cpp:44-79), EOS/material table, BC table,
 * dx = L/n.  Allocates the scratch (flux planes, THINC history planes).
 * Replaces the set-up of main.cpp:29-95. */
int tj_bind_level(tj_ctx* ctx, const tj_level_desc* desc);
/* Tuning: rows per marching segment of the y/z flux passes (seglen > 0
 * fixes it; 0, the default, picks it per pass from the level's shape and the
 * kernel's occupancy so that the segments fill the device's wave slots
 * evenly), and the form of an RK stage of a fluid level:
 *   fuse = 2 (default): on two-dimensional levels without a mixture, the
 *            stage kernels -- the x pass with the x part of the update, then
 *            the y pass over LDS tiles with the rest of the update -- and no
 *            flux planes; other levels as fuse = 0;
 *   fuse = 1: flux planes, the update inside the last directional pass;
 *   fuse = 0: flux planes and an update kernel of its own.
 * Every choice gives identical results; a value < 0 leaves that setting
 * unchanged. */
int tj_set_tuning(tj_ctx* ctx, int seglen, int fuse);
/* Number of components of the bound layout and the first component of a
 * reference Variable (structdefinitions.h:10-32 numbering); -1 if absent. */
int tj_ncomp(const tj_ctx* ctx);
int tj_comp_start(const tj_ctx* ctx, int variable);

/* TimeStep::getTimeStep (timestep.cpp:36-70): getSoundSpeed on the valid
 * cells of U (writes the SOUNDSPEED component), then CFL * min over cells
 * and directions of dx_d / (a + |u_d|).  The result stays in the context's
 * device-resident dt word; if dt_out is non-NULL the call synchronises and
 * returns it on the host. */
int tj_compute_dt(tj_ctx* ctx, const tj_mf* U, double* dt_out);
/* main.cpp:135-143: dt <- min(dt, finalT - t); dt *= 0.2 for step n < 5;
 * operates on the device-resident words (t, n, dt). */
int tj_driver_dt(tj_ctx* ctx, double final_time);
/* Sets the device-resident dt (and resets t, n when reset_clock != 0). */
int tj_set_dt(tj_ctx* ctx, double dt, int reset_clock);
/* Reads the device-resident (t, dt, n). Synchronises. */
int tj_get_clock(tj_ctx* ctx, double* t, double* dt, int* n);

/* One Euler stage of HLLCadvance (hllc.cpp:579-695) with the device dt:
 *   stage 1: Uout = c2p(Uin + dt L(Uin))
 *   stage 2: Uout = 1/2 Ubase + 1/2 c2p(Uin + dt L(Uin))   (advance, hllc.cpp:699-707)
 * Uout may alias Ubase in stage 2.  Only the persistent components are
 * written (alpha, alphaRho, rho_k, rho, rhoU, u, E, p; V and epsilon,
 * alphaRhoEpsilon in solid runs); sigma, devH, HJ2, a are derived. */
int tj_rk_stage(tj_ctx* ctx, int stage, const tj_mf* Uin, const tj_mf* Ubase, tj_mf* Uout, int prec);

/* Post-step chain of main.cpp:154-171 (+ the next getTimeStep), fused per cell:
 * ops is a TJ_OP_* mask, applied in the reference's order. */
int tj_post_step(tj_ctx* ctx, tj_mf* U, unsigned ops);

/* One full time step of main.cpp:124-171 on S (S^n -> S^{n+1}) using T as the
 * stage-1 buffer: tj_rk_stage(1), tj_rk_stage(2) with the post-step chain and
 * the next dt fused into stage 2.  dt is the device-resident dt. */
int tj_step(tj_ctx* ctx, tj_mf* S, tj_mf* T, int prec);

/* Start of the time loop: t = 0, n = 0, status cleared, and the CFL minimum of
 * the initial state S primed for the first tj_driver_dt. */
int tj_begin_run(tj_ctx* ctx, const tj_mf* S, double final_time);
/* nsteps iterations of the loop body of main.cpp:121-191 (tj_driver_dt then
 * tj_step), queued without synchronisation; iterations after the loop test
 * fails (t >= final_time) or after a NaN stop are no-ops on the device. */
int tj_advance_steps(tj_ctx* ctx, tj_mf* S, tj_mf* T, int nsteps, int prec);

/* hipGraph form of tj_advance_steps (the loop body of main.cpp:121-191 as one
 * graph): tj_graph_capture records nsteps loop bodies on S and T into a
 * graph owned by the context (replacing any earlier one); each
 * tj_graph_launch replays them on the context's stream.  The graph holds S
 * and T by address: they must stay allocated, in place, until the next
 * capture or tj_destroy.  With THINC the replay must start from the history
 * phase it was captured at (true after whole steps; TJ_EINVAL otherwise). */
int tj_graph_capture(tj_ctx* ctx, tj_mf* S, tj_mf* T, int nsteps, int prec);
int tj_graph_launch(tj_ctx* ctx);

/* The time loop of main.cpp:121-191 on the device words: dt, clamp, step,
 * NaN stop.  Returns steps taken in *n_out and final time in *t_out. */
int tj_run(tj_ctx* ctx, tj_mf* S, tj_mf* T, double final_time, int max_steps, int prec,
           int check_every, double* t_out, int* n_out);

/* ---- The time loop without device allocation --------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): tj_run on the level described by desc, with no context.
 * Every buffer tj_bind_level would allocate (the clock words, box-edge flags,
 * flux planes, THINC history planes) lives in one device array of
 * TJ_STATIC_SCRATCH_BYTES that the library holds on each GPU it runs on until
 * the process ends; the call allocates nothing, on the device or the host.
 *   - S: the initial state (after tj_init_from_prims), advanced in place; T: a
 *     distinct stage buffer; both caller-owned level arrays of the bound
 *     layout (a tj_fabset view needs a context: TJ_EINVAL).
 *   - Work runs on `stream` (a hipStream_t, NULL for the null stream) on the
 *     current device, with the fixed marching segment length 16; results are
 *     bit-identical to tj_bind_level + tj_run (every segment length gives the
 *     same values).
 *   - Returns TJ_ESCRATCH, before touching the device, when the level needs
 *     more than TJ_STATIC_SCRATCH_BYTES (about 330 bytes per cell for a 2D
 *     fluid level with THINC in the mixed form: a 512 x 256 level fits).
 *   - Every call on a GPU shares that GPU's array: calls that overlap on one
 *     GPU (other host threads, other streams) race.  The history planes are
 *     cleared at the start of each call. */
#define TJ_STATIC_SCRATCH_BYTES (64ll << 20)
int tj_run_static(const tj_level_desc* desc, tj_mf* S, tj_mf* T, double final_time, int max_steps, int prec,
                  int check_every, void* stream, double* t_out, int* n_out);

/* primitiveToConservative over the valid cells (boxaccesscellarray.cpp:124-174),
 * with the thermal-to-total pressure shift of initialise.cpp:104 when
 * shift_pressure != 0 (setInitialConditions, initialise.cpp:111-123). */
int tj_init_from_prims(tj_ctx* ctx, tj_mf* U, int shift_pressure);
/* conservativeToPrimitive over the valid cells (boxaccesscellarray.cpp:67-122). */
int tj_c2p(tj_ctx* ctx, tj_mf* U);
/* CellArray RK algebra (cellarray.cpp:80-99, the operators behind
 * U1 = ((U*(1.0/2.0))+(U2*(1.0/2.0))) of hllc.cpp:705): Z = (X*a) + (Y*b) over
 * every component of the views and every valid cell, products rounded before
 * the sum; Y may be NULL (Z = X*a).  Z may alias X or Y. */
int tj_lincomb(tj_ctx* ctx, const tj_mf* X, double a, const tj_mf* Y, double b, tj_mf* Z);
/* Derived components sigma (and devH, HJ2 in solid runs) from the state. */
int tj_derive(tj_ctx* ctx, tj_mf* U);

/* Per-kernel timing with HIP events recorded on the context's stream around
 * every launch the library makes (no reference counterpart; a measurement
 * hook).  tj_prof_report writes "name count total_ms\n" lines; synchronises. */
int tj_prof_enable(tj_ctx* ctx, int on);
int tj_prof_report(tj_ctx* ctx, char* buf, int len);
/* Restrict the event timing to launches of one name (as tj_prof_report names
 * them); NULL or "" times every launch again.  Each timed launch adds two
 * event records to the stream (a few microseconds), so a measurement of
 * throughput times only the kernel it needs (no reference counterpart). */
int tj_prof_only(tj_ctx* ctx, const char* name);
/* Of the launches that tj_prof_enable / tj_prof_only would time, time the
 * first and then every `every`-th one (1: all of them, the default), so a
 * long run carries fewer event records; the count in tj_prof_report is the
 * number timed.  TJ_EINVAL for every < 1 (no reference counterpart). */
int tj_prof_sample(tj_ctx* ctx, int every);

/* Status word (TJ_ST_*), OR-ed since the last call with clear != 0. Synchronises. */
int tj_status(tj_ctx* ctx, uint32_t* status, int clear);
/* Zeroes the THINC history planes.  The reference's ULStar/URStar arrays hold
 * allocation contents before their first write; zero is used here. */
int tj_reset_history(tj_ctx* ctx);
int tj_synchronize(tj_ctx* ctx);

/* A MultiFab as one FAB per box: the library reads and writes each cell in
 * the FAB that owns it, at that FAB's strides (no copy into a level array;
 * the FAB buffers stay the caller's).  Replaces the per-box Array4 access of
 * BoxAccessCellArray(MFIter&, const Box&, CellArray&) and
 * BoxAccessCellArray(const Box&, FArrayBox&, CellArray&) (cellarray.h:46-72,
 * boxaccesscellarray.cpp:3-23) for the whole level at once.
 *   - the boxes must be exactly the bound level's BoxArray::maxSize
 *     decomposition, in any order (TJ_EINVAL otherwise);
 *   - ghost cells are neither read nor written: a stencil that leaves a box
 *     reads the neighbouring box's valid cells, and the domain boundary is
 *     applied on the fly, as for the level array (FillBoundary and
 *     setBoundaryConditions are part of the kernels);
 *   - ncomp: components in every FAB (at least tj_ncomp).
 * The set holds device tables of the FAB records (about 40 B per box); it
 * belongs to ctx's device and must be destroyed before ctx. */
int tj_fabset_create(tj_ctx* ctx, int nfab, const tj_fab* fabs, int ncomp, tj_fabset** out);
/* The bound level's boxes (BoxArray::maxSize of main.cpp:34-36, the cut the
 * kernels' box-edge semantics follow): *nbox boxes, and with lohi != NULL six
 * ints per box (lo[3], hi[3]), x fastest.  For allocating FABs that match. */
int tj_level_boxes(const tj_ctx* ctx, int* nbox, int* lohi);
/* Cell-wise operations on one box [lo, hi] of the level (inclusive level
 * indices), the rest of the level untouched: the per-box methods of
 * BoxAccessCellArray (cellarray.h:75-101: primitiveToConservative,
 * conservativeToPrimitive, getSoundSpeed, stressTensor + getHenckyJ2 +
 * getDeviatoricHenckyStrain, cleanUpV, cleanUpAlpha, contains_nan),
 * TimeStep::boxTimeStepFunction (timestep.cpp:7-34) and
 * PlasticEOS::boxPlasticUpdate (plastic.cpp:83-225).  ops: TJ_BOX_* and the
 * cell-wise TJ_OP_* bits (PLASTIC, CLEAN_V, CLEAN_ALPHA, NANCHECK, REACTIVE,
 * DT), applied in the order P2C, C2P, REACTIVE, PLASTIC, CLEAN_V, CLEAN_ALPHA,
 * NANCHECK, DERIVE, DT.  With TJ_OP_DT, *dt_out = CFL x the box's minimum of
 * dx/(|u| + a) (the sound speed written on the way); the level's clock
 * (t, dt, n, CFL minimum) is left as it was, status bits accumulate as for
 * any call.  U may be a level array or a tj_fabset view; with a set the box
 * must lie in one FAB.  TJ_OP_RADIAL needs the cell's radius on the level:
 * TJ_EUNSUPPORTED.  Synchronises. */
enum {
    TJ_BOX_P2C = 1 << 8,     /* primitiveToConservative (no thermal-pressure shift) */
    TJ_BOX_C2P = 1 << 9,     /* conservativeToPrimitive */
    TJ_BOX_DERIVE = 1 << 10  /* stress, HJ2 and devH (solid levels; sigma = -p I for fluids) */
};
int tj_box_op(tj_ctx* ctx, tj_mf* U, const int lo[3], const int hi[3], unsigned ops, double* dt_out);
/* A tj_mf naming the set (fabs = set, ncomp = the set's). */
int tj_fabset_view(const tj_fabset* set, tj_mf* out);
void tj_fabset_destroy(tj_fabset* set);

/* ---- Batched ensemble -------------------------------------------------
 * Many independent fluid levels (1D and 2D, uneven sizes, each with its own
 * materials, boundary conditions and time step) advanced together: each
 * problem runs the loop body of main.cpp:121-191 with its own device clock,
 * and every pass of a step is one launch over all problems of a dimension
 * group.  A problem's results are bit-identical to running it alone with
 * tj_run.  The reference has no batched driver; this serves ensembles of
 * small problems that would each leave most of the GPU idle.
 * Restrictions: fluid levels (SOLID = 0), two materials, one THINC setting
 * for the whole batch, at most TJ_BATCH_MAX_PROBLEMS problems (TJ_EINVAL
 * beyond).  S[p] and T[p] are caller-owned device views (state
 * and stage buffer) of problem p, held by address until tj_batch_destroy.
 * On an error from tj_batch_create, *batch may hold a batch for
 * tj_batch_last_error; destroy it. */
typedef struct tj_batch tj_batch;
#define TJ_BATCH_MAX_PROBLEMS 65535
int tj_batch_create(tj_batch** batch, int device, int nprob, const tj_level_desc* descs, const tj_mf* S,
                    const tj_mf* T);
void tj_batch_destroy(tj_batch* batch);
const char* tj_batch_last_error(const tj_batch* batch);
/* Zeroes every problem's THINC history planes (as tj_reset_history). */
int tj_batch_reset_history(tj_batch* batch);
/* Rows per marching segment of problem p's y pass in the flux-plane form
 * (tuning; default 16). */
int tj_batch_set_seglen(tj_batch* batch, int prob, int seglen);
/* The y-tile shape of problem prob's stage kernels (2D problems): 0 = 16
 * columns x 32 rows per lane segment (the default), 1 = 8 x 32, 2 = 16 x 16.
 * Every shape gives the same results; the autotuner picks one per size
 * bucket.  Drops a captured graph. */
int tj_batch_set_tile(tj_batch* batch, int prob, int tile);
/* 2D problems: the stage kernels (on = 1, the default: each RK stage as an
 * x and a y launch over the tiles, no face fluxes in HBM) or the flux-plane
 * passes (on = 0).  The same results either way.  Drops a captured graph. */
int tj_batch_set_stage2d(tj_batch* batch, int on);
/* tj_init_from_prims for every problem. */
int tj_batch_init_from_prims(tj_batch* batch, int shift_pressure);
/* tj_begin_run for every problem (one final time for all). */
int tj_batch_begin_run(tj_batch* batch, double final_time);
/* nsteps loop bodies for every problem (problems past their final time or
 * stopped by a NaN are no-ops on the device). */
int tj_batch_advance(tj_batch* batch, int nsteps, int prec);
/* hipGraph form of tj_batch_advance (as tj_graph_capture / tj_graph_launch). */
int tj_batch_graph_capture(tj_batch* batch, int nsteps, int prec);
int tj_batch_graph_launch(tj_batch* batch);
/* Per-problem t, n and status words (arrays of nprob; any may be NULL). Synchronises. */
int tj_batch_clocks(tj_batch* batch, double* t, int* n, uint32_t* status);
int tj_batch_synchronize(tj_batch* batch);

/* ---- Tangent-linear (differentiated) form -----------------------------
 * Forward-mode derivatives of the time loop (main.cpp:121-191 with advance,
 * plasticUpdate, cleanUpV, cleanUpAlpha and getTimeStep, as tj_run) with
 * respect to per-material EOS and plasticity parameters: every value is the
 * FP64 form's, bit for bit, and each direction carries d(state)/d(parameter)
 * through init, flux, update, plastic, V clean-up and the CFL time step.  The
 * reference has no differentiated form; this serves gradient-based
 * calibration of material parameters against probe traces.
 * Levels: SOLID levels (5-wave path, no RADIAL), and fluid levels with MUSCL
 * (+ THINC/BVD, whose history planes carry derivative parts) and HLLC (no
 * mixture, no RADIAL).  Directions run in groups of
 * `chunk` (1, 2, 3 or 6; 0 = all in one group up to 6): each group re-runs
 * the trajectory from the initial primitives.
 * On an error from tj_tangent_create, *tg is NULL. */
enum {
    TJ_TP_GAMMA0 = 0, /* Gruneisen coefficient adiabaticIndex - 1 (Gamma0 of a Us-up solid) */
    TJ_TP_S = 1,      /* Us-up Hugoniot slope s (Us-up solids) */
    TJ_TP_YIELD = 2,  /* yield stress (plastic.cpp:132-135; plastic Us-up solids) */
    TJ_TP_PREF = 3    /* pref of a MieGruneisen (stiffened-gas) material (equationofstate.cpp:37-115) */
};
typedef struct tj_tangent tj_tangent;
int tj_tangent_create(tj_ctx* ctx, int ndir, const int* mat, const int* param, int chunk, tj_tangent** tg);
void tj_tangent_destroy(tj_tangent* tg);
/* From the primitive state P0 (as given to tj_init_from_prims, shift_pressure
 * alike): for each group, copy P0 to S, initialise, begin a run to final_time
 * and advance nsteps loop bodies (S state, T stage buffer; both caller-owned
 * views of the bound layout).  After every step the value and derivatives of
 * component probe_comp of cell (pi, pj, pk) go to row n of trace (host,
 * nsteps x (1 + ndir) doubles: value, then one column per direction; may be
 * NULL).  *steps_done: loop bodies run (fewer than nsteps after the final
 * time or a NaN stop).  S ends in the final state.  P0, S and T are three
 * distinct buffers. */
int tj_tangent_run(tj_tangent* tg, const tj_mf* P0, int shift_pressure, tj_mf* S, tj_mf* T, double final_time,
                   int nsteps, int probe_comp, int pi, int pj, int pk, double* trace, int* steps_done);
/* The final-state derivative of direction dir (of the last group run) into
 * the caller's device view D of the bound layout. */
int tj_tangent_state(tj_tangent* tg, int dir, tj_mf* D);

#ifdef __cplusplus
}
#endif

#endif /* TENJIN_H */
