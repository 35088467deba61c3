/*
 * awcm_capi.h -- C ABI of the MI355X (gfx950) AWCM compressible-Euler solver.
 *
 * The reference (doseongju/C_pp-library) is a header-only C++ library with no
 * FFI; its "API" is the class surface its drivers call.  Each entry point
 * below replaces one reference call (file:line cited on the entry point), or
 * a whole driver loop, with a bulk operation on device-resident state.  The
 * C++ header API in c_pp-library_amd/include/awcm/ (AWCM_1D, AWCM_2D, ...)
 * is built on these calls.
 *
 * Conventions
 *   - Plain C types only; no HIP or torch types cross this boundary.
 *   - Every call returns an int status: AWCM_OK (0) or a negative AWCM_E*.
 *   - State layout at the boundary is the reference's (FIELD_UNIFORM_1D.h:51-55,
 *     FIELD_UNIFORM_2D.h:42-58, ARRAY_2D.h:112): AoS, ghost width 3,
 *       1-D: q[(i+3)*3 + c],                i = -3 .. Nx+2
 *       2-D: q[((i+3) + (Nx+6)*(j+3))*4 + c]
 *     masks are uint8 with the same ghosted indexing, levels int32 without
 *     ghosts (lvl[i] / lvl[i + Nx*j]).
 *   - A handle owns one HIP stream; calls on one handle are ordered and not
 *     thread-safe; different handles are independent.
 *   - FP64 throughout; results are bit-identical to the reference's on the
 *     same inputs (same masks, levels and fields).
 */
#ifndef AWCM_CAPI_H
#define AWCM_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    AWCM_OK = 0,
    AWCM_EINVAL = -1,     /* bad argument (sizes, levels, null pointers)        */
    AWCM_EDEVICE = -2,    /* HIP runtime / kernel launch failure                */
    AWCM_ENOMEM = -3,     /* device allocation failed                           */
    AWCM_ENONFINITE = -4, /* state not finite / non-positive (sqrt of negative) */
    AWCM_ERANGE = -5,     /* a mask scatter left the ghosted domain             */
    AWCM_ESTATE = -6,     /* call not valid in the handle's current state       */
    AWCM_EDISABLED = -7   /* entry point switched off in this build             */
};

/* Ghost-fill kinds (the per-problem BOUNDARY_MANAGER classes). */
enum {
    AWCM_BC_TRANSMISSIVE = 0, /* zero-order extrapolation: PROBLEM_SHOCKTUBE.h:86-107,
                                 PROBLEM_ENTROPY_INTERACTING.h:81-102,
                                 PROBLEM_QUADRANTS.h:123-203                          */
    AWCM_BC_DOUBLE_MACH = 1   /* BOUNDARY_MANAGER_DMACH::SetBoundaryCondition(t),
                                 PROBLEM_DOUBLEMACH.h:106-245                          */
};

typedef struct awcm_solver awcm_solver;
typedef struct awcm_batch awcm_batch;

/* ---- solver handles ------------------------------------------------------ */

/* MRA_GRID_INT_1D::Initialize (MRA_GRID_UNIFORM_1D.h:70-98) + AWCM_1D(mra)
 * (AWCM_1D.h:21-41).  Nx = (nx_coarse-1)*2^(num_levels-1)+1.  dx is the
 * physical spacing the driver takes from its GRID_UNIFORM_1D.  Requires
 * 2 <= num_levels <= 14 and nx_coarse >= 6. */
int awcm_create_1d(int num_levels, int nx_coarse, double gamma, double dx, awcm_solver** out);

/* MRA_GRID_INT_2D::Initialize (MRA_GRID_UNIFORM_2D.h:87-134) + AWCM_2D(mra)
 * (AWCM_2D.h:24-48).  dx is GRID_UNIFORM_2D::dx (dy must equal it, as the
 * drivers assume); x0 is the x of column 0 (used by the double-Mach ghost
 * fill).  Requires 2 <= num_levels <= 14, nx_coarse, ny_coarse >= 5. */
int awcm_create_2d(int num_levels, int nx_coarse, int ny_coarse, double gamma, double dx, double x0,
                   int bc_kind, awcm_solver** out);

void awcm_destroy(awcm_solver* h);

/* Finest-grid sizes (ny = 1 for 1-D). */
int awcm_dims(const awcm_solver* h, int* nx, int* ny);

/* Host <-> device copies of the conserved field (reference AoS layout). */
int awcm_upload_q(awcm_solver* h, const double* q);
int awcm_download_q(awcm_solver* h, double* q);

/* Replace mask0 (the previous step's active set, ghosted uint8). */
int awcm_set_mask0(awcm_solver* h, const uint8_t* mask0);

/* Read the five masks (mask_temp, mask0..mask3; each ghosted) and the four
 * level fields (level0..level3) -- the public members AWCM_1D/2D expose
 * (AWCM_1D.h:12-13, AWCM_2D.h:15-16). */
int awcm_get_masks(awcm_solver* h, uint8_t* masks, int32_t* levels);

/* AWCM_1D::PrepareEuler (AWCM_1D.h:86-120) / AWCM_2D::PrepareEuler
 * (AWCM_2D.h:114-147): forward transform, threshold, mask refinement, inverse
 * transform and interpolation of the new points, on the device. */
int awcm_prepare_euler(awcm_solver* h, double tol);

/* WAVELET_2D::InterpolateAllGridEuler(Q, mra_grid, mask_from)
 * (Wavelet/WAVELET_2D.h:912-1044), called by Wavelet/main_2d_euler.cpp:334
 * after the time loop: every finest-grid point not in mask_from (ghosted
 * uint8, reference layout) gets its interpolating-wavelet prediction from the
 * coarser levels, all four components, on the device state.  On a 1-D handle
 * its counterpart, the one-mask WAVELET_1D::InterpolateEuler(q, mra,
 * mask_from) (WAVELET_1D.h:433-473), all three components.  Synchronous. */
int awcm_interpolate_all_grid(awcm_solver* h, const uint8_t* mask_from);

/* ---- the 1-D stage flux ---------------------------------------------------
 * The finest-level dF of the 1-D stages is the difference of two face fluxes.
 * AWCM_FLUX_WENOZ: WENOZ_EULER_SOLVER_1D::ComputeFlux (WENOZ.h:160-205), the
 * default.  AWCM_FLUX_WENO5: WENO_EULER_SOLVER_1D::ComputedF (WENO.h:129-184),
 * the WENO5-JS flux pair the shipped driver calls (main_1d_euler.cpp:109, 163,
 * 218).  awcm_set_flux applies to a 1-D handle (AWCM_ESTATE on a 2-D one,
 * whose drivers use WENO-Z), awcm_batch_set_flux to every instance of a batch.
 * The adjoint differentiates the WENO-Z pair: awcm_adjoint_record returns
 * AWCM_ESTATE on a WENO5-JS handle. */
enum { AWCM_FLUX_WENOZ = 0, AWCM_FLUX_WENO5 = 1 };
int awcm_set_flux(awcm_solver* h, int kind);

/* ---- the scalar path ------------------------------------------------------
 * AWCM_1D/2D::Prepare / PostProcess and WAVELET_1D/2D::Interpolate act on a
 * scalar field (FIELD_UNIFORM_1D/2D<double>, ghost width 3: Nx+6 doubles, or
 * (Nx+6) x (Ny+6) with i fastest).  On the device it is the first component
 * of the handle's field; the masks and levels are the handle's. */
int awcm_upload_scalar(awcm_solver* h, const double* q);
int awcm_download_scalar(awcm_solver* h, double* q);

/* AWCM_2D::Prepare (Wavelet/AWCM_2D.h:52-79) / AWCM_1D::Prepare
 * (Wavelet/AWCM_1D.h:45-71): forward transform of the scalar field with
 * threshold tol, banding, adjacent zone, reconstruction check, masks 1..3
 * with their levels, inverse transform.  mask_temp is left as it was. */
int awcm_prepare(awcm_solver* h, double tol);

/* AWCM_2D::PostProcess (AWCM_2D.h:81-84, AWCM_MANAGER_2D::RemoveResidue
 * AWCM_MANAGER_2D.h:257-266) / AWCM_1D::PostProcess (AWCM_1D.h:81-84,
 * AWCM_MANAGER_1D.h:146-157): zero the field at the points outside mask0 (the
 * interior in 2-D; every component of the handle's field). */
int awcm_post_process(awcm_solver* h);

/* WAVELET_2D::Interpolate(q, mra, mask_from, mask_to) (WAVELET_2D.h:360-473),
 * as main_2d_advect.cpp:219 calls it, and WAVELET_1D::Interpolate
 * (WAVELET_1D.h:198-275; its one-mask form is mask_to = every point): the
 * scalar field is predicted from the coarser levels at the points of
 * CheckReconstruction(mask_to minus mask_from) that are not in mask_from.
 * Masks ghosted uint8, reference layout.  Synchronous. */
int awcm_interpolate(awcm_solver* h, const uint8_t* mask_from, const uint8_t* mask_to);

/* WAVELET_2D::InterpolateEuler(q, mra, mask_from, mask_to)
 * (WAVELET_2D.h:784-907) / WAVELET_1D::InterpolateEuler(q, mra, from, to)
 * (WAVELET_1D.h:393-431): the same on momentum and energy of the Euler field.
 * Synchronous. */
int awcm_interpolate_euler(awcm_solver* h, const uint8_t* mask_from, const uint8_t* mask_to);

/* AWCM_*::PostProcessEuler (AWCM_1D.h:122-125, AWCM_2D.h:149-152). */
int awcm_post_process_euler(awcm_solver* h);

/* The scalar driver's time loop, Wavelet/main_2d_advect.cpp:46-199, on the
 * device (2-D handles): n_steps iterations of AWCM_2D::Prepare(q, tol), the
 * periodic ghost fill of PROBLEM_BOX_ADVECTION.h's BOUNDARY_MANAGER (sides
 * only), three SSP-RK3 stages on mask2, mask1, mask0 with velocity (1, 1) --
 * WENO5 right reconstructions on the finest level, AWCM_2D::Differentiate and
 * its one-sided forms at the point's spacing below it, h = dx * spacing with
 * the handle's dx -- the combine on mask0 and AWCM_2D::PostProcess.  The
 * field is the handle's scalar field (awcm_upload_scalar).  The three stage
 * fields are kept in the handle between calls (zero when first used), as the
 * driver keeps Q_2, Q_1, Q_0.  AWCM_ERANGE: a coarse stencil would leave the
 * ghosted grid (the reference reads outside its arrays there).  Synchronous. */
int awcm_run_scalar(awcm_solver* h, double tol, double dt, int n_steps);

/* ---- AWCM_MANAGER_1D / AWCM_MANAGER_2D mask algebra on caller arrays ------
 * The static members of AWCM_MANAGER_1D (Wavelet/AWCM_MANAGER_1D.h:14-157)
 * and AWCM_MANAGER_2D (Wavelet/AWCM_MANAGER_2D.h:14-280), on the GPU, for
 * masks, levels and fields the caller owns, on the handle's MRA grid (only its
 * geometry is used; the handle's own state is left unchanged).  Layouts are
 * the reference's: a mask is ghosted uint8 ((Nx+6) or (Nx+6)(Ny+6) bytes, in
 * and out), a level field int32 without ghosts (Nx or Nx*Ny entries, each in
 * 0 .. num_levels-1), a field AoS ghosted with ncomp components.  A mask
 * update that would leave the ghosted domain returns AWCM_ERANGE (the
 * reference writes out of bounds there).  Synchronous. */
/* AddAdjacentZone (AWCM_MANAGER_1D.h:14-43, AWCM_MANAGER_2D.h:14-61) */
int awcm_add_adjacent_zone(awcm_solver* h, uint8_t* mask);
/* CheckReconstruction (AWCM_MANAGER_1D.h:45-71, AWCM_MANAGER_2D.h:63-158) */
int awcm_check_reconstruction(awcm_solver* h, uint8_t* mask);
/* AddPointForDifferentiation (AWCM_MANAGER_1D.h:73-107, AWCM_MANAGER_2D.h:160-213) */
int awcm_add_point_for_differentiation(awcm_solver* h, uint8_t* mask, const int32_t* level);
/* ComputeLevel (AWCM_MANAGER_1D.h:109-133, AWCM_MANAGER_2D.h:215-241): level is
 * in/out (a point with no active neighbour keeps its value, as in the
 * reference) */
int awcm_compute_level(awcm_solver* h, int32_t* level, const uint8_t* mask);
/* RemoveResidue (AWCM_MANAGER_1D.h:135-157, both overloads; AWCM_MANAGER_2D.h:
 * 257-266) and RemoveResidueEuler (AWCM_MANAGER_2D.h:267-280): zero every
 * component outside mask -- 1-D at the finest-grid points, 2-D at the interior
 * points (the outer ring is left alone, as in the reference).  ncomp: 1
 * (scalar field) or the Euler field's 3 (1-D) / 4 (2-D). */
int awcm_remove_residue(awcm_solver* h, double* q, int ncomp, const uint8_t* mask);

/* Output-time schedule (arr_printing_times of the drivers,
 * main_1d_euler.cpp:43-48): dt is clipped so steps land on these times. */
int awcm_set_schedule(awcm_solver* h, const double* print_times, int n_print);

/* Time loop of Wavelet/main_1d_euler.cpp:61-260 (1-D, WENO-Z flux pair) or
 * main_2d_euler(dmach|quad).cpp:56-350 (2-D): up to max_steps iterations of
 * PrepareEuler, CFL dt (clipped to the schedule), three SSP-RK3 stages with
 * ghost fills, and PostProcessEuler, stopping once t >= t_final.
 * t and cnt (the print counter, 1 on the first call) are in/out; steps_done
 * and updates (sum over steps of |mask0|) may be NULL. */
int awcm_run(awcm_solver* h, double tol, double cfl, double t_final, int max_steps, double* t, int* cnt,
             int* steps_done, long long* updates);

/* ---- discrete adjoint (1-D) ------------------------------------------------
 * The reference has no adjoint; these entry points differentiate its 1-D time
 * loop (the awcm_run path) with the adaptive data frozen per step.
 *
 * awcm_adjoint_record: as awcm_run (same arguments, bit-identical results),
 * and records for every step its masks mask_temp/mask0..3, levels level0..2,
 * dt and the three stage alphas, plus the state at the start of every
 * checkpoint_every-th step.  Replaces any earlier recording.  max_steps also
 * sizes the tape (about 9 (Nx+6) bytes per step plus 24 (Nx+6) bytes per
 * checkpoint).  Replay and the segment recompute keep the masks in LDS, so
 * replay and awcm_adjoint return AWCM_EINVAL for instances longer than one
 * workgroup's LDS holds (about 14 800 finest points on gfx950); recording has
 * no such limit. */
int awcm_adjoint_record(awcm_solver* h, double tol, double cfl, double t_final, int max_steps, int checkpoint_every,
                        double* t, int* cnt, int* steps_done);

/* Frozen forward: replays the recorded steps from the current state (upload it
 * first) with the recorded masks, levels, dt and alphas.  From the recording's
 * initial state it reproduces the recorded run bit for bit; from a perturbed
 * state it is the map whose derivative awcm_adjoint computes. */
int awcm_adjoint_replay(awcm_solver* h);

/* dJ/dQ0 = (dQ_final/dQ0)^T dJ/dQ_final through the recorded steps (frozen
 * masks, levels, dt and alphas), by checkpoint-and-recompute: the steps are
 * reversed in segments of checkpoint_every steps (a positive multiple of the
 * recording's spacing; 0 = the recording's), each recomputed from its
 * checkpoint (one workgroup, the frozen replay) with its stage inputs kept in
 * HBM, then reversed step by step by grid-wide gather kernels (one reversed
 * step is a HIP graph replayed for every step).  Both arrays use the layout of
 * awcm_upload_q; ghost entries of dJ_dQfinal are ignored and those of dJ_dQ0
 * are returned as 0.  The device state Q is left unchanged.  Deterministic:
 * the result does not depend on the segment length and is identical between
 * calls.  Synchronous. */
int awcm_adjoint(awcm_solver* h, const double* dJ_dQfinal, int checkpoint_every, double* dJ_dQ0);

/* As awcm_run, but all n_steps are queued with no host synchronisation in
 * between (steps past t_final are no-ops on the device) and elapsed_ms is the
 * device time between two HIP events recorded on the handle's stream around
 * them.  Inputs must already be resident (awcm_upload_q). */
int awcm_run_timed(awcm_solver* h, double tol, double cfl, double t_final, int n_steps, double* t, int* cnt,
                   int* steps_done, long long* updates, double* elapsed_ms);

/* Per-kernel HIP-event timing (2-D handles).  While enabled, every queued step
 * records event pairs around the kernels of each class; awcm_profile_read
 * returns the summed device time, the number of launches and the number of
 * points those launches processed, then clears the counters. */
enum {
    AWCM_KERNEL_STAGE_WENO = 0,   /* finest-level WENO-Z/Roe stage updates     */
    AWCM_KERNEL_STAGE_COARSE = 1, /* coarse-level central-difference updates    */
    AWCM_KERNEL_PREPARE = 2       /* PrepareEuler + compaction (whole sequence) */
};
int awcm_profile_enable(awcm_solver* h, int enable);
int awcm_profile_read(awcm_solver* h, int kernel, double* total_ms, long long* launches, long long* points);

/* Launch tuning of a 2-D handle (results are identical for every setting):
 *   AWCM_TUNE_TAIL_CELLS  the coarsest levels whose lattice has at most this many
 *                         cells run each level cascade inside one workgroup
 *                         (default 0: every level is a launch of its own);
 *   AWCM_TUNE_GRAPH       1: replay the step as a captured HIP graph (default),
 *                         0: launch the kernels eagerly. */
enum { AWCM_TUNE_TAIL_CELLS = 0, AWCM_TUNE_GRAPH = 1 };
int awcm_set_tuning(awcm_solver* h, int param, int value);

/* Block until the handle's queued device work is finished. */
int awcm_synchronize(awcm_solver* h);

/* Diagnostic, no counterpart in the reference: q[i] = a[i] / b[i] computed on
 * device 0 the way the kernels form quotients that share a divisor (a
 * reciprocal RN(1/b) and one exact-residual correction per quotient; struct
 * Div in awcm/physics.h).  For operands and quotients in the normal range the
 * result is the correctly rounded quotient, the host's a[i] / b[i] bit for
 * bit.  Host arrays of n entries; n >= 0. */
int awcm_divide_shared(const double* a, const double* b, double* q, long long n);

/* ---- initial conditions (host side, reference layout) ------------------- */

/* SHOCKTUBE::Initialize (PROBLEM_SHOCKTUBE.h:14-54) on [xl, xu]: cell-centred
 * grid of res points, state `left` (rho,u,p) where x <= split, else `right`.
 * q holds (res+6)*3 doubles; dx and x0 (x of point 0) are returned. */
int awcm_ic_riemann_1d(int res, double xl, double xu, double split, const double* left, const double* right,
                       double gamma, double* q, double* dx, double* x0);
/* ENTROPY_INTERACTING::Initialize (PROBLEM_ENTROPY_INTERACTING.h:14-49):
 * `left` where x < split, else rho = 1 + amp*sin(5x), u = 0, p = 1. */
int awcm_ic_shu_osher_1d(int res, double xl, double xu, double split, const double* left, double amp,
                         double gamma, double* q, double* dx, double* x0);
/* QUADRANTS::Initialize (PROBLEM_QUADRANTS.h:15-81) on [0,1]^2, res x res
 * points, split at (split, split); states[4*k..4*k+3] = (rho,u,v,p) of
 * quadrant k+1 (NE, NW, SW, SE). */
int awcm_ic_quadrants_2d(int res, double split, const double* states, double gamma, double* q, double* dx);
/* DOUBLEMACH::Initialize (PROBLEM_DOUBLEMACH.h:13-63): ((res-1)*4+1) x res
 * points on [0,4]x[0,1], pre-shock density rho_pre. */
int awcm_ic_double_mach_2d(int res, double rho_pre, double gamma, double* q, double* dx);
/* BOX_ADVECTION::Initialize (PROBLEM_BOX_ADVECTION.h:15-47): the scalar field
 * exp(-50 (x-1/2)^2 - 50 (y-1/2)^2) on the res x res cell centres of [0,1]^2,
 * ghosts included ((res+6)^2 doubles); dx of that grid. */
int awcm_ic_box_advection_2d(int res, double* q, double* dx);

/* ---- batched 1-D form (no reference counterpart: replaces N sequential
 *      main_1d_euler.cpp runs) ------------------------------------------- */

typedef struct {
    int num_levels;
    int nx_coarse;
    double gamma;
    double dx;
    double tol;
    double cfl;
    double t_final;
} awcm_instance_1d;

/* n independent 1-D solves sharing one output-time schedule. */
int awcm_batch_create_1d(int n, const awcm_instance_1d* params, const double* print_times, int n_print,
                         awcm_batch** out);
void awcm_batch_destroy(awcm_batch* b);
int awcm_batch_upload_q(awcm_batch* b, int idx, const double* q);
int awcm_batch_download_q(awcm_batch* b, int idx, double* q);
int awcm_batch_get_masks(awcm_batch* b, int idx, uint8_t* masks, int32_t* levels);
/* Launch geometry: threads per workgroup (64..1024, multiple of 64) and
 * workgroups per launch (0 = one per instance). */
int awcm_batch_set_launch(awcm_batch* b, int threads, int workgroups);
/* Run every instance to its t_final (or max_steps steps) in one persistent
 * launch; instances are dispatched largest-first from a device work queue.
 * An instance whose state breaks down (a wave speed not finite) stops at its
 * last finite time with status AWCM_ENONFINITE in awcm_batch_state; that is
 * not an error of the call.
 * The launch (work-counter reset + solve kernel) is captured once into a HIP
 * graph per launch shape and step budget and replayed. */
int awcm_batch_run(awcm_batch* b, int max_steps);

/* Stage face fluxes of every instance (AWCM_FLUX_WENOZ or AWCM_FLUX_WENO5, as
 * awcm_set_flux); applies from the next run. */
int awcm_batch_set_flux(awcm_batch* b, int kind);
/* Project autotuner for the batched launch: times every candidate shape
 * (workgroup size 256 / 512 / 1024 x one workgroup per instance or 1x / 2x
 * the CU count of persistent workers) on probe_steps steps of every instance
 * from the current state, restores that state after each probe, and keeps the
 * fastest (as awcm_batch_set_launch would).  probe_ms, if not NULL, receives
 * the 9 probe times in that order (thread-major).  No counterpart in the
 * reference, which has no batched form. */
int awcm_batch_autotune(awcm_batch* b, int probe_steps, int* threads, int* workgroups, double* probe_ms);
/* The same solves, one launch per instance, in index order (the baseline the
 * batched form is measured against). */
int awcm_batch_run_sequential(awcm_batch* b, int max_steps);
int awcm_batch_state(awcm_batch* b, int idx, double* t, int* cnt, int* steps, long long* updates, int* status);
int awcm_batch_synchronize(awcm_batch* b);

/* ---- the 1-D solve without allocation ------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): awcm_run's 1-D time loop (same kernel, bit-identical
 * results) on caller-owned device buffers, allocating nothing.  The library's
 * working memory (descriptor, schedule, work counter, RK copy, stage input,
 * face fluxes) is one fixed device array per GPU, about 0.6 MB, held until the
 * process ends.  Limits: Nx = (nx_coarse-1)*2^(num_levels-1)+1 <= 8193 and
 * n_print <= 64; larger requests return AWCM_EINVAL.  Calls that overlap on
 * one GPU share that array and race; each call runs on the null stream and
 * returns after the solve.
 *   p            levels, coarse points, gamma, dx, tol, cfl, t_final
 *   flux         AWCM_FLUX_WENOZ or AWCM_FLUX_WENO5
 *   print_times  host array of n_print output times (the schedule)
 *   q            device, in/out: 3 planes (rho, rho u, E) of Nx+6 doubles,
 *                q[c*(Nx+6) + i+3] for i = -3 .. Nx+2
 *   masks        device, in/out: mask_temp, mask0..mask3, each Nx+6 bytes
 *                (ghosted); a new solve starts from mask0 all 1, the others 0
 *   levels       device, in/out: level0..level3, each Nx bytes
 *   t, cnt, steps_done, updates   as awcm_run */
#define AWCM_NOALLOC_MAX_NX 8193
#define AWCM_NOALLOC_MAX_PRINT 64
int awcm_run_1d_noalloc(const awcm_instance_1d* p, int flux, const double* print_times, int n_print, double* q,
                        uint8_t* masks, uint8_t* levels, int max_steps, double* t, int* cnt, int* steps_done,
                        long long* updates);

#ifdef __cplusplus
}
#endif

#endif /* AWCM_CAPI_H */
