/* hmhd.h -- C ABI of the MI355X per-step update of hosono-2DMHD.
 *
 * The reference has no FFI: its hot path is two Fortran module procedures
 * called from the driver's main loop,
 *     dt = cfl_computedt(grid_block, grid_state)      src/mhd_driver.f90:78
 *     call timeStep_advanceSolution(dt, grid_block, grid_state)
 *                                                     src/mhd_driver.f90:81
 * plus the set-up calls simulation_init / GP_init / grid_init
 * (src/mhd_driver.f90:40-52).  The entry points below replace exactly those
 * calls; a Fortran driver binds them with iso_c_binding (the interfaces in
 * hosono-2dmhd_amd/fortran/hmhd_c.f90).
 *
 * Array layouts are the reference's Fortran layouts on the padded grid
 * (NGC = R + 2 guard cells per side, src/grid.f90:71-81):
 *     U(8, mx, my), V(10, mx, my)   FP64, column-major, mx = nx + 2 NGC
 *     scheme(mx, my)                int32
 * Every function returns HMHD_OK (0) or an error code; hmhd_last_error()
 * gives the message.  A handle is used by one host thread at a time; any
 * number of handles may coexist.
 */
#ifndef HMHD_H
#define HMHD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HMHD_OK 0
#define HMHD_ERR_ARG 1     /* invalid argument or configuration (the reference stops) */
#define HMHD_ERR_DEVICE 2  /* HIP runtime error */
#define HMHD_ERR_GP 3      /* GP covariance not positive definite (src/linAlgQuadPrecision.f90:136) */
#define HMHD_ERR_TIMEOUT 4 /* hmhd_server_wait: the request did not finish in time */

#define HMHD_BC_PERIODIC 0 /* sim_BC = 'periodic' */
#define HMHD_BC_OUTFLOW 1  /* sim_BC = 'outflow'  */

#define HMHD_RIEMANN_HLLC 1 /* sim_riemannSolver = 'hllc' | 'hllcmhd' (src/simulation.f90:62-72) */

/* The par-file quantities the step reads (src/simulation.f90:43-58,
 * src/grid.f90:58-63, src/GP.f90:58-69). */
typedef struct hmhd_config {
    int32_t nx, ny;             /* grid_Nx, grid_Ny */
    double xbeg, xend;          /* grid_xBeg, grid_xEnd */
    double ybeg, yend;          /* grid_yBeg, grid_yEnd */
    int32_t R;                  /* GP_maxRadius, 1..3 */
    double ell_over_delta;      /* GP_ellOverDelta */
    double gamma;               /* sim_gamma */
    double small_p, small_rho;  /* sim_smallPressure, sim_smallDensity */
    double small_B;             /* sim_smallB (read by the reference, unused on the path) */
    int32_t force_hydro;        /* sim_forceHydro */
    int32_t bc;                 /* HMHD_BC_* */
    double cfl;                 /* sim_cfl */
    int32_t riemann;            /* HMHD_RIEMANN_HLLC */
} hmhd_config;

typedef struct hmhd_handle hmhd_t;

/* Replaces simulation_init + GP_init + grid_init (src/mhd_driver.f90:40-52):
 * validates cfg, builds the GP prediction vectors in binary128
 * (src/GP.f90:37-200) and allocates the device state.  Requires
 * nx, ny >= 2 (R + 2). */
int hmhd_create(const hmhd_config* cfg, hmhd_t** out);
void hmhd_destroy(hmhd_t* h);
/* Message of the last failed call on h; with h == NULL, of the last failed
 * hmhd_create on this thread. */
const char* hmhd_last_error(const hmhd_t* h);

/* NGC and the padded extents mx = nx + 2 NGC, my = ny + 2 NGC. */
int hmhd_dims(const hmhd_t* h, int32_t* ngc, int32_t* mx, int32_t* my);

/* GP tables in the reference's module layouts (src/GP.f90:12-21):
 * predvec(nsMax, nqMax, 4, R), quadw(nqMax, R), stenc(2, nsMax, R), with
 * nsMax = 1 + 2R(R+1), nqMax = R + 1.  Any pointer may be NULL. */
int hmhd_get_gp(const hmhd_t* h, double* predvec, double* quadw, int32_t* stenc);
/* The same tables straight from a configuration, without a device or a
 * handle: the binary128 construction of GP_init (src/GP.f90:37-200). */
int hmhd_gp_compute(const hmhd_config* cfg, double* predvec, double* quadw, int32_t* stenc);
/* Replace the GP tables with ones the caller computed (same layouts). */
int hmhd_set_gp(hmhd_t* h, const double* predvec, const double* quadw);

/* Host -> device: the state initialCondition_set leaves in grid_state
 * (src/initialCondition.f90:23-84).  U is required; V may be NULL.  When V
 * is given, the next hmhd_cfl_dt reads it, as cfl_computedt reads the IC's
 * primitives on the first step; otherwise it derives V from U. */
int hmhd_upload_state(hmhd_t* h, const double* U, const double* V);

/* Replaces cfl_computedt (src/cfl.f90:14-52).  Before it returns it also
 * enqueues the part of the next hmhd_advance that does not depend on dt
 * (the first pass's face work of stage 0), which then overlaps the
 * caller's handling of dt; hmhd_download in between is unaffected, and an
 * upload or GP-table change in between discards that work. */
int hmhd_cfl_dt(hmhd_t* h, double* dt);

/* Replaces timeStep_advanceSolution (src/timeStep.f90:22-43): one SSP-RK3
 * step of size dt with the full MOOD loop in every stage.  Enqueues and
 * returns after the last MOOD pass has been decided; the state stays on
 * the device. */
int hmhd_advance(hmhd_t* h, double dt);

/* Device -> host for output steps: U and V with guard cells as the
 * reference holds them after a step (src/timeStep.f90:32-41), and the
 * final scheme of the last stage (read by src/output.f90:251-253).  Any
 * pointer may be NULL. */
int hmhd_download(hmhd_t* h, double* U, double* V, int32_t* scheme);

/* MOOD statistics of the last hmhd_advance: passes and inner
 * admissibility iterations per RK stage (3 each). */
int hmhd_stats(const hmhd_t* h, int32_t* passes, int32_t* inner);

/* When on != 0, every later hmhd_advance records the final scheme of each
 * of its three stages; hmhd_stage_schemes copies them out as
 * scheme(mx, my, 3). */
int hmhd_record_stage_schemes(hmhd_t* h, int32_t on);
int hmhd_stage_schemes(const hmhd_t* h, int32_t* out);

/* Kernel timing with HIP events on the handle's stream.  `families` is a
 * bit mask over the kernel families (bit f: family f in the order below;
 * HMHD_PROFILE_ALL for every one, 0 to stop): each launch of a selected
 * family is bracketed by events.  hmhd_profile_read returns, per family,
 * the total milliseconds and launch counts since the last reset (order:
 * face_full, face_flux, update, recon, flags, cfl, tiles; "recon" times
 * nothing (the inner-loop lowering runs in the "flags" family's scan
 * launch) and keeps its place in the order; "tiles" holds counts, not
 * times: ms[6] counts MOOD iterations (host round trips) and launches[6]
 * the active tiles they visited). */
#define HMHD_PROFILE_ALL 0x7f
int hmhd_profile(hmhd_t* h, int32_t families);
int hmhd_profile_read(hmhd_t* h, double* ms, int64_t* launches, int32_t reset);

/* Wait for all work enqueued on the handle. */
int hmhd_synchronize(hmhd_t* h);

/* The first pass's face kernel (the dominant one) comes in tuned forms:
 * form v = 0..HMHD_FACE_FORMS-1 is a tile of 16 x 16, 32 x 8 or 32 x 16
 * cells compiled for two waves per SIMD (v = 0, 1, 2) or 16 x 16, 32 x 8
 * compiled for one (v = 3, 4), for every GP radius.  Every form computes
 * the same bits.  Process-wide, for every handle, batch and server:
 * mode 1 (the default) measures every form on the first first pass of each
 * (radius, launch size class) and keeps the fastest; mode 0 always runs
 * form 0; mode 2 + v always runs form v.  No counterpart in the reference
 * (build-owned tuning). */
#define HMHD_FACE_FORMS 5
typedef struct hmhd_face_tune {
    int32_t R;
    int32_t size_class;              /* padded cells of the launch: < 2^16, < 2^19, < 2^21, or more */
    int32_t form;                    /* the form kept */
    int32_t tile_x, tile_y, waves;   /* its shape and wave target */
    double ms[8];                    /* per form: fastest of three launches, milliseconds (0: not run) */
} hmhd_face_tune;
int hmhd_face_tuning(int32_t mode);
/* Copies up to max entries of the tuning table; *n receives their number. */
int hmhd_face_tuning_read(hmhd_face_tune* out, int32_t max, int32_t* n);
int hmhd_face_tuning_clear(void);

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the dt of hmhd_cfl_dt (the same
 * arithmetic and bits, src/cfl.f90:14-52) for a state the caller keeps in
 * device memory, without a handle and without allocating.  U is
 * U(8, mx, my) and V, optional, V(10, mx, my) in the reference's layouts
 * (read as the first step's cfl_computedt reads the IC's primitives; NULL:
 * derived from U); dt is one double of device memory, written by the
 * last of two kernels enqueued on `stream` (a hipStream_t; NULL for the
 * null stream).  The block maxima go to a fixed 16 KiB device array per
 * GPU: calls that overlap on one GPU race.  At most HMHD_NOALLOC_MAX_CELLS
 * padded cells; larger grids, and errors, return a code with the message
 * in hmhd_last_error(NULL). */
#define HMHD_NOALLOC_MAX_CELLS (1LL << 30)
int hmhd_cfl_dt_noalloc(const hmhd_config* cfg, const double* U, const double* V, double* dt, void* stream);

/* ------------------------------------------------------------------ *
 * Batch: many independent runs of the driver loop of src/mhd_driver.f90:
 * 76-95 (cfl_computedt, validTimeStep, timeStep_advanceSolution until
 * sim_tmax or sim_nStep) advanced together, one kernel launch per family
 * covering every run's tiles.  Runs may differ in size, R, boundary
 * condition and physics; each keeps its own dt and t, and its results are
 * bit-identical to the same run on its own handle.  The reference has no
 * batched form: this replaces running its executable once per case.
 * A batch is used by one host thread at a time.
 */
typedef struct hmhd_batch_handle hmhd_batch_t;

#define HMHD_RUN_ACTIVE 0  /* still stepping */
#define HMHD_RUN_DONE 1    /* reached tmax or nstepmax (src/mhd_driver.f90:76) */
#define HMHD_RUN_ABORTED 2 /* dt negative or NaN (src/mhd_driver.f90:128-141) */

int hmhd_batch_create(hmhd_batch_t** out);
void hmhd_batch_destroy(hmhd_batch_t* b);
const char* hmhd_batch_last_error(const hmhd_batch_t* b);
/* Add a run from its initial state (layouts as hmhd_upload_state; V may be
 * NULL); it stops at t = tmax or after nstepmax steps (nstepmax <= 0: no
 * step limit).  *id names the run in the calls below. */
int hmhd_batch_add(hmhd_batch_t* b, const hmhd_config* cfg, const double* U, const double* V, double tmax,
                   int32_t nstepmax, int64_t* id);
/* Advance every unfinished run by up to max_steps steps (max_steps <= 0:
 * until all runs have finished). */
int hmhd_batch_run(hmhd_batch_t* b, int32_t max_steps);
/* t, steps taken and HMHD_RUN_* state of one run. */
int hmhd_batch_status(const hmhd_batch_t* b, int64_t id, double* t, int32_t* steps, int32_t* state);
int hmhd_batch_dims(const hmhd_batch_t* b, int64_t id, int32_t* ngc, int32_t* mx, int32_t* my);
/* As hmhd_download, for one run of the batch. */
int hmhd_batch_download(hmhd_batch_t* b, int64_t id, double* U, double* V, int32_t* scheme);
/* As hmhd_profile / hmhd_profile_read, over every run of the batch. */
int hmhd_batch_profile(hmhd_batch_t* b, int32_t families);
int hmhd_batch_profile_read(hmhd_batch_t* b, double* ms, int64_t* launches, int32_t reset);

/* ------------------------------------------------------------------ *
 * Server: runs submitted at any time are admitted in arrival order while
 * their device footprint fits the HBM budget (later, smaller requests are
 * admitted past one that does not fit yet), kept resident on the device,
 * stepped with the other admitted runs, and released when they finish.
 * Admitted runs go to lanes that overlap on the device, each an engine
 * with its own stream and worker thread: one lane steps any number of runs
 * together, three more hold one large run (512^2 cells or more) each.  The
 * calls below are thread-safe.  Results are bit-identical to the same run
 * on its own handle.  A run that finds all GP table sets in use waits for
 * a running one to finish.  The reference has no served form.
 */
typedef struct hmhd_server_handle hmhd_server_t;

#define HMHD_REQ_QUEUED 0
#define HMHD_REQ_RUNNING 1
#define HMHD_REQ_DONE 2    /* reached tmax or nstepmax */
#define HMHD_REQ_ABORTED 3 /* dt negative or NaN */
#define HMHD_REQ_FAILED 4  /* footprint above the budget, or a device error */

typedef struct hmhd_result {
    int32_t state;         /* HMHD_REQ_* */
    int32_t steps;         /* steps taken */
    double t;              /* simulation time reached */
    int64_t zone_updates;  /* steps * nx * ny */
    double submit_s, admit_s, done_s; /* seconds since the server started */
} hmhd_result;

typedef struct hmhd_server_stats {
    int64_t submitted, completed;
    int32_t active, queued, peak_active, reserved;
    uint64_t bytes_in_use, peak_bytes; /* device bytes held by admitted runs */
    int64_t engine_steps;              /* engine steps, summed over the lanes */
    int64_t zone_updates;              /* of completed runs */
    double busy_s, elapsed_s;          /* time stepping, summed over the lanes; time since start */
} hmhd_server_stats;

/* Device bytes one run of this configuration holds while admitted. */
uint64_t hmhd_solve_footprint(const hmhd_config* cfg);
/* max_active <= 0: no limit besides the budget. */
int hmhd_server_create(uint64_t hbm_budget_bytes, int32_t max_active, hmhd_server_t** out);
/* Stops the worker after its current step; unfinished runs are dropped. */
void hmhd_server_destroy(hmhd_server_t* s);
const char* hmhd_server_last_error(const hmhd_server_t* s);
/* Queue a run (the state is copied; layouts as hmhd_upload_state). */
int hmhd_server_submit(hmhd_server_t* s, const hmhd_config* cfg, const double* U, const double* V, double tmax,
                       int32_t nstepmax, int64_t* req);
int hmhd_server_poll(hmhd_server_t* s, int64_t req, hmhd_result* res);
/* Block until the run has finished (timeout_s <= 0: no limit), copy its
 * result and final U, V, scheme (any pointer may be NULL) and forget it. */
int hmhd_server_wait(hmhd_server_t* s, int64_t req, double timeout_s, hmhd_result* res, double* U, double* V,
                     int32_t* scheme);
int hmhd_server_stats_read(hmhd_server_t* s, hmhd_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif
