/*
 * fv3d_hip.h -- C ABI of libfv3d_hip, the MI355X (gfx950) implementation of
 * fv3d's per-timestep finite-volume hydro path.
 *
 * Conventions
 *  - Every entry point returns 0 on success and a non-zero FV3D_E* code on
 *    failure; fv3d_last_error(ctx) (or fv3d_last_error(NULL) for calls that
 *    take no context) gives a message.  Nothing throws across the ABI.
 *  - A context binds one device, one HIP stream and all device buffers of one
 *    box.  Calls on a context are stream-ordered and not thread-safe.
 *  - Host arrays use the reference's host layout: (Ntz, Nty, Ntx, 5) with
 *    ghosts, or (Nz, Ny, Nx, 5) without, variable fastest
 *    (Kokkos::View<double****> LayoutRight, SimInfo.h:14, States.h:5-18).
 *  - On the device the state is conserved variables in SoA layout
 *    [5][Nz][Ny][Nx] (domain only, x fastest); ghosts are produced inside the
 *    kernels from the boundary index maps.
 */
#ifndef FV3D_HIP_H
#define FV3D_HIP_H

#include <stddef.h>
#include "fv3d_params.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FV3D_OK 0
#define FV3D_EINVAL 1      /* bad argument or unsupported parameter combination */
#define FV3D_EHIP 2        /* a HIP runtime call failed */
#define FV3D_ENOMEM 3      /* device allocation failed */
#define FV3D_ENONFINITE 4  /* the state produced NaN/inf (pressure < 0, density 0, ...) */
#define FV3D_EFLOOR 5      /* fv3d_adjoint_run: the density floor acted in the run it differentiates */

#define FV3D_FP64 0        /* bit-equal to the reference's FP64 arithmetic */
#define FV3D_FP32 1        /* FP32 state and fluxes; RK2 accumulation Kahan-compensated */
#define FV3D_FP32_PLAIN 2  /* FP32 state and fluxes, uncompensated */

typedef struct fv3d_ctx fv3d_ctx;

/* library identification: ABI version and sizeof(fv3d_params) it was built with */
int fv3d_abi_version(void);
size_t fv3d_params_size(void);
const char* fv3d_last_error(const fv3d_ctx* ctx);

/* Context = the state owned by the reference's main() (main.cpp:28-30) plus
 * the functors built from Params (main.cpp:38-41): UpdateFunctor
 * (Update.h:45), ComputeDtFunctor (ComputeDt.h:13), BoundaryManager
 * (BoundaryConditions.h:150).  precision is FV3D_FP64, FV3D_FP32 or
 * FV3D_FP32_PLAIN; the dt reduction is FP64 in every precision. */
int fv3d_create(const fv3d_params* p, int precision, int device, fv3d_ctx** out);
int fv3d_destroy(fv3d_ctx* ctx);

/* primToCons(Q, U) after init (main.cpp:43-44, SimInfo.h:323-331): Q is a
 * host primitive array, with ghosts if with_ghosts (ghost values are not
 * read).  Resets the step counter and computes the first inv_dt_hyp. */
int fv3d_set_prim(fv3d_ctx* ctx, const double* Q, int with_ghosts);
/* consToPrim(U, Q) (SimInfo.h:313-321) of the current state; with ghosts the
 * ghost shell is filled by fillBoundaries (BoundaryConditions.h:146-215),
 * which is what the reference's Q holds inside update(). */
int fv3d_get_prim(fv3d_ctx* ctx, double* Q, int with_ghosts);
/* The reference's Q after update() returns (main.cpp:76 snapshot source):
 * RK2: prim(U*) of the last step (Update.h:194); Euler: prim(U^n) of the
 * last step.  Ghosts filled as above. */
int fv3d_get_stage_prim(fv3d_ctx* ctx, double* Q, int with_ghosts);
/* conserved state in host layout (checkpoint / restart). */
int fv3d_set_cons(fv3d_ctx* ctx, const double* U, int with_ghosts);
int fv3d_get_cons(fv3d_ctx* ctx, double* U, int with_ghosts);

/* ComputeDtFunctor::computeDt (ComputeDt.h:17-63) on prim(U) of the current
 * state: inv[3] = {inv_dt_hyp, inv_dt_par_tc, inv_dt_par_visc}, *dt =
 * CFL/max(inv).  Synchronises; returns FV3D_ENONFINITE if a cell produced a
 * non-finite inverse time step (negative pressure or zero density).  inv and
 * *dt are written in that case too, from the finite cells: the reference's
 * Max reducer drops NaN and its run goes on, so a caller that follows the
 * reference treats this code as a warning and keeps stepping. */
int fv3d_compute_dt(fv3d_ctx* ctx, double inv[3], double* dt);

/* UpdateFunctor::update (Update.h:180-206) with dt = computeDt(prim(U)):
 * one Euler or SSP-RK2 step per params.time_stepping.  dt stays on the
 * device (no host synchronisation); read it with fv3d_compute_dt before the
 * step as main.cpp:57 does. */
int fv3d_step(fv3d_ctx* ctx);
/* UpdateFunctor::update(Q, Unew, dt) (Update.h:180-206) with the caller's
 * dt instead of the device-side computeDt value: one Euler or SSP-RK2 step
 * of the fused path.  The step after it reads dt from the ring again (the
 * state this step produced reduces it, as after fv3d_step).  dt must be
 * finite (FV3D_EINVAL otherwise). */
int fv3d_step_dt(fv3d_ctx* ctx, double dt);
/* nsteps steps back to back (no host synchronisation in between). */
int fv3d_steps(fv3d_ctx* ctx, int nsteps);
/* The number of steps taken since fv3d_set_prim / fv3d_set_cons. */
long fv3d_step_count(const fv3d_ctx* ctx);
int fv3d_synchronize(fv3d_ctx* ctx);

/* Timing on the context's stream with HIP events: runs nsteps steps and
 * returns the elapsed milliseconds of the whole window and of each stage
 * kernel launch (stage_ms must hold 2*nsteps floats for RK2, nsteps for
 * Euler; may be NULL). */
int fv3d_time_steps(fv3d_ctx* ctx, int nsteps, float* total_ms, float* stage_ms);

/* ---- the reference's functors one pass at a time, on device arrays ----
 *
 * An fv3d_array is one Array of the reference (Kokkos::View<double****> of
 * extents (Ntz, Nty, Ntx, 5), SimInfo.h:14) in device memory: FP64 with
 * ghosts, stored variable-major [5][Ntz][Nty][Ntx] with x fastest.  It
 * belongs to the context it was made for, uses that context's stream and
 * is zero when created, like a new View.  Each entry point below is one
 * pass of the reference over the same index range (range_tot, range_dom,
 * range_slopes; SimInfo.h:298-303), one kernel per pass, so a caller that
 * drives the functors one by one gets the reference's values pass for pass.
 * fv3d_step is the fused form of the same arithmetic.  Calls are
 * stream-ordered; only upload, download and fv3d_compute_dt_array wait. */
typedef struct fv3d_array fv3d_array;

int fv3d_array_create(fv3d_ctx* ctx, fv3d_array** out);
int fv3d_array_destroy(fv3d_array* a);
/* device address of the data, [5][Ntz][Nty][Ntx] doubles */
double* fv3d_array_data(const fv3d_array* a);
/* host array in the reference's host layout (Ntz, Nty, Ntx, 5) <-> device */
int fv3d_array_upload(fv3d_array* a, const double* host);
int fv3d_array_download(const fv3d_array* a, double* host);
/* Kokkos::deep_copy(dst, src) */
int fv3d_array_copy(fv3d_array* dst, const fv3d_array* src);

/* consToPrim(U, Q) / primToCons(Q, U) over every cell, ghosts included
 * (SimInfo.h:313-331) */
int fv3d_cons_to_prim_array(fv3d_ctx* ctx, const fv3d_array* U, fv3d_array* Q);
int fv3d_prim_to_cons_array(fv3d_ctx* ctx, const fv3d_array* Q, fv3d_array* U);
/* BoundaryManager::fillBoundaries(Q) in place (BoundaryConditions.h:146-215) */
int fv3d_fill_boundaries_array(fv3d_ctx* ctx, fv3d_array* Q);
/* UpdateFunctor::computeSlopes(Q) (Update.h:55-88) into sx, sy, sz */
int fv3d_compute_slopes(fv3d_ctx* ctx, const fv3d_array* Q, fv3d_array* sx, fv3d_array* sy,
                        fv3d_array* sz);
/* UpdateFunctor::computeFluxesAndUpdate(Q, Unew, dt) (Update.h:90-162)
 * with the given slopes */
int fv3d_fluxes_and_update(fv3d_ctx* ctx, const fv3d_array* Q, const fv3d_array* sx,
                           const fv3d_array* sy, const fv3d_array* sz, fv3d_array* Unew,
                           double dt);
/* ThermalConductionFunctor::applyThermalConduction(Q, Unew, dt)
 * (ThermalConduction.h:32-97) */
int fv3d_apply_tc(fv3d_ctx* ctx, const fv3d_array* Q, fv3d_array* Unew, double dt);
/* ViscosityFunctor::applyViscosity(Q, Unew, dt) (Viscosity.h:22-155) */
int fv3d_apply_visc(fv3d_ctx* ctx, const fv3d_array* Q, fv3d_array* Unew, double dt);
/* UpdateFunctor::euler_step(Q, Unew, dt) (Update.h:164-178): ghost fill of
 * Q, slopes (PLM), flux update, then conduction and viscosity if active;
 * the slopes live in the context, as UpdateFunctor owns its slopesX/Y/Z */
int fv3d_euler_step(fv3d_ctx* ctx, fv3d_array* Q, fv3d_array* Unew, double dt);
/* UpdateFunctor::update(Q, Unew, dt) (Update.h:180-206): Euler, or SSP-RK2
 * with U0 and Ustar of the full (Ntz, Nty, Ntx, 5) shape (the reference
 * gives them three extents, which aborts in deep_copy); mutates Q as the
 * reference does.  U0 and Ustar are kept by the context after first use. */
int fv3d_update(fv3d_ctx* ctx, fv3d_array* Q, fv3d_array* Unew, double dt);
/* ComputeDtFunctor::computeDt(Q, ...) (ComputeDt.h:17-63) on a primitive
 * array; synchronises.  FV3D_ENONFINITE as for fv3d_compute_dt. */
int fv3d_compute_dt_array(fv3d_ctx* ctx, const fv3d_array* Q, double inv[3], double* dt);
/* the fused path's state <-> the domain of a conserved array, on the
 * device (ghosts of U untouched); set resets the step count and dt as
 * fv3d_set_cons does */
int fv3d_get_cons_array(fv3d_ctx* ctx, fv3d_array* U);
int fv3d_set_cons_array(fv3d_ctx* ctx, const fv3d_array* U);

/* ---- functor-level entry points on host arrays (unit parity) ---- */

/* BoundaryManager::fillBoundaries (BoundaryConditions.h:146-215) on a host
 * primitive array with ghosts, run on the device. */
int fv3d_fill_boundaries(const fv3d_params* p, double* Q);
/* State primToCons / consToPrim (States.h:20-45) of n states [n][5] with
 * the given gamma0, run on the device. */
int fv3d_prim_to_cons_states(double gamma0, long n, const double* q, double* u);
int fv3d_cons_to_prim_states(double gamma0, long n, const double* u, double* q);
/* hll / hllc (RiemannSolvers.h:6,39) on n face-frame state pairs [n][5]. */
int fv3d_riemann(int solver, double gamma0, long n, const double* qL, const double* qR,
                 double* flux, double* pout);

/* fv3d_riemann without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight host loops, harnesses that forbid hipMalloc).  The faces
 * are staged in a fixed device array of FV3D_NOALLOC_MAX_FACES faces
 * (8 MiB) that each device holds until the process ends; n beyond that is
 * rejected with FV3D_EINVAL.  Runs on the current device's null stream and
 * returns when the results are in flux and pout.  Calls that overlap on one
 * device share that array and race. */
#define FV3D_NOALLOC_MAX_FACES 65536
int fv3d_riemann_noalloc(int solver, double gamma0, long n, const double* qL, const double* qR,
                         double* flux, double* pout);

/* Measurement support (no counterpart in the reference): a plain device copy
 * of nbytes (16-byte accesses), run reps times; *ms_per_copy is the mean
 * time of one copy from HIP events.  Gives the measured copy bandwidth and
 * calibrates the profiler's byte counters (nbytes read + nbytes written). */
int fv3d_stream_copy(size_t nbytes, int reps, float* ms_per_copy);

/* ---- frozen-dt runs and the discrete adjoint (no counterpart in the
 * reference) ----
 *
 * fv3d_forward_frozen advances the context's state by nsteps steps of
 * UpdateFunctor::update (Update.h:180-206) with the given dt sequence
 * instead of computeDt's (the adjoint's forward map); *ms (may be NULL)
 * gets the device time.  Afterwards the context restarts its step count and
 * dt ring from the new state, as after fv3d_set_cons.  FP32 contexts run
 * the uncompensated FP32 RK2.
 *
 * fv3d_adjoint_run differentiates J(U^n) for the run of nsteps frozen-dt
 * steps that starts from the context's current state: given dJ_dUT =
 * dJ/dU^n (conserved, host layout without ghosts (Nz, Ny, Nx, 5)), it
 * returns dJ_dU0 = dJ/dU^0 (same layout) and, if not NULL, dJ/dkappa and
 * dJ/dmu (constant conductivity and viscosity).  The forward pass keeps a
 * checkpoint every K steps; each K-step segment is recomputed in the
 * backward pass.  Branches (minmod selection, Riemann wave region) are
 * those of the forward evaluation; the density floor is taken as inactive,
 * and a run in which it changes a density fails with FV3D_EFLOOR (no
 * gradient is written).
 * Gravity and the well-balanced variants are rejected (FV3D_EINVAL).  The
 * context's state is unchanged.
 * fv3d_adjoint_times gives the device times of the last fv3d_adjoint_run:
 * ms[0] forward pass, ms[1] backward pass (recomputation included),
 * ms[2] total. */
int fv3d_forward_frozen(fv3d_ctx* ctx, int nsteps, const double* dts, float* ms);
int fv3d_adjoint_run(fv3d_ctx* ctx, int nsteps, int K, const double* dts, const double* dJ_dUT,
                     double* dJ_dU0, double* dJ_dkappa, double* dJ_dmu);
int fv3d_adjoint_times(const fv3d_ctx* ctx, float ms[3]);
/* Device memory (bytes) that later fv3d_adjoint_run calls on this context
 * may use to keep whole segments from the forward pass: a kept segment's
 * U^s and stage states are stored as the forward pass computes them, and
 * the backward pass sweeps it without recomputing it.  The last segments
 * are kept first; 0 (the default) recomputes every segment.  A segment of
 * K RK2 steps takes (2K+1) * 5 * Nx*Ny*Nz * sizeof(real) bytes (K+1 for
 * Euler).  The result does not depend on it. */
int fv3d_set_adjoint_store(fv3d_ctx* ctx, size_t bytes);

/* ---- ensembles of independent boxes (no counterpart in the reference) ----
 *
 * nbox boxes, each a complete fv3d run with its own fv3d_params (sizes,
 * gamma, boundaries, gravity, coefficients, CFL) and its own dt, advanced
 * together.  All boxes must share reconstruction, riemann_solver,
 * time_stepping and the conduction / viscosity switches: they select the
 * kernel.  Each box's step is the fv3d_step of a context with the same
 * parameters and precision, so a box's result is bit-equal to that
 * context's.  The reference runs one box per process; this replaces a loop
 * over such runs (main.cpp:51-76 per box). */
typedef struct fv3d_batch fv3d_batch;

#define FV3D_BATCH_GRAPH 0       /* one persistent launch per RK stage, replayed from a hipGraph */
#define FV3D_BATCH_PERSISTENT 1  /* the same launches, issued from the host each step */
#define FV3D_BATCH_LOOP 2        /* one single-box launch per box per stage */

int fv3d_batch_create(int nbox, const fv3d_params* params, int precision, int device,
                      fv3d_batch** out);
int fv3d_batch_destroy(fv3d_batch* b);
const char* fv3d_batch_last_error(const fv3d_batch* b);
/* primToCons of box `box` from a host primitive array (layout as for
 * fv3d_set_prim) and its first inv_dt_hyp; the other boxes are untouched. */
int fv3d_batch_set_prim(fv3d_batch* b, int box, const double* Q, int with_ghosts);
/* consToPrim of box `box`'s current state, ghosts filled if with_ghosts. */
int fv3d_batch_get_prim(fv3d_batch* b, int box, double* Q, int with_ghosts);
/* dts[nbox]: each box's CFL / max(inv_dt_hyp, inv_dt_par_tc, inv_dt_par_visc)
 * of its current state (ComputeDt.h:17-63).  Synchronises. */
int fv3d_batch_compute_dt(fv3d_batch* b, double* dts);
/* nsteps steps of every box, each with its own dt; mode is FV3D_BATCH_*. */
int fv3d_batch_step(fv3d_batch* b, int nsteps, int mode);
/* nsteps steps timed with HIP events around the window (*ms). */
int fv3d_batch_time(fv3d_batch* b, int nsteps, int mode, float* ms);
long fv3d_batch_step_count(const fv3d_batch* b);
/* total interior cells over all boxes */
long long fv3d_batch_cells(const fv3d_batch* b);
/* launch shape of the persistent kernels: z-chunk cap per work item (0:
 * min(Nz, 32)) and workgroups per launch (0: one per CU). */
int fv3d_batch_configure(fv3d_batch* b, int nzc_cap, int grid);
/* Stage tile width of every box in the persistent kernels: 64 or 32
 * columns, or 0 for the default (32 when 64-column tiles would cover at
 * least 10% more cells over the ensemble, else 64).  Results do not depend
 * on it.  No reference counterpart: the reference has no ensemble. */
int fv3d_batch_set_tile_width(fv3d_batch* b, int wx);
/* the tile width in use (-1 on a NULL batch) */
int fv3d_batch_tile_width(const fv3d_batch* b);
/* Autotuner: times nsteps graph steps for each tile width in {64, 32}, z-chunk
 * cap in {8,16,32,64,128} and grid in {1,2,4} x CUs on a copy of the current
 * state (restored afterwards), then configures the fastest.  cache_path (may
 * be NULL) is a text file of "signature nzc grid wx" lines: a hit skips the
 * search, a miss appends the result.  FV3D_TILE_WX=64 or 32 in the
 * environment fixes the width. */
int fv3d_batch_tune_shape(fv3d_batch* b, int nsteps, const char* cache_path, int* nzc_out,
                          int* grid_out, int* wx_out);
/* fv3d_batch_tune_shape without the width output */
int fv3d_batch_tune(fv3d_batch* b, int nsteps, const char* cache_path, int* nzc_out,
                    int* grid_out);

#ifdef __cplusplus
}
#endif

#endif /* FV3D_HIP_H */
