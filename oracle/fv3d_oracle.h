/*
 * fv3d_oracle.h -- CPU restatement of fv3d's per-timestep hydro path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests, the smoke
 * check and the CPU-baseline leg of bench.py call it to produce expected
 * values and CPU timings.  Nothing in the product (fv3d_amd/, libfv3d_hip)
 * links, loads or calls it.
 *
 * Arrays use the reference's host layout: Kokkos::View<double****>
 * (Ntz, Nty, Ntx, 5), LayoutRight, so element (k, j, i, var) lives at
 * ((k*Nty + j)*Ntx + i)*5 + var, ghosts included (SimInfo.h:14, States.h:5-18).
 *
 * Parity pins (what checks each function besides the GPU path): see the
 * table at the top of fv3d_oracle.c.
 */
#ifndef FV3D_ORACLE_H
#define FV3D_ORACLE_H

#include "fv3d_params.h"

#ifdef __cplusplus
extern "C" {
#endif

/* problem ids, in the order of InitType (Init.h:217-226) */
enum { ORC_SOD_X = 0, ORC_SOD_Y, ORC_SOD_Z, ORC_BLAST, ORC_RAYLEIGH_TAYLOR, ORC_DIFFUSION,
       ORC_H84, ORC_C91 };

/* ---- state-level functions (States.h, RiemannSolvers.h) ---- */
void orc_state_prim_to_cons(const fv3d_params* p, const double q[5], double u[5]);
void orc_state_cons_to_prim(const fv3d_params* p, const double u[5], double q[5]);
void orc_compute_flux(const fv3d_params* p, const double q[5], double f[5]);
void orc_hll(const fv3d_params* p, const double qL[5], const double qR[5], double flux[5], double* pout);
void orc_hllc(const fv3d_params* p, const double qL[5], const double qR[5], double flux[5], double* pout);
/* batched Riemann solves over n state pairs stored as [n][5] */
void orc_riemann_batch(const fv3d_params* p, int solver, long n, const double* qL, const double* qR,
                       double* flux, double* pout);

/* ---- array passes (one per Kokkos kernel of the reference) ---- */
void orc_cons_to_prim(const fv3d_params* p, const double* U, double* Q);
void orc_prim_to_cons(const fv3d_params* p, const double* Q, double* U);
void orc_fill_boundaries(const fv3d_params* p, double* Q);
void orc_compute_slopes(const fv3d_params* p, const double* Q, double* sx, double* sy, double* sz);
void orc_fluxes_and_update(const fv3d_params* p, const double* Q, const double* sx, const double* sy,
                           const double* sz, double* Unew, double dt);
void orc_thermal_conduction(const fv3d_params* p, const double* Q, double* Unew, double dt);
void orc_viscosity(const fv3d_params* p, const double* Q, double* Unew, double dt);

/* work must hold 5 arrays of Ntz*Nty*Ntx*5 doubles (slopes x/y/z, U0, Ustar). */
long orc_work_size(const fv3d_params* p);
void orc_euler_step(const fv3d_params* p, double* Q, double* Unew, double dt, double* work);
void orc_update(const fv3d_params* p, double* Q, double* Unew, double dt, double* work);
/* returns dt; inv[3] receives inv_dt_hyp, inv_dt_par_tc, inv_dt_par_visc */
double orc_compute_dt(const fv3d_params* p, const double* Q, double inv[3]);

/* ---- problem setup (Init.h) ---- */
int orc_init(const fv3d_params* p, int problem, double* Q);

/* ---- the time loop of main.cpp:51-74, run for at most max_steps steps ----
 * U (cons, ghosts as set by prim_to_cons at init), Q and Unew are full arrays.
 * Writes the dt of every step into dts (if non-null, capacity max_steps).
 * Returns the number of steps taken; *t_out receives the final time.
 * With stop_at_tend=0 the loop ignores tend and runs exactly max_steps. */
long orc_run(const fv3d_params* p, double* U, double* Q, double* Unew, double* work, long max_steps,
             int stop_at_tend, double* dts, double* t_out);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif
