/*
 * fv3d_oracle.c -- FP64 CPU restatement of fv3d's per-timestep hydro path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check and
 * bench.py's CPU-baseline leg load this library.  The product never does.
 *
 * It keeps the reference's pass structure (one loop nest per Kokkos kernel,
 * separate c2p / ghost fill / slopes / flux update / conduction / viscosity /
 * RK combine passes and the state copies of main.cpp) and the reference's
 * C expression trees, so that compiled with -O2 -ffp-contract=off and no
 * -march it produces the values the reference's own OpenMP build would.
 * The one deliberate change is the RK2 scratch allocation: Update.h:184-185
 * builds U0/Ustar with three extents for a rank-4 View, which aborts in
 * deep_copy; here they are full (Ntz,Nty,Ntx,5) arrays.
 *
 * The reference cannot be compiled here (Kokkos, HighFive and inih are not
 * vendored), so this restatement is pinned by other means (tests/):
 *   state prim<->cons, computeFlux  closed forms, round trips
 *   hll                             textbook HLL with Davis speeds; qL==qR;
 *                                   supersonic selection; Toro test states
 *   hllc                            textbook HLLC in 1D (v=w=0); qL==qR; the
 *                                   dropped-velocity energy (RiemannSolvers.h:52-55)
 *                                   as a closed form at qL==qR; the passive w is
 *                                   upwinded and changes no other flux bit
 *   ghost fill                      brute-force index maps on tiny grids
 *   slopes                          hand cases, mirror antisymmetry
 *   flux update + RK2               exact Sod solution (L1 convergence),
 *                                   conservation, mirror symmetry, sod_x == sod_y^T
 *   dt                              closed form at t=0
 *   conduction                      linear-T zero divergence, periodic conservation
 *   viscosity                       uniform flow -> 0; x/y branches against the
 *                                   textbook stress with the missing 1/d; the
 *                                   z-branch index and component quirks
 *                                   (Viscosity.h:122-129) by hand-derived closed
 *                                   forms on bilinear fields
 *   gravity, PCM_WB, WB wall flux   a discrete hydrostatic state stays at rest
 *                                   (and moves without the shift or the wall
 *                                   flux); sum(E - g rho z) is conserved
 *   C91 walls                       no mass through the walls; ghost values by
 *                                   hand on a 4-level column (the gravity-flag
 *                                   pressure of BoundaryConditions.h:104)
 *   B02 conductivity                one conductivity per face (energy conserved
 *                                   between insulating walls); kappa at zmid and
 *                                   the tanh slope as closed forms; ComputeDt's
 *                                   integer-index use (ComputeDt.h:39-41) as a
 *                                   closed form of inv_dt_par_tc
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "fv3d_oracle.h"

#define IDX(p, i, j, k) ((((long)(k) * (p)->Nty + (j)) * (p)->Ntx + (i)) * 5)

static inline void swap_component(double q[5], int dir) {  /* States.h:103-111 */
  double t;
  if (dir == FV3D_IY) { t = q[1]; q[1] = q[2]; q[2] = t; }
  else if (dir == FV3D_IZ) { t = q[1]; q[1] = q[3]; q[3] = t; }
}

static inline void get_pos(const fv3d_params* p, int i, int j, int k, double pos[3]) {  /* SimInfo.h:162-167 */
  pos[0] = p->xmin + (i - p->ibeg + 0.5) * p->dx;
  pos[1] = p->ymin + (j - p->jbeg + 0.5) * p->dy;
  pos[2] = p->zmin + (k - p->kbeg + 0.5) * p->dz;
}

/* ------------------------------------------------------------------------ */
/* State-level functions                                                     */
/* ------------------------------------------------------------------------ */

void orc_state_prim_to_cons(const fv3d_params* p, const double q[5], double u[5]) {  /* States.h:20-31 */
  double r[5];
  r[0] = q[0];
  r[1] = q[0] * q[1];
  r[2] = q[0] * q[2];
  r[3] = q[0] * q[3];
  double Ek = 0.5 * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3]) / q[0];
  r[4] = (Ek + q[4] / (p->gamma0 - 1.0));
  memcpy(u, r, sizeof r);
}

void orc_state_cons_to_prim(const fv3d_params* p, const double u[5], double q[5]) {  /* States.h:34-45 */
  double r[5];
  r[0] = u[0];
  r[1] = u[1] / u[0];
  r[2] = u[2] / u[0];
  r[3] = u[3] / u[0];
  double Ek = 0.5 * r[0] * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
  r[4] = (u[4] - Ek) * (p->gamma0 - 1.0);
  memcpy(q, r, sizeof r);
}

static inline double speed_of_sound(const fv3d_params* p, const double q[5]) {  /* States.h:47-50 */
  return sqrt(q[4] * p->gamma0 / q[0]);
}

void orc_compute_flux(const fv3d_params* p, const double q[5], double f[5]) {  /* States.h:113-127 */
  const double Ek = 0.5 * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  const double E = (q[4] / (p->gamma0 - 1.0) + Ek);
  f[0] = q[0] * q[1];
  f[1] = q[0] * q[1] * q[1] + q[4];
  f[2] = q[0] * q[1] * q[2];
  f[3] = q[0] * q[1] * q[3];
  f[4] = (q[4] + E) * q[1];
}

void orc_hll(const fv3d_params* p, const double qL[5], const double qR[5], double flux[5],
             double* pout) {  /* RiemannSolvers.h:5-36 */
  const double aL = speed_of_sound(p, qL);
  const double aR = speed_of_sound(p, qR);
  const double sminL = qL[1] - aL;
  const double smaxL = qL[1] + aL;
  const double sminR = qR[1] - aR;
  const double smaxR = qR[1] + aR;
  const double SL = fmin(sminL, sminR);
  const double SR = fmax(smaxL, smaxR);
  double FL[5], FR[5];
  orc_compute_flux(p, qL, FL);
  orc_compute_flux(p, qR, FR);
  if (SL >= 0.0) {
    memcpy(flux, FL, sizeof FL);
    *pout = qL[4];
  } else if (SR <= 0.0) {
    memcpy(flux, FR, sizeof FR);
    *pout = qR[4];
  } else {
    double uL[5], uR[5];
    orc_state_prim_to_cons(p, qL, uL);
    orc_state_prim_to_cons(p, qR, uR);
    *pout = 0.5 * (qL[4] + qR[4]);
    /* (SR*FL - SL*FR + SL*SR*(uR-uL)) / (SR-SL), State operators elementwise */
    for (int n = 0; n < 5; ++n)
      flux[n] = (FL[n] * SR - FR[n] * SL + (uR[n] - uL[n]) * (SL * SR)) / (SR - SL);
  }
}

void orc_hllc(const fv3d_params* p, const double qL[5], const double qR[5], double flux[5],
              double* pout) {  /* RiemannSolvers.h:38-112 */
  const double rL = qL[0], uL = qL[1], vL = qL[2], pL = qL[4];
  const double rR = qR[0], uR = qR[1], vR = qR[2], pR = qR[4];
  const double entho = 1.0 / (p->gamma0 - 1.0);
  const double ekL = 0.5 * rL * (uL * uL + vL * vL);
  const double EL = ekL + pL * entho;
  const double ekR = 0.5 * rR * (uR * uR + vR * vR);
  const double ER = ekR + pR * entho;
  const double cfastL = speed_of_sound(p, qL);
  const double cfastR = speed_of_sound(p, qR);
  const double SL = fmin(uL, uR) - fmax(cfastL, cfastR);
  const double SR = fmax(uL, uR) + fmax(cfastL, cfastR);
  const double rcL = rL * (uL - SL);
  const double rcR = rR * (SR - uR);
  const double uS = (rcR * uR + rcL * uL + (pL - pR)) / (rcR + rcL);
  const double pS = (rcR * pL + rcL * pR + rcL * rcR * (uL - uR)) / (rcR + rcL);
  const double rSL = rL * (SL - uL) / (SL - uS);
  const double ESL = ((SL - uL) * EL - pL * uL + pS * uS) / (SL - uS);
  const double rSR = rR * (SR - uR) / (SR - uS);
  const double ESR = ((SR - uR) * ER - pR * uR + pS * uS) / (SR - uS);
  double st[5], E;
  if (SL > 0.0) {
    memcpy(st, qL, sizeof st); E = EL; *pout = pL;
  } else if (uS > 0.0) {
    st[0] = rSL; st[1] = uS; st[2] = qL[2]; st[3] = qL[3]; st[4] = pS; E = ESL; *pout = pS;
  } else if (SR > 0.0) {
    st[0] = rSR; st[1] = uS; st[2] = qR[2]; st[3] = qR[3]; st[4] = pS; E = ESR; *pout = pS;
  } else {
    memcpy(st, qR, sizeof st); E = ER; *pout = pR;
  }
  flux[0] = st[0] * st[1];
  flux[1] = st[0] * st[1] * st[1] + st[4];
  flux[2] = flux[0] * st[2];
  flux[3] = flux[0] * st[3];
  flux[4] = (E + st[4]) * st[1];
}

void orc_riemann_batch(const fv3d_params* p, int solver, long n, const double* qL, const double* qR,
                       double* flux, double* pout) {
  for (long s = 0; s < n; ++s) {
    if (solver == FV3D_HLL) orc_hll(p, qL + 5 * s, qR + 5 * s, flux + 5 * s, pout + s);
    else orc_hllc(p, qL + 5 * s, qR + 5 * s, flux + 5 * s, pout + s);
  }
}

/* ------------------------------------------------------------------------ */
/* Array passes                                                              */
/* ------------------------------------------------------------------------ */

void orc_cons_to_prim(const fv3d_params* p, const double* U, double* Q) {  /* SimInfo.h:313-321 */
  const long n = (long)p->Ntx * p->Nty * p->Ntz;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) orc_state_cons_to_prim(p, U + 5 * c, Q + 5 * c);
}

void orc_prim_to_cons(const fv3d_params* p, const double* Q, double* U) {  /* SimInfo.h:323-331 */
  const long n = (long)p->Ntx * p->Nty * p->Ntz;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) orc_state_prim_to_cons(p, Q + 5 * c, U + 5 * c);
}

/* BoundaryConditions.h:21-53 */
static void fill_reflecting(const fv3d_params* p, const double* Q, int i, int j, int k, int iref,
                            int jref, int kref, int dir, double q[5]) {
  int isym, jsym, ksym;
  if (dir == FV3D_IX) {
    int ipiv = (i < iref ? p->ibeg : p->iend);
    isym = 2 * ipiv - i - 1; jsym = j; ksym = k;
  } else if (dir == FV3D_IY) {
    int jpiv = (j < jref ? p->jbeg : p->jend);
    isym = i; jsym = 2 * jpiv - j - 1; ksym = k;
  } else {
    int kpiv = (k < kref ? p->kbeg : p->kend);
    isym = i; jsym = j; ksym = 2 * kpiv - k - 1;
  }
  memcpy(q, Q + IDX(p, isym, jsym, ksym), 5 * sizeof(double));
  if (dir == FV3D_IX) q[1] *= -1.0;
  else if (dir == FV3D_IY) q[2] *= -1.0;
  else q[3] *= -1.0;
}

/* BoundaryConditions.h:59-81 */
static void fill_periodic(const fv3d_params* p, const double* Q, int i, int j, int k, int dir,
                          double q[5]) {
  if (dir == FV3D_IX) { if (i < p->ibeg) i += p->Nx; else i -= p->Nx; }
  else if (dir == FV3D_IY) { if (j < p->jbeg) j += p->Ny; else j -= p->Ny; }
  else { if (k < p->kbeg) k += p->Nz; else k -= p->Nz; }
  memcpy(q, Q + IDX(p, i, j, k), 5 * sizeof(double));
}

/* BoundaryConditions.h:86-134 (z only) */
static void fill_c91(const fv3d_params* p, const double* Q, int i, int j, int k, int kref,
                     double res[5]) {
  int kpiv = (k < kref ? p->kbeg : p->kend);
  int ksym = 2 * kpiv - k - 1;
  double pos[3];
  get_pos(p, i, j, k, pos);
  const double* qref = Q + IDX(p, i, j, kref);
  const double* qsym = Q + IDX(p, i, j, ksym);
  const double grav = (double)(p->gravity != 0);  /* the reference multiplies by the bool */
  if (k < p->kbeg) {
    double dz = pos[2] - 0.5 * p->dz;
    double T = 1.0 + pos[2] * p->theta1;
    res[4] = qref[4] + dz * qref[0] * grav;
    res[0] = res[4] / T;
    res[4] = fmax(res[4], 1.0e-6);
    res[0] = fmax(res[0], 1.0e-6);
    const double dratio = qsym[0] / res[0];
    res[1] = qsym[1];
    res[2] = qsym[2];
    res[3] = -qsym[3] * dratio;
  } else {
    const double Tref = qref[4] / qref[0];
    const double dz = pos[2] - (1.0 - 0.5 * p->dz);
    const double T = Tref + dz * p->theta1;
    const double P = qref[4] + dz * qref[0] * grav;
    const double rho = P / T;
    res[4] = fmax(P, 1.0e-6);
    res[0] = fmax(rho, 1.0e-6);
    const double dratio = qsym[0] / rho;
    res[1] = qsym[1] * dratio;
    res[2] = qsym[2] * dratio;
    res[3] = -qsym[3] * dratio;
  }
}

/* BoundaryConditions.h:146-215: three ordered passes X -> Y -> Z.  Within a
 * pass every item reads only cells outside that pass's ghost slab, so the
 * loops below may run in any order. */
void orc_fill_boundaries(const fv3d_params* p, double* Q) {
  const int Ng = p->Ng;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = p->kbeg; k < p->kend; ++k)
    for (int j = p->jbeg; j < p->jend; ++j)
      for (int i = 0; i < Ng; ++i) {
        const int il = i, ir = p->iend + i;
        const int refl = p->ibeg, refr = p->iend - 1;
        double ql[5], qr[5];
        switch (p->boundary_x) {
          case FV3D_BC_ABSORBING:
            memcpy(ql, Q + IDX(p, refl, j, k), sizeof ql);
            memcpy(qr, Q + IDX(p, refr, j, k), sizeof qr);
            break;
          case FV3D_BC_REFLECTING:
            fill_reflecting(p, Q, il, j, k, refl, j, k, FV3D_IX, ql);
            fill_reflecting(p, Q, ir, j, k, refr, j, k, FV3D_IX, qr);
            break;
          default:
            fill_periodic(p, Q, il, j, k, FV3D_IX, ql);
            fill_periodic(p, Q, ir, j, k, FV3D_IX, qr);
        }
        memcpy(Q + IDX(p, il, j, k), ql, sizeof ql);
        memcpy(Q + IDX(p, ir, j, k), qr, sizeof qr);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = p->kbeg; k < p->kend; ++k)
    for (int j = 0; j < Ng; ++j)
      for (int i = 0; i < p->Ntx; ++i) {
        const int jt = j, jb = p->jend + j;
        const int reft = p->jbeg, refb = p->jend - 1;
        double qt[5], qb[5];
        switch (p->boundary_y) {
          case FV3D_BC_ABSORBING:
            memcpy(qt, Q + IDX(p, i, reft, k), sizeof qt);
            memcpy(qb, Q + IDX(p, i, refb, k), sizeof qb);
            break;
          case FV3D_BC_REFLECTING:
            fill_reflecting(p, Q, i, jt, k, i, reft, k, FV3D_IY, qt);
            fill_reflecting(p, Q, i, jb, k, i, refb, k, FV3D_IY, qb);
            break;
          default:
            fill_periodic(p, Q, i, jt, k, FV3D_IY, qt);
            fill_periodic(p, Q, i, jb, k, FV3D_IY, qb);
        }
        memcpy(Q + IDX(p, i, jt, k), qt, sizeof qt);
        memcpy(Q + IDX(p, i, jb, k), qb, sizeof qb);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < Ng; ++k)
    for (int j = 0; j < p->Nty; ++j)
      for (int i = 0; i < p->Ntx; ++i) {
        const int kf = k, kb = p->kend + k;
        const int reff = p->kbeg, refb = p->kend - 1;
        double qf[5], qb[5];
        switch (p->boundary_z) {
          case FV3D_BC_ABSORBING:
            memcpy(qf, Q + IDX(p, i, j, reff), sizeof qf);
            memcpy(qb, Q + IDX(p, i, j, refb), sizeof qb);
            break;
          case FV3D_BC_REFLECTING:
            fill_reflecting(p, Q, i, j, kf, i, j, reff, FV3D_IZ, qf);
            fill_reflecting(p, Q, i, j, kb, i, j, refb, FV3D_IZ, qb);
            break;
          case FV3D_BC_C91:
            fill_c91(p, Q, i, j, kf, reff, qf);
            fill_c91(p, Q, i, j, kb, refb, qb);
            break;
          default:
            fill_periodic(p, Q, i, j, kf, FV3D_IZ, qf);
            fill_periodic(p, Q, i, j, kb, FV3D_IZ, qb);
        }
        memcpy(Q + IDX(p, i, j, kf), qf, sizeof qf);
        memcpy(Q + IDX(p, i, j, kb), qb, sizeof qb);
      }
}

static inline double minmod(double dL, double dR) {  /* Update.h:73-80 */
  if (dL * dR < 0.0) return 0.0;
  else if (fabs(dL) < fabs(dR)) return dL;
  else return dR;
}

void orc_compute_slopes(const fv3d_params* p, const double* Q, double* sx, double* sy,
                        double* sz) {  /* Update.h:55-88, over range_slopes */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = p->kbeg - 1; k < p->kend + 1; ++k)
    for (int j = p->jbeg - 1; j < p->jend + 1; ++j)
      for (int i = p->ibeg - 1; i < p->iend + 1; ++i)
        for (int v = 0; v < 5; ++v) {
          const long c = IDX(p, i, j, k) + v;
          double dL = Q[c] - Q[IDX(p, i - 1, j, k) + v];
          double dR = Q[IDX(p, i + 1, j, k) + v] - Q[c];
          double dU = Q[c] - Q[IDX(p, i, j - 1, k) + v];
          double dD = Q[IDX(p, i, j + 1, k) + v] - Q[c];
          double dF = Q[IDX(p, i, j, k + 1) + v] - Q[c];
          double dB = Q[c] - Q[IDX(p, i, j, k - 1) + v];
          sx[c] = minmod(dL, dR);
          sy[c] = minmod(dU, dD);
          sz[c] = minmod(dF, dB);
        }
}

/* Update.h:14-33 */
static void reconstruct(const fv3d_params* p, const double* Q, const double* slopes, int i, int j,
                        int k, double sign, int dir, double res[5]) {
  const double* q = Q + IDX(p, i, j, k);
  const double* s = slopes + IDX(p, i, j, k);
  switch (p->reconstruction) {
    case FV3D_PLM:
      for (int n = 0; n < 5; ++n) res[n] = q[n] + s[n] * sign * 0.5;
      break;
    case FV3D_PCM_WB:
      res[0] = q[0]; res[1] = q[1]; res[2] = q[2]; res[3] = q[3];
      res[4] = (dir == FV3D_IX || dir == FV3D_IY ? q[4] : q[4] + sign * q[0] * p->g * p->dz * 0.5);
      break;
    default:
      memcpy(res, q, 5 * sizeof(double));
  }
  swap_component(res, dir);
}

static inline void riemann(const fv3d_params* p, const double qL[5], const double qR[5],
                           double f[5], double* pout) {
  if (p->riemann_solver == FV3D_HLL) orc_hll(p, qL, qR, f, pout);
  else orc_hllc(p, qL, qR, f, pout);
}

void orc_fluxes_and_update(const fv3d_params* p, const double* Q, const double* sx,
                           const double* sy, const double* sz, double* Unew,
                           double dt) {  /* Update.h:90-162 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = p->kbeg; k < p->kend; ++k)
    for (int j = p->jbeg; j < p->jend; ++j)
      for (int i = p->ibeg; i < p->iend; ++i) {
        for (int dir = 0; dir < 3; ++dir) {
          const double* slopes = (dir == FV3D_IX ? sx : (dir == FV3D_IY ? sy : sz));
          int dm[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
          dm[dir] = -1; dp[dir] = 1;
          double qCL[5], qCR[5], qL[5], qR[5];
          reconstruct(p, Q, slopes, i, j, k, -1.0, dir, qCL);
          reconstruct(p, Q, slopes, i, j, k, 1.0, dir, qCR);
          reconstruct(p, Q, slopes, i + dm[0], j + dm[1], k + dm[2], 1.0, dir, qL);
          reconstruct(p, Q, slopes, i + dp[0], j + dp[1], k + dp[2], -1.0, dir, qR);
          double fluxL[5], fluxR[5], poutL, poutR;
          riemann(p, qL, qCL, fluxL, &poutL);
          riemann(p, qCR, qR, fluxR, &poutR);
          swap_component(fluxL, dir);
          swap_component(fluxR, dir);
          const double rho_c = Q[IDX(p, i, j, k) + 0];
          if (p->well_balanced_flux_at_z_bc && (k == p->kbeg || k == p->kend - 1) && dir == FV3D_IZ) {
            if (k == p->kbeg) {
              fluxL[0] = 0.0; fluxL[1] = 0.0; fluxL[2] = 0.0;
              fluxL[3] = poutR - rho_c * p->g * p->dz; fluxL[4] = 0.0;
            } else {
              fluxR[0] = 0.0; fluxR[1] = 0.0; fluxR[2] = 0.0;
              fluxR[3] = poutL + rho_c * p->g * p->dz; fluxR[4] = 0.0;
            }
          } else if (dir == FV3D_IZ && (k == p->kbeg || k == p->kend - 1) &&
                     p->boundary_z == FV3D_BC_C91) {
            if (k == p->kbeg) {
              fluxL[0] = 0.0; fluxL[1] = 0.0; fluxL[2] = 0.0; fluxL[3] = poutL; fluxL[4] = 0.0;
            } else {
              fluxR[0] = 0.0; fluxR[1] = 0.0; fluxR[2] = 0.0; fluxR[3] = poutR; fluxR[4] = 0.0;
            }
          }
          double* un = Unew + IDX(p, i, j, k);
          const double d = (dir == FV3D_IX ? p->dx : (dir == FV3D_IY ? p->dy : p->dz));
          for (int n = 0; n < 5; ++n) un[n] += (fluxL[n] - fluxR[n]) * dt / d;
          if (dir == FV3D_IZ && p->gravity) {
            un[3] += dt * rho_c * p->g;
            un[4] += dt * 0.5 * (fluxL[0] + fluxR[0]) * p->g;
          }
        }
        double* un = Unew + IDX(p, i, j, k);
        un[0] = fmax(1.0e-6, un[0]);
      }
}

/* ThermalConduction.h:7-22 */
static inline double compute_kappa(const fv3d_params* p, double x, double y, double z) {
  (void)x; (void)y;
  if (p->thermal_conductivity_mode == FV3D_TCM_B02) {
    const double tr = (tanh((z - p->b02_zmid) / p->b02_thickness) + 1.0) * 0.5;
    return p->kappa * (p->b02_kappa1 * (1.0 - tr) + p->b02_kappa2 * tr);
  }
  return p->kappa;
}

void orc_thermal_conduction(const fv3d_params* p, const double* Q, double* Unew,
                            double dt) {  /* ThermalConduction.h:32-97 */
  const double dx = p->dx, dy = p->dy, dz = p->dz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = p->kbeg; k < p->kend; ++k)
    for (int j = p->jbeg; j < p->jend; ++j)
      for (int i = p->ibeg; i < p->iend; ++i) {
        double pos[3];
        get_pos(p, i, j, k, pos);
        const double x = pos[0], y = pos[1], z = pos[2];
        double kappaL = 0.5 * (compute_kappa(p, x, y, z) + compute_kappa(p, x - dx, y, z));
        double kappaR = 0.5 * (compute_kappa(p, x, y, z) + compute_kappa(p, x + dx, y, z));
        double kappaU = 0.5 * (compute_kappa(p, x, y, z) + compute_kappa(p, x, y - dy, z));
        double kappaD = 0.5 * (compute_kappa(p, x, y, z) + compute_kappa(p, x, y + dy, z));
        double kappaF = 0.5 * (compute_kappa(p, x, y, z) + compute_kappa(p, x, y, z - dz));
        double kappaB = 0.5 * (compute_kappa(p, x, y, z) + compute_kappa(p, x, y, z + dz));
#define T_AT(ii, jj, kk) (Q[IDX(p, ii, jj, kk) + 4] / Q[IDX(p, ii, jj, kk) + 0])
        double TC = T_AT(i, j, k);
        double TL = T_AT(i - 1, j, k);
        double TR = T_AT(i + 1, j, k);
        double TU = T_AT(i, j - 1, k);
        double TD = T_AT(i, j + 1, k);
        double TF = T_AT(i, j, k - 1);
        double TB = T_AT(i, j, k + 1);
#undef T_AT
        double FL = kappaL * (TC - TL) / dx;
        double FR = kappaR * (TR - TC) / dx;
        double FU = kappaU * (TC - TU) / dy;
        double FD = kappaD * (TD - TC) / dy;
        double FF = kappaF * (TC - TF) / dz;
        double FB = kappaB * (TB - TC) / dz;
        if (k == p->kbeg && p->bctc_zmin != FV3D_BCTC_NONE) {
          if (p->bctc_zmin == FV3D_BCTC_FIXED_TEMPERATURE)
            FF = kappaF * 2.0 * (TC - p->bctc_zmin_value) / dz;
          else if (p->bctc_zmin == FV3D_BCTC_FIXED_GRADIENT)
            FF = kappaF * p->bctc_zmin_value;
        }
        if (k == p->kend - 1 && p->bctc_zmax != FV3D_BCTC_NONE) {
          if (p->bctc_zmax == FV3D_BCTC_FIXED_TEMPERATURE)
            FB = kappaB * 2.0 * (p->bctc_zmax_value - TC) / dz;
          else if (p->bctc_zmax == FV3D_BCTC_FIXED_GRADIENT)
            FB = kappaB * p->bctc_zmax_value;
        }
        Unew[IDX(p, i, j, k) + 4] += dt / dx * (FR - FL) + dt / dy * (FD - FU) + dt / dz * (FB - FF);
      }
}

void orc_viscosity(const fv3d_params* p, const double* Q, double* Unew,
                   double dt) {  /* Viscosity.h:22-155 */
  const double dx = p->dx, dy = p->dy, dz = p->dz;
  const double four_thirds = 4.0 / 3.0;
  const double two_thirds = 2.0 / 3.0;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = p->kbeg; k < p->kend; ++k)
    for (int j = p->jbeg; j < p->jend; ++j)
      for (int i = p->ibeg; i < p->iend; ++i) {
        double st[3][3][3][5];  /* stencil[dk+1][dj+1][di+1] */
        for (int di = -1; di < 2; ++di)
          for (int dj = -1; dj < 2; ++dj)
            for (int dk = -1; dk < 2; ++dk)
              memcpy(st[dk + 1][dj + 1][di + 1], Q + IDX(p, i + di, j + dj, k + dk),
                     5 * sizeof(double));
        const double one_over_dx = 1.0 / dx;
        const double one_over_dy = 1.0 / dy;
        const double one_over_dz = 1.0 / dz;
        const double mu = p->mu;  /* computeMu, Viscosity.h:7-11 */
        double vf[3][5];
        for (int dir = 0; dir < 3; ++dir) {
          double flux[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
          for (int side = 1; side < 3; ++side) {
            const double sign = (side == 1 ? -1.0 : 1.0);
            if (dir == FV3D_IX) {
              double qi[5];
              for (int n = 0; n < 5; ++n) qi[n] = 0.5 * (st[1][1][side][n] + st[1][1][side - 1][n]);
              double dudx = one_over_dx * (st[1][1][side][1] - st[1][1][side - 1][1]);
              double dvdx = one_over_dx * (st[1][1][side][2] - st[1][1][side - 1][2]);
              double dwdx = one_over_dx * (st[1][1][side][3] - st[1][1][side - 1][3]);
              double dudy = 0.25 * one_over_dy * (st[1][2][side][1] - st[1][0][side][1]
                                                + st[1][2][side - 1][1] - st[1][0][side - 1][1]);
              double dvdy = 0.25 * one_over_dy * (st[1][2][side][2] - st[1][0][side][2]
                                                + st[1][2][side - 1][2] - st[1][0][side - 1][2]);
              double dudz = 0.25 * one_over_dz * (st[2][1][side][1] - st[0][1][side][1]
                                                + st[2][1][side - 1][1] - st[0][1][side - 1][1]);
              double dwdz = 0.25 * one_over_dz * (st[2][1][side][3] - st[0][1][side][3]
                                                + st[2][1][side - 1][3] - st[0][1][side - 1][3]);
              const double tau_xx = four_thirds * dudx - two_thirds * (dvdy + dwdz);
              const double tau_xy = dvdx + dudy;
              const double tau_xz = dwdx + dudz;
              flux[1] += sign * mu * tau_xx;
              flux[2] += sign * mu * tau_xy;
              flux[3] += sign * mu * tau_xz;
              flux[4] += sign * mu * (tau_xx * qi[1] + tau_xy * qi[2] + tau_xz * qi[3]);
            } else if (dir == FV3D_IY) {
              double qi[5];
              for (int n = 0; n < 5; ++n) qi[n] = 0.5 * (st[1][side][1][n] + st[1][side - 1][1][n]);
              double dudy = one_over_dy * (st[1][side][1][1] - st[1][side - 1][1][1]);
              double dvdy = one_over_dy * (st[1][side][1][2] - st[1][side - 1][1][2]);
              double dwdy = one_over_dy * (st[1][side][1][3] - st[1][side - 1][1][3]);
              double dudx = 0.25 * one_over_dx * (st[1][side][2][1] - st[1][side][0][1]
                                                + st[1][side - 1][2][1] - st[1][side - 1][0][1]);
              double dvdx = 0.25 * one_over_dx * (st[1][side][2][2] - st[1][side][0][2]
                                                + st[1][side - 1][2][2] - st[1][side - 1][0][2]);
              double dwdz = 0.25 * one_over_dz * (st[2][side][1][3] - st[0][side][1][3]
                                                + st[2][side - 1][1][3] - st[0][side - 1][1][3]);
              double dvdz = 0.25 * one_over_dz * (st[2][side][1][2] - st[0][side][1][2]
                                                + st[2][side - 1][1][2] - st[0][side - 1][1][2]);
              const double tau_yy = four_thirds * dvdy - two_thirds * (dudx + dwdz);
              const double tau_xy = dvdx + dudy;
              const double tau_yz = dvdz + dwdy;
              flux[1] += sign * mu * tau_xy;
              flux[2] += sign * mu * tau_yy;
              flux[3] += sign * mu * tau_yz;
              flux[4] += sign * mu * (tau_xy * qi[1] + tau_yy * qi[2] + tau_yz * qi[3]);
            } else {
              double qi[5];
              for (int n = 0; n < 5; ++n) qi[n] = 0.5 * (st[side][1][1][n] + st[side - 1][1][1][n]);
              double dudz = one_over_dz * (st[side][1][1][1] - st[side - 1][1][1][1]);
              double dvdz = one_over_dz * (st[side][1][1][2] - st[side - 1][1][1][2]);
              double dwdz = one_over_dz * (st[side][1][1][3] - st[side - 1][1][1][3]);
              /* index and component mix-ups kept as in Viscosity.h:122-133 */
              double dudx = 0.25 * one_over_dx * (st[side][1][2][1] - st[1][side][0][1]
                                                + st[side - 1][1][2][1] - st[1][side - 1][0][1]);
              double dwdx = 0.25 * one_over_dx * (st[side][1][2][2] - st[1][side][0][3]
                                                + st[side - 1][1][2][2] - st[1][side - 1][0][3]);
              double dvdy = 0.25 * one_over_dz * (st[side][2][1][3] - st[side][0][1][2]
                                                + st[side - 1][2][1][3] - st[side - 1][0][1][2]);
              double dwdy = 0.25 * one_over_dz * (st[side][2][1][2] - st[side][0][1][3]
                                                + st[side - 1][2][1][2] - st[side - 1][0][1][3]);
              const double tau_zz = four_thirds * dwdz - two_thirds * (dudx + dvdy);
              const double tau_xz = dudz + dwdx;
              const double tau_yz = dvdz + dwdy;
              flux[1] += sign * mu * tau_xz;
              flux[2] += sign * mu * tau_yz;
              flux[3] += sign * mu * tau_zz;
              flux[4] += sign * mu * (tau_xz * qi[1] + tau_yz * qi[2] + tau_zz * qi[3]);
            }
          }
          memcpy(vf[dir], flux, sizeof flux);
        }
        double* un = Unew + IDX(p, i, j, k);
        for (int n = 0; n < 5; ++n) un[n] += (vf[0][n] + vf[1][n] + vf[2][n]) * dt;
      }
}

long orc_work_size(const fv3d_params* p) { return 5L * p->Ntx * p->Nty * p->Ntz * 5; }

void orc_euler_step(const fv3d_params* p, double* Q, double* Unew, double dt,
                    double* work) {  /* Update.h:164-178 */
  const long n = (long)p->Ntx * p->Nty * p->Ntz * 5;
  double *sx = work, *sy = work + n, *sz = work + 2 * n;
  orc_fill_boundaries(p, Q);
  if (p->reconstruction == FV3D_PLM) orc_compute_slopes(p, Q, sx, sy, sz);
  orc_fluxes_and_update(p, Q, sx, sy, sz, Unew, dt);
  if (p->thermal_conductivity_active) orc_thermal_conduction(p, Q, Unew, dt);
  if (p->viscosity_active) orc_viscosity(p, Q, Unew, dt);
}

void orc_update(const fv3d_params* p, double* Q, double* Unew, double dt,
                double* work) {  /* Update.h:180-206 */
  if (p->time_stepping == FV3D_TS_EULER) {
    orc_euler_step(p, Q, Unew, dt, work);
    return;
  }
  const long n = (long)p->Ntx * p->Nty * p->Ntz * 5;
  double *U0 = work + 3 * n, *Ustar = work + 4 * n;
  memcpy(U0, Unew, n * sizeof(double));
  memcpy(Ustar, Unew, n * sizeof(double));
  orc_euler_step(p, Q, Ustar, dt, work);
  memcpy(Unew, Ustar, n * sizeof(double));
  orc_cons_to_prim(p, Ustar, Q);
  orc_euler_step(p, Q, Unew, dt, work);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = p->kbeg; k < p->kend; ++k)
    for (int j = p->jbeg; j < p->jend; ++j)
      for (int i = p->ibeg; i < p->iend; ++i)
        for (int v = 0; v < 5; ++v) {
          const long c = IDX(p, i, j, k) + v;
          Unew[c] = 0.5 * (U0[c] + Unew[c]);
        }
}

double orc_compute_dt(const fv3d_params* p, const double* Q, double inv[3]) {  /* ComputeDt.h:17-63 */
  double ih = -1.0e308, itc = -1.0e308, ivi = -1.0e308;
#pragma omp parallel
  {
    double lh = -1.0e308, ltc = -1.0e308, lvi = -1.0e308;
#pragma omp for collapse(2) schedule(static) nowait
    for (int k = p->kbeg; k < p->kend; ++k)
      for (int j = p->jbeg; j < p->jend; ++j)
        for (int i = p->ibeg; i < p->iend; ++i) {
          const double* q = Q + IDX(p, i, j, k);
          double cs = speed_of_sound(p, q);
          double h = (cs + fabs(q[1])) / p->dx + (cs + fabs(q[2])) / p->dy + (cs + fabs(q[3])) / p->dz;
          double tc = p->epsilon;
          if (p->thermal_conductivity_active) {
            /* computeKappa receives the integer indices here (ComputeDt.h:39-41) */
            double kap = compute_kappa(p, (double)i, (double)j, (double)k);
            tc = fmax(fmax(2.0 * kap / (p->dx * p->dx), 2.0 * kap / (p->dy * p->dy)),
                      2.0 * kap / (p->dz * p->dz));
          }
          double vi = p->epsilon;
          if (p->viscosity_active)
            vi = fmax(fmax(2.0 * p->mu / (p->dx * p->dx), 2.0 * p->mu / (p->dy * p->dy)),
                      2.0 * p->mu / (p->dz * p->dz));
          lh = fmax(lh, h);
          ltc = fmax(ltc, tc);
          lvi = fmax(lvi, vi);
        }
#pragma omp critical
    {
      ih = fmax(ih, lh);
      itc = fmax(itc, ltc);
      ivi = fmax(ivi, lvi);
    }
  }
  if (inv) { inv[0] = ih; inv[1] = itc; inv[2] = ivi; }
  double m = ih;
  if (itc > m) m = itc;
  if (ivi > m) m = ivi;
  return p->CFL / m;
}

/* ------------------------------------------------------------------------ */
/* Problem setup (Init.h).  H84 and C91 draw their perturbation from a       */
/* counter-based generator keyed by (seed, cell) instead of Kokkos'          */
/* Random_XorShift64_Pool, whose draws depend on the thread mapping.         */
/* ------------------------------------------------------------------------ */

static inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

static inline double cell_uniform(int seed, long cell, double lo, double hi) {
  uint64_t r = splitmix64(((uint64_t)(uint32_t)seed << 40) ^ (uint64_t)cell);
  double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
  return lo + (hi - lo) * u;
}

int orc_init(const fv3d_params* p, int problem, double* Q) {
  if (problem < ORC_SOD_X || problem > ORC_C91) return -1;
  for (int k = p->kbeg; k < p->kend; ++k)
    for (int j = p->jbeg; j < p->jend; ++j)
      for (int i = p->ibeg; i < p->iend; ++i) {
        double* q = Q + IDX(p, i, j, k);
        double pos[3];
        get_pos(p, i, j, k, pos);
        const double x = pos[0], y = pos[1], z = pos[2];
        switch (problem) {
          case ORC_SOD_X: case ORC_SOD_Y: case ORC_SOD_Z: {  /* Init.h:18-67 */
            const double c = pos[problem - ORC_SOD_X];
            if (c <= 0.5) { q[0] = 1.0; q[4] = 1.0; }
            else { q[0] = 0.125; q[4] = 0.1; }
            q[1] = 0.0; q[2] = 0.0; q[3] = 0.0;
            break;
          }
          case ORC_BLAST: {  /* Init.h:72-99 */
            double xmid = 0.5 * (p->xmin + p->xmax);
            double ymid = 0.5 * (p->ymin + p->ymax);
            double zmid = 0.5 * (p->zmin + p->zmax);
            double xr = xmid - x, yr = ymid - y, zr = zmid - z;
            double r = sqrt(xr * xr + yr * yr + zr * zr);
            if (r < 0.2) { q[0] = 1.0; q[4] = 10.0; }
            else { q[0] = 1.2; q[4] = 0.1; }
            q[1] = 0.0; q[2] = 0.0; q[3] = 0.0;
            break;
          }
          case ORC_DIFFUSION: {  /* Init.h:155-178 */
            double xmid = 0.5 * (p->xmin + p->xmax);
            double ymid = 0.5 * (p->ymin + p->ymax);
            double zmid = 0.5 * (p->zmin + p->zmax);
            double x0 = x - xmid, y0 = y - ymid, z0 = z - zmid;
            double r = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
            q[0] = (r < 0.2) ? 1.0 : 0.1;
            q[4] = 1.0; q[1] = 1.0; q[2] = 1.0; q[3] = 1.0;
            break;
          }
          case ORC_RAYLEIGH_TAYLOR: {  /* Init.h:183-209 */
            double zmid = 0.5 * (p->zmin + p->zmax);
            const double P0 = 2.5;
            q[0] = (z < zmid) ? 1.0 : 2.0;
            q[1] = 0.0; q[2] = 0.0; q[3] = 0.0;
            q[4] = P0 + 0.1 * p->g * z;
            if (z > -1.0 / 3.0 && z < 1.0 / 3.0)
              q[3] = 0.01 * (1.0 + cos(4 * M_PI * x)) * (1.0 + cos(4.0 * M_PI * y)) *
                     (1 + cos(3.0 * M_PI * z)) / 4.0;
            /* outside that band w keeps its allocation value, zero */
            break;
          }
          case ORC_H84: {  /* Init.h:104-123 */
            const long cell = ((long)k * p->Nty + j) * p->Ntx + i;
            q[0] = pow(z, p->m1);
            q[1] = 0.0; q[2] = 0.0;
            q[3] = p->h84_pert * cell_uniform(p->seed, cell, -0.5, 0.5);
            q[4] = pow(z, p->m1 + 1.0);
            break;
          }
          case ORC_C91: {  /* Init.h:128-150 */
            const long cell = ((long)k * p->Nty + j) * p->Ntx + i;
            double T = (1.0 + p->theta1 * z);
            double rho = pow(T, p->m1);
            double prs = pow(T, p->m1 + 1.0);
            double pert = p->c91_pert * cell_uniform(p->seed, cell, -0.5, 0.5);
            prs = prs * (1.0 + pert);
            q[0] = rho; q[1] = 0.0; q[2] = 0.0; q[3] = 0.0; q[4] = prs;
            break;
          }
        }
      }
  orc_fill_boundaries(p, Q);
  return 0;
}

/* main.cpp:51-74 without the snapshot calls */
long orc_run(const fv3d_params* p, double* U, double* Q, double* Unew, double* work, long max_steps,
             int stop_at_tend, double* dts, double* t_out) {
  const long n = (long)p->Ntx * p->Nty * p->Ntz * 5;
  double t = 0.0;
  long s = 0;
  while (s < max_steps && (!stop_at_tend || t + p->epsilon < p->tend)) {
    memcpy(Unew, U, n * sizeof(double));
    orc_cons_to_prim(p, U, Q);
    double dt = orc_compute_dt(p, Q, NULL);
    orc_update(p, Q, Unew, dt, work);
    memcpy(U, Unew, n * sizeof(double));
    t += dt;
    if (dts) dts[s] = dt;
    ++s;
  }
  if (t_out) *t_out = t;
  return s;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
