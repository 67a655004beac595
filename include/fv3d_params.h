/*
 * fv3d_params.h -- plain-C mirror of fv3d's run parameters.
 *
 * Replaces `struct fv3d::Params` (SimInfo.h:72-159) at the C-ABI boundary:
 * every numeric member keeps its name and meaning, enums become ints with the
 * reference's enumerator values (SimInfo.h:17-69), and the two strings
 * (`problem`, `filename_out`) stay on the host side of the boundary.
 * The derived members (Ntx.., ibeg.., dx..) are filled by
 * fv3d_params_finalize(), which restates readInifile's arithmetic
 * (SimInfo.h:186-198).
 */
#ifndef FV3D_PARAMS_H
#define FV3D_PARAMS_H

#ifdef __cplusplus
extern "C" {
#endif

/* IVar (SimInfo.h:23-30): primitive slots; IE shares slot 4 with IP. */
enum { FV3D_IR = 0, FV3D_IU = 1, FV3D_IV = 2, FV3D_IW = 3, FV3D_IP = 4, FV3D_IE = 4 };
enum { FV3D_NFIELDS = 5 };
/* IDir (SimInfo.h:17-21) */
enum { FV3D_IX = 0, FV3D_IY = 1, FV3D_IZ = 2 };
/* RiemannSolver (SimInfo.h:32-35) */
enum { FV3D_HLL = 0, FV3D_HLLC = 1 };
/* BoundaryType (SimInfo.h:37-42) */
enum { FV3D_BC_ABSORBING = 0, FV3D_BC_REFLECTING = 1, FV3D_BC_PERIODIC = 2, FV3D_BC_C91 = 3 };
/* TimeStepping (SimInfo.h:44-47) */
enum { FV3D_TS_EULER = 0, FV3D_TS_RK2 = 1 };
/* ReconstructionType (SimInfo.h:49-53) */
enum { FV3D_PCM = 0, FV3D_PCM_WB = 1, FV3D_PLM = 2 };
/* ThermalConductivityMode (SimInfo.h:55-58) */
enum { FV3D_TCM_CONSTANT = 0, FV3D_TCM_B02 = 1 };
/* BCTC_Mode (SimInfo.h:61-65) */
enum { FV3D_BCTC_NONE = 0, FV3D_BCTC_FIXED_TEMPERATURE = 1, FV3D_BCTC_FIXED_GRADIENT = 2 };
/* ViscosityMode (SimInfo.h:67-69) */
enum { FV3D_VSC_CONSTANT = 0 };

typedef struct fv3d_params {
  /* mesh (SimInfo.h:92-114) */
  int Nx, Ny, Nz, Ng;
  int Ntx, Nty, Ntz;
  int ibeg, iend, jbeg, jend, kbeg, kend;
  double xmin, xmax, ymin, ymax, zmin, zmax;
  double dx, dy, dz;

  /* run (SimInfo.h:73-82) */
  double save_freq, tend;
  int boundary_x, boundary_y, boundary_z;
  int reconstruction;
  int riemann_solver;
  int time_stepping;
  double CFL;

  /* run and physics (SimInfo.h:116-123) */
  double epsilon;
  double gamma0;
  int gravity;
  double g;
  int well_balanced_flux_at_z_bc;
  int well_balanced;

  /* thermal conduction (SimInfo.h:125-131) */
  int thermal_conductivity_active;
  int thermal_conductivity_mode;
  double kappa;
  int bctc_zmin, bctc_zmax;
  double bctc_zmin_value, bctc_zmax_value;

  /* viscosity (SimInfo.h:133-136) */
  int viscosity_active;
  int viscosity_mode;
  double mu;

  /* polytropes, H84, C91, B02 (SimInfo.h:138-154) */
  double m1, theta1, m2, theta2;
  double h84_pert;
  double c91_pert;
  double b02_zmid, b02_kappa1, b02_kappa2, b02_thickness;

  /* misc (SimInfo.h:156-158) */
  int seed;
  int log_frequency;
} fv3d_params;

/* Fills the defaults readInifile applies when a key is absent
 * (SimInfo.h:175-294) and then calls fv3d_params_finalize. */
static inline void fv3d_params_defaults(fv3d_params* p) {
  fv3d_params z = {0};
  *p = z;
  p->Nx = 32; p->Ny = 32; p->Nz = 32; p->Ng = 2;
  p->xmin = 0.0; p->xmax = 1.0; p->ymin = 0.0; p->ymax = 1.0; p->zmin = 0.0; p->zmax = 1.0;
  p->tend = 1.0; p->save_freq = 1.0e-1;
  p->boundary_x = p->boundary_y = p->boundary_z = FV3D_BC_REFLECTING;
  p->reconstruction = FV3D_PCM;
  p->riemann_solver = FV3D_HLLC;
  p->time_stepping = FV3D_TS_EULER;
  p->CFL = 0.8;
  p->epsilon = 1.0e-6;
  p->gamma0 = 5.0 / 3.0;
  p->gravity = 0; p->g = 0.0;
  p->m1 = 1.0; p->theta1 = 10.0; p->m2 = 1.0; p->theta2 = 10.0;
  p->well_balanced_flux_at_z_bc = 0;
  p->thermal_conductivity_active = 0;
  p->thermal_conductivity_mode = FV3D_TCM_CONSTANT;
  p->kappa = 0.0;
  p->bctc_zmin = p->bctc_zmax = FV3D_BCTC_NONE;
  p->bctc_zmin_value = p->bctc_zmax_value = 1.0;
  p->viscosity_active = 0;
  p->viscosity_mode = FV3D_VSC_CONSTANT;
  p->mu = 0.0;
  p->h84_pert = 1.0e-4;
  p->c91_pert = 1.0e-3;
  p->seed = 12345;
  p->log_frequency = 10;
}

/* Derived mesh members, as readInifile computes them (SimInfo.h:186-198). */
static inline void fv3d_params_finalize(fv3d_params* p) {
  p->Ntx = p->Nx + 2 * p->Ng;
  p->Nty = p->Ny + 2 * p->Ng;
  p->Ntz = p->Nz + 2 * p->Ng;
  p->ibeg = p->Ng; p->iend = p->Ng + p->Nx;
  p->jbeg = p->Ng; p->jend = p->Ng + p->Ny;
  p->kbeg = p->Ng; p->kend = p->Ng + p->Nz;
  p->dx = (p->xmax - p->xmin) / p->Nx;
  p->dy = (p->ymax - p->ymin) / p->Ny;
  p->dz = (p->zmax - p->zmin) / p->Nz;
}

#ifdef __cplusplus
}
#endif

#endif /* FV3D_PARAMS_H */
