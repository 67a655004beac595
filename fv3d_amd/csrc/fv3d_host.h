// fv3d_host.h -- host-side helpers shared by the single-box and batch C ABI:
// parameter checks, kernel parameters, and the host layout transposes.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fv3d_hip.h"
#include "fv3d_kernels.h"

namespace fv3d_hip {

// records msg as the library-wide last error (fv3d_last_error(NULL)); returns code
int record_error(int code, const std::string& msg);

inline bool is_pow2(double d) {
  if (!(d > 0.0) || !std::isfinite(d)) return false;
  int e;
  return std::frexp(d, &e) == 0.5;
}

// inv_dt_par for a constant coefficient (ComputeDt.h:37-47)
inline double inv_dt_par(bool active, double coef, const fv3d_params& p) {
  if (!active) return p.epsilon;
  return std::fmax(std::fmax(2.0 * coef / (p.dx * p.dx), 2.0 * coef / (p.dy * p.dy)),
                   2.0 * coef / (p.dz * p.dz));
}

// z-chunk length of the stage kernel.  One stage workgroup fits on a CU
// (LDS and VGPRs), so tiles_xy * tz workgroups run in
// ceil(tiles_xy*tz/cus) rounds (256 CUs on MI355X), each marching a chunk of
// nzc planes after a warm-up worth about 1.5 planes (measured on the 256^3
// blast: 1, 2 and 4 rounds of 128, 64 and 32 planes took 1.588, 1.607 and
// 1.630 ms per step).  The chunk count minimises rounds x (nzc + 1.5);
// FV3D_NZC overrides.
inline int tiles_z_for(const fv3d_params& p, int tiles_xy, int cus) {
  int nzc = 0;
  if (const char* e = std::getenv("FV3D_NZC")) nzc = std::atoi(e);
  if (nzc <= 0) {
    const int kCUs = cus > 0 ? cus : 256;
    double best = 0.0;
    for (int tz = 1; tz <= std::min(p.Nz, 64); ++tz) {
      const int n = (p.Nz + tz - 1) / tz;
      if (tz > 1 && (p.Nz + n - 1) / n < tz) continue;  // the same chunks as a smaller tz
      const long long rounds = ((long long)tiles_xy * tz + kCUs - 1) / kCUs;
      const double cost = (double)rounds * (n + 1.5);
      if (nzc == 0 || cost < best) { best = cost; nzc = n; }
    }
  }
  return std::max(1, std::min(nzc, p.Nz));
}

// computeKappa (ThermalConduction.h:7-22), evaluated on the host with the
// host's tanh as the reference's OpenMP build does
inline double compute_kappa(const fv3d_params& p, double z) {
  if (p.thermal_conductivity_mode == FV3D_TCM_B02) {
    const double tr = (std::tanh((z - p.b02_zmid) / p.b02_thickness) + 1.0) * 0.5;
    return p.kappa * (p.b02_kappa1 * (1.0 - tr) + p.b02_kappa2 * tr);
  }
  return p.kappa;
}

// B02 face conductivities per z level, [3][Nz]: the x/y faces (kappaL..D),
// the lower and the upper z face (kappaF, kappaB), ThermalConduction.h:44-49
inline std::vector<double> b02_face_table(const fv3d_params& p) {
  std::vector<double> t(3 * (size_t)p.Nz);
  for (int k = 0; k < p.Nz; ++k) {
    const double z = p.zmin + (k + 0.5) * p.dz;  // getPos, SimInfo.h:162-167
    const double kc = compute_kappa(p, z);
    t[k] = 0.5 * (kc + compute_kappa(p, z));
    t[p.Nz + k] = 0.5 * (kc + compute_kappa(p, z - p.dz));
    t[2 * p.Nz + k] = 0.5 * (kc + compute_kappa(p, z + p.dz));
  }
  return t;
}

inline int check_params(const fv3d_params& p, std::string& why) {
  if (p.Nx < 2 || p.Ny < 2 || p.Nz < 2) { why = "Nx, Ny, Nz must be >= 2"; return 1; }
  if (p.Ng < (p.reconstruction == FV3D_PLM ? 2 : 1)) { why = "too few ghost layers"; return 1; }
  if (p.thermal_conductivity_active && p.thermal_conductivity_mode == FV3D_TCM_B02 &&
      !(p.b02_thickness != 0.0))
    { why = "conductivity_mode=B02 needs a non-zero b02 thickness"; return 1; }
  if (p.riemann_solver != FV3D_HLL && p.riemann_solver != FV3D_HLLC) { why = "riemann_solver"; return 1; }
  if (p.reconstruction < FV3D_PCM || p.reconstruction > FV3D_PLM) { why = "reconstruction"; return 1; }
  if (!(p.dx > 0 && p.dy > 0 && p.dz > 0)) { why = "non-positive cell size"; return 1; }
  return 0;
}

// the compute units of a device (256 on MI355X; fewer on a partitioned one)
inline int device_cus(int device) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      cus < 1)
    cus = 256;
  return cus;
}

// Cells of the stage tiles that cover a box of tile shape (wx, ty),
// overhang included.
inline long long padded_cells(const fv3d_params& p, int wx, int ty) {
  return (long long)((p.Nx + wx - 1) / wx) * wx * ((p.Ny + ty - 1) / ty) * ty * p.Nz;
}

// Stage tile width for a set of boxes run with one kernel (one box, or an
// ensemble): 32 columns when 64-column tiles would cover at least 10% more
// cells (overhang lanes and rows do no work), else 64.  At equal cover 64
// columns run faster: 256^3 blast FP64 11.29 against 10.93 Gcell/s, FP32
// compensated 26.1 against 24.7; and 64 x 4 is 4% faster than 32 x 12 with
// viscosity in FP64.  FV3D_TILE_WX=64 or 32 overrides.
inline int stage_wx_for(const fv3d_params* ps, int n, bool fp64) {
  if (const char* e = std::getenv("FV3D_TILE_WX")) {
    const int w = std::atoi(e);
    if (w == 32 || w == 64) return w;
  }
  long long c64 = 0, c32 = 0;
  for (int i = 0; i < n; ++i) {
    const bool visc = ps[i].viscosity_active != 0;
    c64 += padded_cells(ps[i], 64, stage_ty(visc, fp64, 64));
    c32 += padded_cells(ps[i], 32, stage_ty(visc, fp64, 32));
  }
  return (double)c64 >= 1.10 * (double)c32 ? 32 : 64;
}

// kernel parameters of a box for FP64 or FP32 state; cus sizes the z
// chunks (tiles_z_for); wx: stage tile width (64 or 32; 0: stage_wx_for)
inline DevParams make_devparams(const fv3d_params& p, int nzc_override, int cus, bool fp64,
                                int wx = 0) {
  DevParams d;
  std::memset(&d, 0, sizeof d);
  d.Nx = p.Nx; d.Ny = p.Ny; d.Nz = p.Nz;
  d.sy = p.Nx;
  d.sz = (long long)p.Nx * p.Ny;
  d.sf = (long long)p.Nx * p.Ny * p.Nz;
  d.bcx = p.boundary_x; d.bcy = p.boundary_y; d.bcz = p.boundary_z;
  d.pcm_wb = p.reconstruction == FV3D_PCM_WB;
  d.wb_flux_zbc = p.well_balanced_flux_at_z_bc != 0;
  d.c91_flux_zbc = p.boundary_z == FV3D_BC_C91;
  d.gravity = p.gravity != 0;
  d.gamma = p.gamma0;
  d.gm1 = p.gamma0 - 1.0;
  d.entho = 1.0 / (p.gamma0 - 1.0);
  d.dx = p.dx; d.dy = p.dy; d.dz = p.dz;
  d.idx = 1.0 / p.dx; d.idy = 1.0 / p.dy; d.idz = 1.0 / p.dz;
  d.pow2x = is_pow2(p.dx); d.pow2y = is_pow2(p.dy); d.pow2z = is_pow2(p.dz);
  d.g = p.g;
  d.kface = 0.5 * (p.kappa + p.kappa);
  d.bctc_zmin = p.bctc_zmin; d.bctc_zmax = p.bctc_zmax;
  d.bctc_zmin_value = p.bctc_zmin_value; d.bctc_zmax_value = p.bctc_zmax_value;
  d.mu = p.mu;
  d.CFL = p.CFL;
  d.inv_tc = inv_dt_par(p.thermal_conductivity_active != 0, p.kappa, p);
  if (p.thermal_conductivity_active && p.thermal_conductivity_mode == FV3D_TCM_B02) {
    // ComputeDt.h:39-41 passes the ghosted integer indices as coordinates
    double m = 0.0;
    for (int kg = p.kbeg; kg < p.kend; ++kg) {
      const double kap = compute_kappa(p, (double)kg);
      const double v = std::fmax(std::fmax(2.0 * kap / (p.dx * p.dx), 2.0 * kap / (p.dy * p.dy)),
                                 2.0 * kap / (p.dz * p.dz));
      if (kg == p.kbeg || v > m) m = v;
    }
    d.inv_tc = m;
  }
  d.kztab = nullptr;  // set by the owner of the table
  d.inv_visc = inv_dt_par(p.viscosity_active != 0, p.mu, p);
  d.zmin = p.zmin;
  d.theta1 = p.theta1;
  d.c91g = nullptr;  // set by the owner of the ghost-plane buffer
  d.twx = (wx == 32 || wx == 64) ? wx : stage_wx_for(&p, 1, fp64);
  d.tty = stage_ty(p.viscosity_active != 0, fp64, d.twx);
  d.tiles_x = (p.Nx + d.twx - 1) / d.twx;
  d.tiles_y = (p.Ny + d.tty - 1) / d.tty;
  d.nzc = nzc_override > 0 ? std::max(1, std::min(nzc_override, p.Nz))
                           : tiles_z_for(p, d.tiles_x * d.tiles_y, cus);
  d.tiles_z = (p.Nz + d.nzc - 1) / d.nzc;
  return d;
}

// ---- host layout helpers --------------------------------------------------
// host AoS (k, j, i, var) with or without ghosts  <->  SoA domain [5][Nz][Ny][Nx]
inline void aos_to_soa_domain(const fv3d_params& p, const double* src, int with_ghosts,
                              std::vector<double>& dst) {
  const long long n = (long long)p.Nx * p.Ny * p.Nz;
  dst.resize(5 * n);
  const int g = with_ghosts ? p.Ng : 0;
  const long long tx = p.Nx + 2 * g, ty = p.Ny + 2 * g;
  for (int k = 0; k < p.Nz; ++k)
    for (int j = 0; j < p.Ny; ++j)
      for (int i = 0; i < p.Nx; ++i) {
        const double* s = src + (((long long)(k + g) * ty + (j + g)) * tx + (i + g)) * 5;
        const long long d = i + (long long)p.Nx * (j + (long long)p.Ny * k);
        for (int f = 0; f < 5; ++f) dst[f * n + d] = s[f];
      }
}

inline void soa_domain_to_aos(const fv3d_params& p, const double* src, int with_ghosts,
                              double* dst) {
  const long long n = (long long)p.Nx * p.Ny * p.Nz;
  const int g = with_ghosts ? p.Ng : 0;
  const long long tx = p.Nx + 2 * g, ty = p.Ny + 2 * g;
  for (int k = 0; k < p.Nz; ++k)
    for (int j = 0; j < p.Ny; ++j)
      for (int i = 0; i < p.Nx; ++i) {
        double* d = dst + (((long long)(k + g) * ty + (j + g)) * tx + (i + g)) * 5;
        const long long s = i + (long long)p.Nx * (j + (long long)p.Ny * k);
        for (int f = 0; f < 5; ++f) d[f] = src[f * n + s];
      }
}

// ghosted SoA [5][Ntz][Nty][Ntx]  <->  ghosted AoS (Ntz, Nty, Ntx, 5)
inline void soa_ghosted_to_aos(const fv3d_params& p, const double* src, double* dst) {
  const long long nt = (long long)p.Ntx * p.Nty * p.Ntz;
  for (long long c = 0; c < nt; ++c)
    for (int f = 0; f < 5; ++f) dst[c * 5 + f] = src[f * nt + c];
}
inline void aos_to_soa_ghosted(const fv3d_params& p, const double* src, double* dst) {
  const long long nt = (long long)p.Ntx * p.Nty * p.Ntz;
  for (long long c = 0; c < nt; ++c)
    for (int f = 0; f < 5; ++f) dst[f * nt + c] = src[c * 5 + f];
}

}  // namespace fv3d_hip
