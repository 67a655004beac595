// init.cpp -- host problem generators (Init.h:18-279).
//
// The reference's problems keep their expressions, so Sod, blast, diffusion
// and Rayleigh-Taylor initial conditions equal the reference's.  H84 and
// C91 draw their perturbation from a counter-based generator keyed by
// (seed, cell) -- splitmix64 -- instead of Kokkos' Random_XorShift64_Pool,
// whose draws depend on the thread mapping (Init.h:114-116, 139-141).
#include "init.h"

#include <cmath>
#include <cstdint>
#include <map>
#include <stdexcept>

namespace fv3d {

static inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

static inline double cell_uniform(int seed, long cell, double lo, double hi) {
  const uint64_t r = splitmix64(((uint64_t)(seed & 0xFFFFFF) << 40) ^ (uint64_t)cell);
  const double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
  return lo + (hi - lo) * u;
}

void initialize(const RunConfig& cfg, HostArray& Q) {
  const fv3d_params& p = cfg.params;
  static const std::map<std::string, int> init_map{{"sod_x", 0}, {"sod_y", 1}, {"sod_z", 2},
                                                   {"blast", 3}, {"rayleigh-taylor", 4},
                                                   {"diffusion", 5}, {"H84", 6}, {"C91", 7}};
  auto it = init_map.find(cfg.problem);
  if (it == init_map.end()) throw std::runtime_error("Error unknown problem " + cfg.problem);
  const int kind = it->second;
  Q.assign((size_t)p.Ntx * p.Nty * p.Ntz * 5, 0.0);
  for (int k = p.kbeg; k < p.kend; ++k)
    for (int j = p.jbeg; j < p.jend; ++j)
      for (int i = p.ibeg; i < p.iend; ++i) {
        const long cell = ((long)k * p.Nty + j) * p.Ntx + i;
        double* q = &Q[(size_t)cell * 5];
        const double x = p.xmin + (i - p.ibeg + 0.5) * p.dx;  // getPos, SimInfo.h:162-167
        const double y = p.ymin + (j - p.jbeg + 0.5) * p.dy;
        const double z = p.zmin + (k - p.kbeg + 0.5) * p.dz;
        switch (kind) {
          case 0: case 1: case 2: {
            const double c = kind == 0 ? x : (kind == 1 ? y : z);
            q[0] = c <= 0.5 ? 1.0 : 0.125;
            q[4] = c <= 0.5 ? 1.0 : 0.1;
            break;
          }
          case 3: {
            const double xr = 0.5 * (p.xmin + p.xmax) - x;
            const double yr = 0.5 * (p.ymin + p.ymax) - y;
            const double zr = 0.5 * (p.zmin + p.zmax) - z;
            const double r = std::sqrt(xr * xr + yr * yr + zr * zr);
            const bool in = r < cfg.blast_radius;
            q[0] = in ? cfg.blast_rho_in : cfg.blast_rho_out;
            q[4] = in ? cfg.blast_p_in : cfg.blast_p_out;
            break;
          }
          case 4: {
            const double zmid = 0.5 * (p.zmin + p.zmax);
            q[0] = z < zmid ? 1.0 : 2.0;
            q[4] = 2.5 + 0.1 * p.g * z;
            if (z > -1.0 / 3.0 && z < 1.0 / 3.0)
              q[3] = 0.01 * (1.0 + std::cos(4 * M_PI * x)) * (1.0 + std::cos(4.0 * M_PI * y)) *
                     (1 + std::cos(3.0 * M_PI * z)) / 4.0;
            break;
          }
          case 5: {
            const double x0 = x - 0.5 * (p.xmin + p.xmax);
            const double y0 = y - 0.5 * (p.ymin + p.ymax);
            const double z0 = z - 0.5 * (p.zmin + p.zmax);
            const double r = std::sqrt(x0 * x0 + y0 * y0 + z0 * z0);
            q[0] = r < 0.2 ? 1.0 : 0.1;
            q[1] = q[2] = q[3] = q[4] = 1.0;
            break;
          }
          case 6:
            q[0] = std::pow(z, p.m1);
            q[3] = p.h84_pert * cell_uniform(p.seed, cell, -0.5, 0.5);
            q[4] = std::pow(z, p.m1 + 1.0);
            break;
          case 7: {
            const double T = 1.0 + p.theta1 * z;
            q[0] = std::pow(T, p.m1);
            q[4] = std::pow(T, p.m1 + 1.0) * (1.0 + p.c91_pert * cell_uniform(p.seed, cell, -0.5, 0.5));
            break;
          }
        }
      }
}

}  // namespace fv3d
