// fv3d_functors.hip -- the reference's functors one pass at a time, on device
// arrays (the fv3d_array part of include/fv3d_hip.h).
//
// An fv3d_array is one Array of the reference, Kokkos::View<double****> of
// extents (Ntz, Nty, Ntx, 5) (SimInfo.h:14), in HBM as FP64 with ghosts,
// variable-major [5][Ntz][Nty][Ntx], x fastest.  Each entry point is one
// reference pass over the same index range (range_tot, range_dom,
// range_slopes, SimInfo.h:298-303) with one thread per cell, so a caller that
// drives the functors one by one -- the reference's own main.cpp and
// UpdateFunctor -- gets the reference's results pass for pass.  fv3d_step is
// the fused form of the same arithmetic; these kernels are for the functor
// interface and for unit parity, not for speed.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fv3d_ctx.h"
#include "fv3d_physics.h"

using namespace fv3d_hip;

struct fv3d_array {
  fv3d_ctx* ctx = nullptr;
  double* d = nullptr;
};

namespace fv3d_hip {

// geometry of a ghosted array
struct ArrGeom {
  int Ntx, Nty, Ntz, Ng;
  int ibeg, iend, jbeg, jend, kbeg, kend;
  long long sy, sz, sf;  // element strides: y, z, variable
};

static ArrGeom arr_geom(const fv3d_params& p) {
  ArrGeom g;
  g.Ntx = p.Ntx; g.Nty = p.Nty; g.Ntz = p.Ntz; g.Ng = p.Ng;
  g.ibeg = p.ibeg; g.iend = p.iend; g.jbeg = p.jbeg; g.jend = p.jend;
  g.kbeg = p.kbeg; g.kend = p.kend;
  g.sy = p.Ntx;
  g.sz = (long long)p.Ntx * p.Nty;
  g.sf = g.sz * p.Ntz;
  return g;
}

// a box of cells [i0, i1) x [j0, j1) x [k0, k1), one thread each (grid-stride)
struct Range3 {
  int i0, j0, k0, ni, nj, nk;
  __device__ __forceinline__ long long count() const { return (long long)ni * nj * nk; }
  __device__ __forceinline__ void cell(long long c, int& i, int& j, int& k) const {
    i = i0 + (int)(c % ni);
    j = j0 + (int)((c / ni) % nj);
    k = k0 + (int)(c / ((long long)ni * nj));
  }
};

static Range3 range_tot(const ArrGeom& g) { return Range3{0, 0, 0, g.Ntx, g.Nty, g.Ntz}; }
static Range3 range_dom(const ArrGeom& g) {
  return Range3{g.ibeg, g.jbeg, g.kbeg, g.iend - g.ibeg, g.jend - g.jbeg, g.kend - g.kbeg};
}
static Range3 range_slopes(const ArrGeom& g) {
  return Range3{g.ibeg - 1, g.jbeg - 1, g.kbeg - 1, g.iend - g.ibeg + 2, g.jend - g.jbeg + 2,
                g.kend - g.kbeg + 2};
}

#define FOR_CELLS(R)                                                                  \
  for (long long c_ = blockIdx.x * (long long)blockDim.x + threadIdx.x; c_ < (R).count(); \
       c_ += (long long)gridDim.x * blockDim.x)

__device__ __forceinline__ long long aidx(const ArrGeom& g, int i, int j, int k) {
  return i + j * g.sy + k * g.sz;
}

__device__ __forceinline__ Prim<double> arr_state(const double* __restrict__ A, const ArrGeom& g,
                                                  long long s) {
  return Prim<double>{A[s], A[s + g.sf], A[s + 2 * g.sf], A[s + 3 * g.sf], A[s + 4 * g.sf]};
}

// consToPrim over range_tot (SimInfo.h:313-321, States.h:34-45)
__global__ void k_arr_cons_to_prim(const double* __restrict__ U, double* __restrict__ Q, ArrGeom g,
                                   Range3 R, double gm1) {
  FOR_CELLS(R) {
    int i, j, k;
    R.cell(c_, i, j, k);
    const long long s = aidx(g, i, j, k);
    const Prim<double> q = cons_to_prim<double>(U[s], U[s + g.sf], U[s + 2 * g.sf],
                                                U[s + 3 * g.sf], U[s + 4 * g.sf], gm1);
    Q[s] = q.r; Q[s + g.sf] = q.u; Q[s + 2 * g.sf] = q.v; Q[s + 3 * g.sf] = q.w;
    Q[s + 4 * g.sf] = q.p;
  }
}

// primToCons over range_tot (SimInfo.h:323-331, States.h:20-31)
__global__ void k_arr_prim_to_cons(const double* __restrict__ Q, double* __restrict__ U, ArrGeom g,
                                   Range3 R, double gm1) {
  FOR_CELLS(R) {
    int i, j, k;
    R.cell(c_, i, j, k);
    const long long s = aidx(g, i, j, k);
    double u[5];
    prim_to_cons(arr_state(Q, g, s), gm1, u);
#pragma unroll
    for (int f = 0; f < 5; ++f) U[s + f * g.sf] = u[f];
  }
}

// UpdateFunctor::computeSlopes over range_slopes (Update.h:55-88): x and y
// take minmod(backward, forward), z minmod(forward, backward)
__global__ void k_arr_slopes(const double* __restrict__ Q, double* __restrict__ sx,
                             double* __restrict__ sy, double* __restrict__ sz, ArrGeom g,
                             Range3 R) {
  FOR_CELLS(R) {
    int i, j, k;
    R.cell(c_, i, j, k);
    const long long s = aidx(g, i, j, k);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double* q = Q + v * g.sf;
      const double dL = q[s] - q[s - 1];
      const double dR = q[s + 1] - q[s];
      const double dU = q[s] - q[s - g.sy];
      const double dD = q[s + g.sy] - q[s];
      const double dF = q[s + g.sz] - q[s];
      const double dB = q[s] - q[s - g.sz];
      sx[s + v * g.sf] = minmod(dL, dR);
      sy[s + v * g.sf] = minmod(dU, dD);
      sz[s + v * g.sf] = minmod(dF, dB);
    }
  }
}

// reconstruct (Update.h:14-33): the face state of cell s on its `sign` side
// along DIR, rotated into the face frame
template <int DIR>
__device__ __forceinline__ Prim<double> reconstruct(const double* __restrict__ Q,
                                                    const double* __restrict__ sl, const ArrGeom& g,
                                                    long long s, double sign, int recon,
                                                    const DevParams& P) {
  const Prim<double> q = arr_state(Q, g, s);
  Prim<double> res = q;
  if (recon == FV3D_PLM) {
    const Prim<double> d = arr_state(sl, g, s);
    res = Prim<double>{q.r + d.r * sign * 0.5, q.u + d.u * sign * 0.5, q.v + d.v * sign * 0.5,
                       q.w + d.w * sign * 0.5, q.p + d.p * sign * 0.5};
  } else if (recon == FV3D_PCM_WB && DIR == FV3D_IZ) {
    res.p = q.p + sign * q.r * P.g * P.dz * 0.5;
  }
  return rotate<double, DIR>(res);
}

// updateAlongDir of computeFluxesAndUpdate (Update.h:101-155) for one cell
template <int DIR>
__device__ __forceinline__ void update_along_dir(const double* __restrict__ Q,
                                                 const double* __restrict__ sl,
                                                 double* __restrict__ Unew, const ArrGeom& g,
                                                 int k, long long s, double dt, int recon,
                                                 int rs, const DevParams& P) {
  const long long st = DIR == FV3D_IX ? 1 : (DIR == FV3D_IY ? g.sy : g.sz);
  const Prim<double> qCL = reconstruct<DIR>(Q, sl, g, s, -1.0, recon, P);
  const Prim<double> qCR = reconstruct<DIR>(Q, sl, g, s, 1.0, recon, P);
  const Prim<double> qL = reconstruct<DIR>(Q, sl, g, s - st, 1.0, recon, P);
  const Prim<double> qR = reconstruct<DIR>(Q, sl, g, s + st, -1.0, recon, P);
  double fL[5], fR[5], poutL, poutR;
  if (rs == FV3D_HLL) {
    hll<double>(qL, qCL, P.gamma, P.gm1, fL, poutL);
    hll<double>(qCR, qR, P.gamma, P.gm1, fR, poutR);
  } else {
    hllc<double>(qL, qCL, P.gamma, P.entho, fL, poutL);
    hllc<double>(qCR, qR, P.gamma, P.entho, fR, poutR);
  }
  unrotate_flux<double, DIR>(fL);
  unrotate_flux<double, DIR>(fR);
  if (DIR == FV3D_IZ && (k == g.kbeg || k == g.kend - 1)) {
    if (P.wb_flux_zbc) {  // Update.h:132-137
      if (k == g.kbeg) {
        fL[0] = 0.0; fL[1] = 0.0; fL[2] = 0.0; fL[3] = poutR - Q[s] * P.g * P.dz; fL[4] = 0.0;
      } else {
        fR[0] = 0.0; fR[1] = 0.0; fR[2] = 0.0; fR[3] = poutL + Q[s] * P.g * P.dz; fR[4] = 0.0;
      }
    } else if (P.c91_flux_zbc) {  // Update.h:138-143
      if (k == g.kbeg) {
        fL[0] = 0.0; fL[1] = 0.0; fL[2] = 0.0; fL[3] = poutL; fL[4] = 0.0;
      } else {
        fR[0] = 0.0; fR[1] = 0.0; fR[2] = 0.0; fR[3] = poutR; fR[4] = 0.0;
      }
    }
  }
  const double d = DIR == FV3D_IX ? P.dx : (DIR == FV3D_IY ? P.dy : P.dz);
#pragma unroll
  for (int n = 0; n < 5; ++n) Unew[s + n * g.sf] = Unew[s + n * g.sf] + dt * (fL[n] - fR[n]) / d;
  if (DIR == FV3D_IZ && P.gravity) {  // Update.h:148-151
    Unew[s + 3 * g.sf] = Unew[s + 3 * g.sf] + dt * Q[s] * P.g;
    Unew[s + 4 * g.sf] = Unew[s + 4 * g.sf] + dt * 0.5 * (fL[0] + fR[0]) * P.g;
  }
}

// UpdateFunctor::computeFluxesAndUpdate over range_dom (Update.h:90-162)
__global__ void k_arr_fluxes_and_update(const double* __restrict__ Q, const double* __restrict__ sx,
                                        const double* __restrict__ sy,
                                        const double* __restrict__ sz, double* __restrict__ Unew,
                                        ArrGeom g, Range3 R, double dt, int recon, int rs,
                                        DevParams P) {
  FOR_CELLS(R) {
    int i, j, k;
    R.cell(c_, i, j, k);
    const long long s = aidx(g, i, j, k);
    update_along_dir<FV3D_IX>(Q, sx, Unew, g, k, s, dt, recon, rs, P);
    update_along_dir<FV3D_IY>(Q, sy, Unew, g, k, s, dt, recon, rs, P);
    update_along_dir<FV3D_IZ>(Q, sz, Unew, g, k, s, dt, recon, rs, P);
    Unew[s] = fmax(1.0e-6, Unew[s]);  // Update.h:160
  }
}

// ThermalConductionFunctor::applyThermalConduction over range_dom
// (ThermalConduction.h:32-97); kztab holds B02's face conductivities per
// level, else every face has 0.5 (kappa + kappa)
__global__ void k_arr_conduction(const double* __restrict__ Q, double* __restrict__ Unew,
                                 ArrGeom g, Range3 R, double dt, DevParams P) {
  FOR_CELLS(R) {
    int i, j, k;
    R.cell(c_, i, j, k);
    const long long s = aidx(g, i, j, k);
    auto T = [&](long long o) { return Q[o + 4 * g.sf] / Q[o]; };
    const int kd = k - g.kbeg;
    double kxy = P.kface, kzF = P.kface, kzB = P.kface;
    if (P.kztab) {
      kxy = P.kztab[kd];
      kzF = P.kztab[P.Nz + kd];
      kzB = P.kztab[2 * P.Nz + kd];
    }
    const double inc = conduction_increment<double>(
        T(s), T(s - 1), T(s + 1), T(s - g.sy), T(s + g.sy), T(s - g.sz), T(s + g.sz), kxy, kzF,
        kzB, k == g.kbeg, k == g.kend - 1, P, false, false, false, dt / P.dx, dt / P.dy,
        dt / P.dz);
    Unew[s + 4 * g.sf] = Unew[s + 4 * g.sf] + inc;
  }
}

// ViscosityFunctor::applyViscosity over range_dom (Viscosity.h:22-155)
__global__ void k_arr_viscosity(const double* __restrict__ Q, double* __restrict__ Unew,
                                ArrGeom g, Range3 R, double dt, DevParams P) {
  FOR_CELLS(R) {
    int i, j, k;
    R.cell(c_, i, j, k);
    const long long s = aidx(g, i, j, k);
    auto S = [&](int dk, int dj, int di, int v) -> double {
      return Q[s + (di - 1) + (dj - 1) * g.sy + (dk - 1) * g.sz + v * g.sf];
    };
    double vx[5] = {0, 0, 0, 0, 0}, vy[5] = {0, 0, 0, 0, 0}, vz[5] = {0, 0, 0, 0, 0};
    viscous_fluxes_by_side<double>(S, 1.0 / P.dx, 1.0 / P.dy, 1.0 / P.dz, P.mu, vx, vy, vz);
    // Viscosity.h:149-151: un += dt * (vf_x + vf_y + vf_z), component 0 included
#pragma unroll
    for (int n = 0; n < 5; ++n)
      Unew[s + n * g.sf] = Unew[s + n * g.sf] + dt * (vx[n] + vy[n] + vz[n]);
  }
}

// the SSP-RK2 correction over range_dom (Update.h:198-204)
__global__ void k_arr_rk2_correct(const double* __restrict__ U0, double* __restrict__ Unew,
                                  ArrGeom g, Range3 R) {
  FOR_CELLS(R) {
    int i, j, k;
    R.cell(c_, i, j, k);
    const long long s = aidx(g, i, j, k);
#pragma unroll
    for (int v = 0; v < 5; ++v) Unew[s + v * g.sf] = 0.5 * (U0[s + v * g.sf] + Unew[s + v * g.sf]);
  }
}

// ComputeDtFunctor::computeDt's hydro term over range_dom (ComputeDt.h:28-35):
// max of (cs+|u|)/dx + (cs+|v|)/dy + (cs+|w|)/dz into acc[0] (an exact
// atomicMax on the bit image of a positive double; acc[0] must start at 0),
// and acc[1] = 1 if a cell's value is not finite (left out of the max, as
// the reference's Max reducer leaves out NaN)
__global__ void __launch_bounds__(256)
k_arr_dt(const double* __restrict__ Q, unsigned long long* acc, ArrGeom g, Range3 R,
         DevParams P) {
  __shared__ double sred[4];
  double m = 0.0;
  int bad = 0;
  FOR_CELLS(R) {
    int i, j, k;
    R.cell(c_, i, j, k);
    const Prim<double> q = arr_state(Q, g, aidx(g, i, j, k));
    const double cs = sound_speed(q.r, q.p, P.gamma);
    const double ih = (cs + fabs(q.u)) / P.dx + (cs + fabs(q.v)) / P.dy + (cs + fabs(q.w)) / P.dz;
    if (ih > m) m = ih;
    if (!(ih <= 1.0e300)) bad = 1;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(m, off);
    m = (o > m) ? o : m;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sred[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) mm = (sred[w] > mm) ? sred[w] : mm;
    if (mm > 0.0) atomicMax(acc, double_as_u64(mm));
  }
  if (bad) atomicOr(acc + 1, 1ull);
}

// state (SoA domain [5][Nz][Ny][Nx], real) <-> the domain of a ghosted FP64 array
template <typename real>
__global__ void k_state_to_arr(const real* __restrict__ U, double* __restrict__ A, ArrGeom g,
                               Range3 R, DevParams P) {
  FOR_CELLS(R) {
    int i, j, k;
    R.cell(c_, i, j, k);
    const long long s = (i - g.ibeg) + (j - g.jbeg) * P.sy + (k - g.kbeg) * P.sz;
    const long long d = aidx(g, i, j, k);
#pragma unroll
    for (int v = 0; v < 5; ++v) A[d + v * g.sf] = (double)U[s + v * P.sf];
  }
}
template <typename real>
__global__ void k_arr_to_state(const double* __restrict__ A, real* __restrict__ U, ArrGeom g,
                               Range3 R, DevParams P) {
  FOR_CELLS(R) {
    int i, j, k;
    R.cell(c_, i, j, k);
    const long long s = (i - g.ibeg) + (j - g.jbeg) * P.sy + (k - g.kbeg) * P.sz;
    const long long d = aidx(g, i, j, k);
#pragma unroll
    for (int v = 0; v < 5; ++v) U[s + v * P.sf] = (real)A[d + v * g.sf];
  }
}

}  // namespace fv3d_hip

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

int fnblocks(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (int)std::max(1LL, b);
}

long long cells_of(const Range3& R) { return (long long)R.ni * R.nj * R.nk; }

int bad_arg(fv3d_ctx* c, const char* what) { return ctx_fail(c, FV3D_EINVAL, what); }

// every array belongs to ctx
bool owns(const fv3d_ctx* c, std::initializer_list<const fv3d_array*> as) {
  for (const fv3d_array* a : as)
    if (!a || a->ctx != c) return false;
  return true;
}

#define FNCHK(ctx, expr)                                                                    \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return ctx_fail((ctx), FV3D_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// scratch of the array path (allocated on first use, kept by the context):
// the slopes arrays UpdateFunctor owns (Update.h:49-51, zero on creation)
// and RK2's U0 and Ustar (Update.h:184-185, allocated per call in the
// reference), plus the dt accumulator
int ensure_fn_scratch(fv3d_ctx* c, bool rk2) {
  const size_t bytes = (size_t)5 * c->p.Ntx * c->p.Nty * c->p.Ntz * sizeof(double);
  const int need = rk2 ? 5 : 3;
  for (int n = 0; n < need; ++n) {
    if (c->fn_arrays[n]) continue;
    FNCHK(c, hipMalloc(&c->fn_arrays[n], bytes));
    FNCHK(c, hipMemsetAsync(c->fn_arrays[n], 0, bytes, c->stream));
  }
  return FV3D_OK;
}

int launch_dt_array(fv3d_ctx* c, const double* Q, double inv[3], double* dt) {
  if (!c->fn_acc) FNCHK(c, hipMalloc(&c->fn_acc, 2 * sizeof(unsigned long long)));
  const ArrGeom g = arr_geom(c->p);
  const Range3 R = range_dom(g);
  FNCHK(c, launch_fill_zero_u64(c->fn_acc, 2, c->stream));
  long long b = (cells_of(R) + 255) / 256;
  hipLaunchKernelGGL(k_arr_dt, dim3((int)std::min(b, 4096LL)), dim3(256), 0, c->stream, Q,
                     c->fn_acc, g, R, c->dp);
  FNCHK(c, hipGetLastError());
  unsigned long long h[2];
  FNCHK(c, hipMemcpyAsync(h, c->fn_acc, sizeof h, hipMemcpyDeviceToHost, c->stream));
  FNCHK(c, hipStreamSynchronize(c->stream));
  double ih;
  std::memcpy(&ih, &h[0], sizeof ih);
  // ComputeDt.h:62: CFL / max(inv_dt_hyp, inv_dt_par_tc, inv_dt_par_visc)
  double m = ih;
  if (m < c->dp.inv_tc) m = c->dp.inv_tc;
  if (m < c->dp.inv_visc) m = c->dp.inv_visc;
  if (inv) { inv[0] = ih; inv[1] = c->dp.inv_tc; inv[2] = c->dp.inv_visc; }
  if (dt) *dt = c->p.CFL / m;
  if (h[1]) return ctx_fail(c, FV3D_ENONFINITE, "non-finite inverse time step in the array");
  return FV3D_OK;
}

// the array passes, on raw device pointers
int pass_fill(fv3d_ctx* c, double* Q) {
  FNCHK(c, launch_fill_boundaries(Q, c->p, c->stream));
  return FV3D_OK;
}
int pass_slopes(fv3d_ctx* c, const double* Q, double* sx, double* sy, double* sz) {
  const ArrGeom g = arr_geom(c->p);
  const Range3 R = range_slopes(g);
  hipLaunchKernelGGL(k_arr_slopes, dim3(fnblocks(cells_of(R))), dim3(256), 0, c->stream, Q, sx,
                     sy, sz, g, R);
  FNCHK(c, hipGetLastError());
  return FV3D_OK;
}
int pass_fluxes(fv3d_ctx* c, const double* Q, const double* sx, const double* sy, const double* sz,
                double* Unew, double dt) {
  const ArrGeom g = arr_geom(c->p);
  const Range3 R = range_dom(g);
  hipLaunchKernelGGL(k_arr_fluxes_and_update, dim3(fnblocks(cells_of(R))), dim3(256), 0,
                     c->stream, Q, sx, sy, sz, Unew, g, R, dt, c->p.reconstruction,
                     c->p.riemann_solver, c->dp);
  FNCHK(c, hipGetLastError());
  return FV3D_OK;
}
int pass_tc(fv3d_ctx* c, const double* Q, double* Unew, double dt) {
  const ArrGeom g = arr_geom(c->p);
  const Range3 R = range_dom(g);
  hipLaunchKernelGGL(k_arr_conduction, dim3(fnblocks(cells_of(R))), dim3(256), 0, c->stream, Q,
                     Unew, g, R, dt, c->dp);
  FNCHK(c, hipGetLastError());
  return FV3D_OK;
}
int pass_visc(fv3d_ctx* c, const double* Q, double* Unew, double dt) {
  const ArrGeom g = arr_geom(c->p);
  const Range3 R = range_dom(g);
  hipLaunchKernelGGL(k_arr_viscosity, dim3(fnblocks(cells_of(R))), dim3(256), 0, c->stream, Q,
                     Unew, g, R, dt, c->dp);
  FNCHK(c, hipGetLastError());
  return FV3D_OK;
}
int pass_c2p(fv3d_ctx* c, const double* U, double* Q) {
  const ArrGeom g = arr_geom(c->p);
  const Range3 R = range_tot(g);
  hipLaunchKernelGGL(k_arr_cons_to_prim, dim3(fnblocks(cells_of(R))), dim3(256), 0, c->stream, U,
                     Q, g, R, c->p.gamma0 - 1.0);
  FNCHK(c, hipGetLastError());
  return FV3D_OK;
}

// UpdateFunctor::euler_step (Update.h:164-178)
int pass_euler_step(fv3d_ctx* c, double* Q, double* Unew, double dt) {
  int rc = ensure_fn_scratch(c, false);
  if (!rc) rc = pass_fill(c, Q);
  double** sl = (double**)c->fn_arrays;
  if (!rc && c->p.reconstruction == FV3D_PLM) rc = pass_slopes(c, Q, sl[0], sl[1], sl[2]);
  if (!rc) rc = pass_fluxes(c, Q, sl[0], sl[1], sl[2], Unew, dt);
  if (!rc && c->p.thermal_conductivity_active) rc = pass_tc(c, Q, Unew, dt);
  if (!rc && c->p.viscosity_active) rc = pass_visc(c, Q, Unew, dt);
  return rc;
}

bool dt_ok(double dt) { return std::isfinite(dt); }

}  // namespace

extern "C" {

int fv3d_array_create(fv3d_ctx* c, fv3d_array** out) {
  if (!c || !out) return ctx_fail(c, FV3D_EINVAL, "null argument");
  *out = nullptr;
  FNCHK(c, hipSetDevice(c->device));
  const size_t bytes = (size_t)5 * c->p.Ntx * c->p.Nty * c->p.Ntz * sizeof(double);
  fv3d_array* a = new fv3d_array();
  a->ctx = c;
  if (hipMalloc(&a->d, bytes) != hipSuccess) {
    delete a;
    return ctx_fail(c, FV3D_ENOMEM, "array allocation failed");
  }
  hipError_t e = hipMemsetAsync(a->d, 0, bytes, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    hipFree(a->d);
    delete a;
    return ctx_fail(c, FV3D_EHIP, std::string("array_create: ") + hipGetErrorString(e));
  }
  *out = a;
  return FV3D_OK;
}

int fv3d_array_destroy(fv3d_array* a) {
  if (!a) return FV3D_OK;
  hipSetDevice(a->ctx->device);
  hipStreamSynchronize(a->ctx->stream);
  hipFree(a->d);
  delete a;
  return FV3D_OK;
}

double* fv3d_array_data(const fv3d_array* a) { return a ? a->d : nullptr; }

int fv3d_array_upload(fv3d_array* a, const double* host) {
  if (!a || !host) return ctx_fail(a ? a->ctx : nullptr, FV3D_EINVAL, "null argument");
  fv3d_ctx* c = a->ctx;
  FNCHK(c, hipSetDevice(c->device));
  const size_t nt = (size_t)c->p.Ntx * c->p.Nty * c->p.Ntz;
  std::vector<double> h(5 * nt);
  aos_to_soa_ghosted(c->p, host, h.data());
  FNCHK(c, hipMemcpyAsync(a->d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  FNCHK(c, hipStreamSynchronize(c->stream));
  return FV3D_OK;
}

int fv3d_array_download(const fv3d_array* a, double* host) {
  if (!a || !host) return ctx_fail(a ? a->ctx : nullptr, FV3D_EINVAL, "null argument");
  fv3d_ctx* c = a->ctx;
  FNCHK(c, hipSetDevice(c->device));
  const size_t nt = (size_t)c->p.Ntx * c->p.Nty * c->p.Ntz;
  std::vector<double> h(5 * nt);
  FNCHK(c, hipMemcpyAsync(h.data(), a->d, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FNCHK(c, hipStreamSynchronize(c->stream));
  soa_ghosted_to_aos(c->p, h.data(), host);
  return FV3D_OK;
}

int fv3d_array_copy(fv3d_array* dst, const fv3d_array* src) {
  if (!dst || !src || dst->ctx != src->ctx)
    return ctx_fail(dst ? dst->ctx : nullptr, FV3D_EINVAL, "arrays of different contexts");
  fv3d_ctx* c = dst->ctx;
  FNCHK(c, hipSetDevice(c->device));
  const size_t bytes = (size_t)5 * c->p.Ntx * c->p.Nty * c->p.Ntz * sizeof(double);
  FNCHK(c, hipMemcpyAsync(dst->d, src->d, bytes, hipMemcpyDeviceToDevice, c->stream));
  return FV3D_OK;
}

int fv3d_cons_to_prim_array(fv3d_ctx* c, const fv3d_array* U, fv3d_array* Q) {
  if (!c || !owns(c, {U, Q})) return bad_arg(c, "arrays must belong to the context");
  FNCHK(c, hipSetDevice(c->device));
  return pass_c2p(c, U->d, Q->d);
}

int fv3d_prim_to_cons_array(fv3d_ctx* c, const fv3d_array* Q, fv3d_array* U) {
  if (!c || !owns(c, {U, Q})) return bad_arg(c, "arrays must belong to the context");
  FNCHK(c, hipSetDevice(c->device));
  const ArrGeom g = arr_geom(c->p);
  const Range3 R = range_tot(g);
  hipLaunchKernelGGL(k_arr_prim_to_cons, dim3(fnblocks(cells_of(R))), dim3(256), 0, c->stream,
                     Q->d, U->d, g, R, c->p.gamma0 - 1.0);
  FNCHK(c, hipGetLastError());
  return FV3D_OK;
}

int fv3d_fill_boundaries_array(fv3d_ctx* c, fv3d_array* Q) {
  if (!c || !owns(c, {Q})) return bad_arg(c, "arrays must belong to the context");
  FNCHK(c, hipSetDevice(c->device));
  return pass_fill(c, Q->d);
}

int fv3d_compute_slopes(fv3d_ctx* c, const fv3d_array* Q, fv3d_array* sx, fv3d_array* sy,
                        fv3d_array* sz) {
  if (!c || !owns(c, {Q, sx, sy, sz})) return bad_arg(c, "arrays must belong to the context");
  FNCHK(c, hipSetDevice(c->device));
  return pass_slopes(c, Q->d, sx->d, sy->d, sz->d);
}

int fv3d_fluxes_and_update(fv3d_ctx* c, const fv3d_array* Q, const fv3d_array* sx,
                           const fv3d_array* sy, const fv3d_array* sz, fv3d_array* Unew,
                           double dt) {
  if (!c || !owns(c, {Q, sx, sy, sz, Unew})) return bad_arg(c, "arrays must belong to the context");
  if (!dt_ok(dt)) return bad_arg(c, "dt must be finite");
  FNCHK(c, hipSetDevice(c->device));
  return pass_fluxes(c, Q->d, sx->d, sy->d, sz->d, Unew->d, dt);
}

int fv3d_apply_tc(fv3d_ctx* c, const fv3d_array* Q, fv3d_array* Unew, double dt) {
  if (!c || !owns(c, {Q, Unew})) return bad_arg(c, "arrays must belong to the context");
  if (!dt_ok(dt)) return bad_arg(c, "dt must be finite");
  FNCHK(c, hipSetDevice(c->device));
  return pass_tc(c, Q->d, Unew->d, dt);
}

int fv3d_apply_visc(fv3d_ctx* c, const fv3d_array* Q, fv3d_array* Unew, double dt) {
  if (!c || !owns(c, {Q, Unew})) return bad_arg(c, "arrays must belong to the context");
  if (!dt_ok(dt)) return bad_arg(c, "dt must be finite");
  FNCHK(c, hipSetDevice(c->device));
  return pass_visc(c, Q->d, Unew->d, dt);
}

int fv3d_euler_step(fv3d_ctx* c, fv3d_array* Q, fv3d_array* Unew, double dt) {
  if (!c || !owns(c, {Q, Unew})) return bad_arg(c, "arrays must belong to the context");
  if (!dt_ok(dt)) return bad_arg(c, "dt must be finite");
  FNCHK(c, hipSetDevice(c->device));
  return pass_euler_step(c, Q->d, Unew->d, dt);
}

int fv3d_update(fv3d_ctx* c, fv3d_array* Q, fv3d_array* Unew, double dt) {
  if (!c || !owns(c, {Q, Unew})) return bad_arg(c, "arrays must belong to the context");
  if (!dt_ok(dt)) return bad_arg(c, "dt must be finite");
  FNCHK(c, hipSetDevice(c->device));
  if (c->p.time_stepping != FV3D_TS_RK2) return pass_euler_step(c, Q->d, Unew->d, dt);
  int rc = ensure_fn_scratch(c, true);
  if (rc) return rc;
  const size_t bytes = (size_t)5 * c->p.Ntx * c->p.Nty * c->p.Ntz * sizeof(double);
  double* U0 = (double*)c->fn_arrays[3];
  double* Ustar = (double*)c->fn_arrays[4];
  // Update.h:188-204 (with U0 and Ustar of the full rank-4 shape, the Q1 fix)
  FNCHK(c, hipMemcpyAsync(U0, Unew->d, bytes, hipMemcpyDeviceToDevice, c->stream));
  FNCHK(c, hipMemcpyAsync(Ustar, Unew->d, bytes, hipMemcpyDeviceToDevice, c->stream));
  rc = pass_euler_step(c, Q->d, Ustar, dt);
  if (rc) return rc;
  FNCHK(c, hipMemcpyAsync(Unew->d, Ustar, bytes, hipMemcpyDeviceToDevice, c->stream));
  rc = pass_c2p(c, Ustar, Q->d);
  if (!rc) rc = pass_euler_step(c, Q->d, Unew->d, dt);
  if (rc) return rc;
  const ArrGeom g = arr_geom(c->p);
  const Range3 R = range_dom(g);
  hipLaunchKernelGGL(k_arr_rk2_correct, dim3(fnblocks(cells_of(R))), dim3(256), 0, c->stream, U0,
                     Unew->d, g, R);
  FNCHK(c, hipGetLastError());
  return FV3D_OK;
}

int fv3d_compute_dt_array(fv3d_ctx* c, const fv3d_array* Q, double inv[3], double* dt) {
  if (!c || !owns(c, {Q})) return bad_arg(c, "arrays must belong to the context");
  FNCHK(c, hipSetDevice(c->device));
  return launch_dt_array(c, Q->d, inv, dt);
}

int fv3d_get_cons_array(fv3d_ctx* c, fv3d_array* U) {
  if (!c || !owns(c, {U})) return bad_arg(c, "arrays must belong to the context");
  if (!c->have_state) return bad_arg(c, "no state set");
  FNCHK(c, hipSetDevice(c->device));
  const ArrGeom g = arr_geom(c->p);
  const Range3 R = range_dom(g);
  if (c->precision == FV3D_FP64)
    hipLaunchKernelGGL(k_state_to_arr<double>, dim3(fnblocks(cells_of(R))), dim3(256), 0,
                       c->stream, (const double*)c->A, U->d, g, R, c->dp);
  else
    hipLaunchKernelGGL(k_state_to_arr<float>, dim3(fnblocks(cells_of(R))), dim3(256), 0,
                       c->stream, (const float*)c->A, U->d, g, R, c->dp);
  FNCHK(c, hipGetLastError());
  return FV3D_OK;
}

int fv3d_set_cons_array(fv3d_ctx* c, const fv3d_array* U) {
  if (!c || !owns(c, {U})) return bad_arg(c, "arrays must belong to the context");
  FNCHK(c, hipSetDevice(c->device));
  const ArrGeom g = arr_geom(c->p);
  const Range3 R = range_dom(g);
  if (c->precision == FV3D_FP64)
    hipLaunchKernelGGL(k_arr_to_state<double>, dim3(fnblocks(cells_of(R))), dim3(256), 0,
                       c->stream, U->d, (double*)c->A, g, R, c->dp);
  else
    hipLaunchKernelGGL(k_arr_to_state<float>, dim3(fnblocks(cells_of(R))), dim3(256), 0,
                       c->stream, U->d, (float*)c->A, g, R, c->dp);
  FNCHK(c, hipGetLastError());
  return ctx_reset_dt(c);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// state-level primToCons / consToPrim (States.h:20-45) on [n][5] states
// ---------------------------------------------------------------------------
namespace fv3d_hip {
__global__ void k_states_convert(int to_cons, long n, double gm1, const double* __restrict__ in,
                                 double* __restrict__ out) {
  const long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double* a = in + 5 * s;
  double* b = out + 5 * s;
  if (to_cons) {
    prim_to_cons(Prim<double>{a[0], a[1], a[2], a[3], a[4]}, gm1, b);
  } else {
    const Prim<double> q = cons_to_prim<double>(a[0], a[1], a[2], a[3], a[4], gm1);
    b[0] = q.r; b[1] = q.u; b[2] = q.v; b[3] = q.w; b[4] = q.p;
  }
}
}  // namespace fv3d_hip

namespace {
int convert_states(int to_cons, double gamma0, long n, const double* in, double* out) {
  if (n < 0 || (n > 0 && (!in || !out))) return record_error(FV3D_EINVAL, "bad argument");
  if (n == 0) return FV3D_OK;
  double *din = nullptr, *dout = nullptr;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    return record_error(FV3D_EHIP, "hipStreamCreate failed");
  const size_t b5 = (size_t)n * 5 * sizeof(double);
  hipError_t e = hipMalloc(&din, b5);
  if (e == hipSuccess) e = hipMalloc(&dout, b5);
  if (e == hipSuccess) e = hipMemcpyAsync(din, in, b5, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_states_convert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       to_cons, n, gamma0 - 1.0, din, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, b5, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (din) hipFree(din);
  if (dout) hipFree(dout);
  hipStreamDestroy(s);
  if (e != hipSuccess) return record_error(FV3D_EHIP, std::string("states: ") + hipGetErrorString(e));
  return FV3D_OK;
}
}  // namespace

extern "C" {

int fv3d_prim_to_cons_states(double gamma0, long n, const double* q, double* u) {
  return convert_states(1, gamma0, n, q, u);
}

int fv3d_cons_to_prim_states(double gamma0, long n, const double* u, double* q) {
  return convert_states(0, gamma0, n, u, q);
}

}  // extern "C"
