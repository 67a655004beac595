// fv3d_aux.hip -- the small kernels around the fused stage: prim<->cons over
// whole arrays (SimInfo.h:313-331), the three ordered ghost-fill passes of
// BoundaryManager::fillBoundaries (BoundaryConditions.h:146-215) on a
// ghosted primitive array, the CFL max-reduction (ComputeDt.h:28-54) and a
// batched Riemann solve used to check the solvers one face at a time.
#include "fv3d_device.h"
#include "fv3d_kernels.h"

namespace fv3d_hip {

// ---------------------------------------------------------------------------
// prim <-> cons
// ---------------------------------------------------------------------------
template <typename real>
__global__ void k_cons_to_prim_ghosted(const real* __restrict__ U, double* __restrict__ Q, DevParams P,
                                       int Ng) {
  const long long n = (long long)P.Nx * P.Ny * P.Nz;
  const long long Ntx = P.Nx + 2 * Ng, Nty = P.Ny + 2 * Ng, Ntz = P.Nz + 2 * Ng;
  const long long qf = Ntx * Nty * Ntz;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c % P.Nx);
    const int j = (int)((c / P.Nx) % P.Ny);
    const int k = (int)(c / ((long long)P.Nx * P.Ny));
    const long long s = i + j * P.sy + k * P.sz;
    real u[5];
    for (int f = 0; f < 5; ++f) u[f] = U[s + f * P.sf];
    const Prim<double> q = cons_to_prim<double>(u[0], u[1], u[2], u[3], u[4], P.gm1);
    const long long d = (i + Ng) + (j + Ng) * Ntx + (k + Ng) * Ntx * Nty;
    Q[d] = q.r; Q[d + qf] = q.u; Q[d + 2 * qf] = q.v; Q[d + 3 * qf] = q.w; Q[d + 4 * qf] = q.p;
  }
}

template <typename real>
__global__ void k_prim_to_cons(const double* __restrict__ Q, real* __restrict__ U, DevParams P) {
  const long long n = (long long)P.Nx * P.Ny * P.Nz;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c % P.Nx);
    const int j = (int)((c / P.Nx) % P.Ny);
    const int k = (int)(c / ((long long)P.Nx * P.Ny));
    const Prim<double> q{Q[c], Q[c + n], Q[c + 2 * n], Q[c + 3 * n], Q[c + 4 * n]};
    double u[5];
    prim_to_cons(q, P.gm1, u);
    const long long s = i + j * P.sy + k * P.sz;
#pragma unroll
    for (int f = 0; f < 5; ++f) U[s + f * P.sf] = (real)u[f];
  }
}

// ---------------------------------------------------------------------------
// ghost fill on Q[5][Ntz][Nty][Ntx] (BoundaryConditions.h)
// ---------------------------------------------------------------------------
struct GhostGeom {
  int Nx, Ny, Nz, Ng, Ntx, Nty, Ntz, ibeg, iend, jbeg, jend, kbeg, kend;
  long long qf;
  int bcx, bcy, bcz;
  double xmin, ymin, zmin, dx, dy, dz, theta1;
  int gravity;
};

__device__ __forceinline__ long long gidx(const GhostGeom& G, int i, int j, int k) {
  return i + (long long)j * G.Ntx + (long long)k * G.Ntx * G.Nty;
}
__device__ __forceinline__ void gcopy(double* Q, const GhostGeom& G, long long dst,
                                      long long src, int flipvar) {
  for (int f = 0; f < 5; ++f) {
    double v = Q[src + f * G.qf];
    if (f == flipvar) v *= -1.0;
    Q[dst + f * G.qf] = v;
  }
}

// fillAbsorbing/fillReflecting/fillPeriodic for one ghost along one axis:
// returns the source index along that axis and whether to negate.
__device__ __forceinline__ int ghost_src(int g, int beg, int end, int N, int ref, int bc,
                                         bool& flip) {
  flip = false;
  if (bc == FV3D_BC_ABSORBING) return ref;
  if (bc == FV3D_BC_REFLECTING) {  // BoundaryConditions.h:21-53
    const int piv = (g < ref ? beg : end);
    flip = true;
    return 2 * piv - g - 1;
  }
  return g < beg ? g + N : g - N;  // periodic (and C91 in x, y)
}

__global__ void k_fill_x(double* Q, GhostGeom G) {
  const long long n = (long long)G.Ng * G.Ny * G.Nz;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c % G.Ng);
    const int j = G.jbeg + (int)((c / G.Ng) % G.Ny);
    const int k = G.kbeg + (int)(c / ((long long)G.Ng * G.Ny));
    bool fl;
    int s = ghost_src(i, G.ibeg, G.iend, G.Nx, G.ibeg, G.bcx, fl);
    gcopy(Q, G, gidx(G, i, j, k), gidx(G, s, j, k), fl ? FV3D_IU : -1);
    const int ir = G.iend + i;
    s = ghost_src(ir, G.ibeg, G.iend, G.Nx, G.iend - 1, G.bcx, fl);
    gcopy(Q, G, gidx(G, ir, j, k), gidx(G, s, j, k), fl ? FV3D_IU : -1);
  }
}

__global__ void k_fill_y(double* Q, GhostGeom G) {
  const long long n = (long long)G.Ntx * G.Ng * G.Nz;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c % G.Ntx);
    const int j = (int)((c / G.Ntx) % G.Ng);
    const int k = G.kbeg + (int)(c / ((long long)G.Ntx * G.Ng));
    bool fl;
    int s = ghost_src(j, G.jbeg, G.jend, G.Ny, G.jbeg, G.bcy, fl);
    gcopy(Q, G, gidx(G, i, j, k), gidx(G, i, s, k), fl ? FV3D_IV : -1);
    const int jb = G.jend + j;
    s = ghost_src(jb, G.jbeg, G.jend, G.Ny, G.jend - 1, G.bcy, fl);
    gcopy(Q, G, gidx(G, i, jb, k), gidx(G, i, s, k), fl ? FV3D_IV : -1);
  }
}

// fillC91 (BoundaryConditions.h:86-134) for ghost plane k of column (i, j)
__device__ void fill_c91(double* Q, const GhostGeom& G, int i, int j, int k, int kref) {
  const int kpiv = (k < kref ? G.kbeg : G.kend);
  const int ksym = 2 * kpiv - k - 1;
  const double z = G.zmin + (k - G.kbeg + 0.5) * G.dz;
  const long long ref = gidx(G, i, j, kref), sym = gidx(G, i, j, ksym);
  const double rr = Q[ref], rp = Q[ref + 4 * G.qf];
  const double sr = Q[sym], su = Q[sym + G.qf], sv = Q[sym + 2 * G.qf], sw = Q[sym + 3 * G.qf];
  const double grav = G.gravity ? 1.0 : 0.0;
  double res[5];
  if (k < G.kbeg) {
    const double dz = z - 0.5 * G.dz;
    const double T = 1.0 + z * G.theta1;
    res[4] = rp + dz * rr * grav;
    res[0] = res[4] / T;
    res[4] = fmax(res[4], 1.0e-6);
    res[0] = fmax(res[0], 1.0e-6);
    const double dratio = sr / res[0];
    res[1] = su;
    res[2] = sv;
    res[3] = -sw * dratio;
  } else {
    const double Tref = rp / rr;
    const double dz = z - (1.0 - 0.5 * G.dz);
    const double T = Tref + dz * G.theta1;
    const double Pp = rp + dz * rr * grav;
    const double rho = Pp / T;
    res[4] = fmax(Pp, 1.0e-6);
    res[0] = fmax(rho, 1.0e-6);
    const double dratio = sr / rho;
    res[1] = su * dratio;
    res[2] = sv * dratio;
    res[3] = -sw * dratio;
  }
  const long long d = gidx(G, i, j, k);
  for (int f = 0; f < 5; ++f) Q[d + f * G.qf] = res[f];
}

__global__ void k_fill_z(double* Q, GhostGeom G) {
  const long long n = (long long)G.Ntx * G.Nty * G.Ng;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c % G.Ntx);
    const int j = (int)((c / G.Ntx) % G.Nty);
    const int k = (int)(c / ((long long)G.Ntx * G.Nty));
    const int kb = G.kend + k;
    if (G.bcz == FV3D_BC_C91) {
      fill_c91(Q, G, i, j, k, G.kbeg);
      fill_c91(Q, G, i, j, kb, G.kend - 1);
      continue;
    }
    bool fl;
    int s = ghost_src(k, G.kbeg, G.kend, G.Nz, G.kbeg, G.bcz, fl);
    gcopy(Q, G, gidx(G, i, j, k), gidx(G, i, j, s), fl ? FV3D_IW : -1);
    s = ghost_src(kb, G.kbeg, G.kend, G.Nz, G.kend - 1, G.bcz, fl);
    gcopy(Q, G, gidx(G, i, j, kb), gidx(G, i, j, s), fl ? FV3D_IW : -1);
  }
}

static int nblocks(long long n, int bs) {
  long long b = (n + bs - 1) / bs;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (int)b;
}

hipError_t launch_fill_boundaries(double* Q, const fv3d_params& p, hipStream_t s) {
  GhostGeom G;
  G.Nx = p.Nx; G.Ny = p.Ny; G.Nz = p.Nz; G.Ng = p.Ng;
  G.Ntx = p.Ntx; G.Nty = p.Nty; G.Ntz = p.Ntz;
  G.ibeg = p.ibeg; G.iend = p.iend; G.jbeg = p.jbeg; G.jend = p.jend;
  G.kbeg = p.kbeg; G.kend = p.kend;
  G.qf = (long long)p.Ntx * p.Nty * p.Ntz;
  G.bcx = p.boundary_x; G.bcy = p.boundary_y; G.bcz = p.boundary_z;
  G.xmin = p.xmin; G.ymin = p.ymin; G.zmin = p.zmin;
  G.dx = p.dx; G.dy = p.dy; G.dz = p.dz; G.theta1 = p.theta1;
  G.gravity = p.gravity;
  const int bs = 256;
  hipLaunchKernelGGL(k_fill_x, dim3(nblocks((long long)G.Ng * G.Ny * G.Nz, bs)), dim3(bs), 0, s,
                     Q, G);
  hipLaunchKernelGGL(k_fill_y, dim3(nblocks((long long)G.Ntx * G.Ng * G.Nz, bs)), dim3(bs), 0, s,
                     Q, G);
  hipLaunchKernelGGL(k_fill_z, dim3(nblocks((long long)G.Ntx * G.Nty * G.Ng, bs)), dim3(bs), 0,
                     s, Q, G);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// CFL reduction (ComputeDt.h:28-54): max over the domain of inv_dt_hyp.
// Positive doubles order like their bit patterns, so the max is an exact,
// order-free 64-bit atomicMax.
// ---------------------------------------------------------------------------
template <typename real>
__global__ void __launch_bounds__(256) k_dt_reduce(const real* __restrict__ U,
                                                   unsigned long long* inv, int* health,
                                                   DevParams P) {
  __shared__ double sred[4];
  const long long n = (long long)P.Nx * P.Ny * P.Nz;
  double m = 0.0;
  int bad = 0;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c % P.Nx);
    const int j = (int)((c / P.Nx) % P.Ny);
    const int k = (int)(c / ((long long)P.Nx * P.Ny));
    const long long s = i + j * P.sy + k * P.sz;
    const Prim<double> q = cons_to_prim<double>(U[s], U[s + P.sf], U[s + 2 * P.sf],
                                                U[s + 3 * P.sf], U[s + 4 * P.sf], P.gm1);
    const double cs = sound_speed(q.r, q.p, P.gamma);
    const double a = cs + fabs(q.u), b = cs + fabs(q.v), cc = cs + fabs(q.w);
    const double ih = (P.pow2x ? a * P.idx : a / P.dx) + (P.pow2y ? b * P.idy : b / P.dy) +
                      (P.pow2z ? cc * P.idz : cc / P.dz);
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
    if (mm > 0.0) atomicMax(inv, double_as_u64(mm));
  }
  if (bad) atomicOr(health, 1);
}

// ---------------------------------------------------------------------------
// batched Riemann solves (RiemannSolvers.h) for unit checks
// ---------------------------------------------------------------------------
__global__ void k_riemann(int solver, double gamma, long n, const double* __restrict__ qL,
                          const double* __restrict__ qR, double* __restrict__ flux,
                          double* __restrict__ pout) {
  const long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (s >= n) return;
  const Prim<double> L{qL[5 * s], qL[5 * s + 1], qL[5 * s + 2], qL[5 * s + 3], qL[5 * s + 4]};
  const Prim<double> R{qR[5 * s], qR[5 * s + 1], qR[5 * s + 2], qR[5 * s + 3], qR[5 * s + 4]};
  double F[5], po;
  const double gm1 = gamma - 1.0, entho = 1.0 / (gamma - 1.0);
  if (solver == FV3D_HLL) hll(L, R, gamma, gm1, F, po);
  else hllc(L, R, gamma, entho, F, po);
  for (int f = 0; f < 5; ++f) flux[5 * s + f] = F[f];
  pout[s] = po;
}

// plain copy of n doubles (16-byte accesses): the STREAM-copy reference
// kernel, and the calibration kernel for the profiler's byte counters
// A plain device copy at the bandwidth HBM gives a streaming pattern: each
// thread keeps four 16-byte loads in flight, and the loads and stores are
// non-temporal (the data is touched once).  Measured 6.1 TB/s over 1 GiB,
// against 4.6 for a grid-stride loop of one 16-byte load per iteration.
constexpr int COPY_UNROLL = 4;
__global__ void __launch_bounds__(256) k_copy(const double2* __restrict__ a, double2* __restrict__ b,
                                              long long n2) {
  const long long base = (long long)blockIdx.x * (256 * COPY_UNROLL) + threadIdx.x;
  double v[COPY_UNROLL][2];
#pragma unroll
  for (int u = 0; u < COPY_UNROLL; ++u) {
    const long long i = base + u * 256;
    if (i < n2) {
      v[u][0] = __builtin_nontemporal_load(&a[i].x);
      v[u][1] = __builtin_nontemporal_load(&a[i].y);
    }
  }
#pragma unroll
  for (int u = 0; u < COPY_UNROLL; ++u) {
    const long long i = base + u * 256;
    if (i < n2) {
      __builtin_nontemporal_store(v[u][0], &b[i].x);
      __builtin_nontemporal_store(v[u][1], &b[i].y);
    }
  }
}

hipError_t launch_copy(const double* a, double* b, long long n, hipStream_t s) {
  const long long n2 = n / 2;
  const long long blocks = (n2 + 256 * COPY_UNROLL - 1) / (256 * COPY_UNROLL);
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, (const double2*)a, (double2*)b, n2);
  return hipGetLastError();
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void k_put_one_u64(unsigned long long* p, unsigned long long v) { *p = v; }

hipError_t launch_put_u64(unsigned long long* p, unsigned long long v, hipStream_t s) {
  hipLaunchKernelGGL(k_put_one_u64, dim3(1), dim3(1), 0, s, p, v);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename real>
hipError_t launch_cons_to_prim_ghosted(const real* U, double* Q, const DevParams& P, int Ng,
                                       hipStream_t s) {
  const long long n = (long long)P.Nx * P.Ny * P.Nz;
  hipLaunchKernelGGL((k_cons_to_prim_ghosted<real>), dim3(nblocks(n, 256)), dim3(256), 0, s, U, Q, P,
                     Ng);
  return hipGetLastError();
}

template <typename real>
hipError_t launch_prim_to_cons(const double* Q, real* U, const DevParams& P, hipStream_t s) {
  const long long n = (long long)P.Nx * P.Ny * P.Nz;
  hipLaunchKernelGGL((k_prim_to_cons<real>), dim3(nblocks(n, 256)), dim3(256), 0, s, Q, U, P);
  return hipGetLastError();
}

template <typename real>
hipError_t launch_dt_reduce(const real* U, unsigned long long* inv, int* health,
                            const DevParams& P, hipStream_t s) {
  const long long n = (long long)P.Nx * P.Ny * P.Nz;
  long long b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  hipLaunchKernelGGL((k_dt_reduce<real>), dim3((int)b), dim3(256), 0, s, U, inv, health, P);
  return hipGetLastError();
}

hipError_t launch_riemann_batch(int solver, double gamma, long n, const double* qL,
                                const double* qR, double* flux, double* pout, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_riemann, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, solver, gamma,
                     n, qL, qR, flux, pout);
  return hipGetLastError();
}

hipError_t launch_fill_zero_u64(unsigned long long* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u64, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
  return hipGetLastError();
}

template hipError_t launch_cons_to_prim_ghosted<double>(const double*, double*, const DevParams&, int,
                                                        hipStream_t);
template hipError_t launch_cons_to_prim_ghosted<float>(const float*, double*, const DevParams&, int,
                                                       hipStream_t);
template hipError_t launch_prim_to_cons<double>(const double*, double*, const DevParams&,
                                                hipStream_t);
template hipError_t launch_prim_to_cons<float>(const double*, float*, const DevParams&,
                                               hipStream_t);
template hipError_t launch_dt_reduce<double>(const double*, unsigned long long*, int*,
                                             const DevParams&, hipStream_t);
template hipError_t launch_dt_reduce<float>(const float*, unsigned long long*, int*,
                                            const DevParams&, hipStream_t);

}  // namespace fv3d_hip
