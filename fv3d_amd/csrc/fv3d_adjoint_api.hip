// fv3d_adjoint_api.hip -- frozen-dt forward runs and the discrete adjoint of
// the time loop (declared in include/fv3d_hip.h).
//
// The forward map of n steps with a given dt sequence is
//   RK2:   U1 = S(U^s),  U^{s+1} = 0.5 (U^s + S(U1))      (Update.h:183-205)
//   Euler: U^{s+1} = S(U^s)                              (Update.h:181-182)
// with S(U) = U + dt_s R(U).  Its adjoint, for lam = dJ/dU^{s+1}:
//   RK2:   w1  = 0.5 lam + dt J_R(U1)^T (0.5 lam)
//          lam <- 0.5 lam + w1 + dt J_R(U^s)^T w1
//   Euler: lam <- lam + dt J_R(U^s)^T lam
// States are checkpointed every K steps in the forward pass; each segment
// is recomputed from its checkpoint (keeping U^s and U1 of every step) and
// then swept backwards.  Segments that fit in the context's store budget
// (fv3d_set_adjoint_store; the last segments first) keep those states from
// the forward pass itself and are not recomputed.  The parameter gradients dJ/dkappa and dJ/dmu are
// the sums over all stages of (adjoint of the stage output) . dS/dtheta.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fv3d_ctx.h"

using namespace fv3d_hip;

namespace {

template <typename real>
struct AdjBuffers {
  std::vector<real*> owned;
  hipError_t alloc(real** p, size_t n) {
    hipError_t e = hipMalloc((void**)p, n * sizeof(real));
    if (e == hipSuccess) owned.push_back(*p);
    return e;
  }
  ~AdjBuffers() {
    for (real* p : owned) hipFree(p);
  }
};

template <typename real>
StageLaunch<real> frozen_stage(const fv3d_ctx* c, const double* dt_slot) {
  StageLaunch<real> L;
  L.recon = c->p.reconstruction == FV3D_PLM ? FV3D_PLM : FV3D_PCM;
  L.riemann = c->p.riemann_solver;
  L.tc = c->p.thermal_conductivity_active != 0;
  L.visc = c->p.viscosity_active != 0;
  L.health = c->health;
  L.P = c->dp;
  L.P.dt_direct = 1;
  L.inv_cur = (const unsigned long long*)dt_slot;
  L.inv_next = nullptr;
  L.inv_zero = nullptr;
  L.Urk = nullptr;
  L.mode = 0;
  return L;
}

// The density floor fmax(1e-6, rho) (Update.h:160) has acted on a stage
// output exactly where its density is the floor value (the floor maps every
// smaller or NaN density there; a density that lands on 1e-6 by itself is a
// kink of the map as well).  Inside an adjoint run, which differentiates the
// map without the floor, every frozen stage output is scanned and *flag set
// if the floor acted.
template <typename real>
__global__ void __launch_bounds__(256) k_floor_scan(const real* __restrict__ rho, long long n,
                                                    int* flag) {
  int hit = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    hit |= rho[i] == (real)1.0e-6;
  if (__any(hit) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

template <typename real>
hipError_t floor_scan(const fv3d_ctx* c, const real* U) {
  if (!c->adj_floor) return hipSuccess;
  const long long n = (long long)c->p.Nx * c->p.Ny * c->p.Nz;
  const long long b = std::min<long long>((n + 255) / 256, 2048);
  hipLaunchKernelGGL((k_floor_scan<real>), dim3((unsigned)b), dim3(256), 0, c->stream, U, n,
                     c->adj_floor);
  return hipGetLastError();
}

// one frozen-dt step: U0 -> U1 (stage state, RK2 only) -> Unext.  Unext may
// alias U0.  For Euler, U1 is unused and Unext must differ from U0.
template <typename real>
hipError_t frozen_step(const fv3d_ctx* c, const double* dt_slot, const real* U0, real* U1,
                       real* Unext) {
  StageLaunch<real> L = frozen_stage<real>(c, dt_slot);
  hipError_t e;
  if (c->p.time_stepping == FV3D_TS_RK2) {
    L.Uin = U0; L.Uout = U1;
    e = launch_stage<real>(L, c->stream);
    if (e == hipSuccess) e = floor_scan<real>(c, U1);
    if (e != hipSuccess) return e;
    L.Uin = U1; L.Urk = U0; L.Uout = Unext;
  } else {
    L.Uin = U0; L.Uout = Unext;
  }
  e = launch_stage<real>(L, c->stream);
  if (e == hipSuccess) e = floor_scan<real>(c, Unext);
  return e;
}

template <typename real>
int forward_frozen_t(fv3d_ctx* c, int nsteps, const double* dts_dev) {
  real* A = (real*)c->A;
  real* B = (real*)c->B;
  for (int s = 0; s < nsteps; ++s) {
    if (c->p.time_stepping == FV3D_TS_RK2) {
      HIPCHK(c, frozen_step<real>(c, dts_dev + s, A, B, A));
    } else {
      HIPCHK(c, frozen_step<real>(c, dts_dev + s, A, nullptr, B));
      std::swap(A, B);
    }
  }
  if (A != (real*)c->A) HIPCHK(c, hipMemcpyAsync(c->A, A, c->state_bytes, hipMemcpyDeviceToDevice,
                                                 c->stream));
  return FV3D_OK;
}

template <typename real>
int adjoint_run_t(fv3d_ctx* c, int nsteps, int K, const double* dts_dev,
                  const std::vector<double>& lamT, std::vector<double>& lam0, double grads[2]) {
  const long long N = (long long)c->p.Nx * c->p.Ny * c->p.Nz;
  const size_t ns = 5 * (size_t)N;
  const bool rk2 = c->p.time_stepping == FV3D_TS_RK2;
  const int ncp = (nsteps + K - 1) / K;
  // segments kept from the forward pass: cp >= first_kept.  A kept segment
  // holds U^s for its K+1 step boundaries, then (RK2) the K stage states.
  const size_t seg_states = rk2 ? 2 * (size_t)K + 1 : (size_t)K + 1;
  const size_t seg_bytes = seg_states * ns * sizeof(real);
  const int keep = (int)std::min<size_t>((size_t)ncp, c->adj_store_bytes / seg_bytes);
  const int first_kept = ncp - keep;
  AdjBuffers<real> buf;
  real *ckpt = nullptr, *seg0 = nullptr, *seg1 = nullptr, *lam = nullptr, *w1 = nullptr,
       *lam2 = nullptr, *Qb = nullptr, *tmp = nullptr, *store = nullptr;
  double* grad = nullptr;
  auto kept0 = [&](int cp) { return store + (size_t)(cp - first_kept) * seg_states * ns; };
  auto kept1 = [&](int cp) { return kept0(cp) + (size_t)(K + 1) * ns; };
  if (keep > 0 && buf.alloc(&store, seg_states * ns * keep) != hipSuccess)
    return ctx_fail(c, FV3D_ENOMEM, "adjoint: segment store allocation failed");
  const bool recompute = first_kept > 0;
  if (buf.alloc(&ckpt, ns * ncp) != hipSuccess ||
      (recompute && buf.alloc(&seg0, ns * (K + 1)) != hipSuccess) ||
      (recompute && rk2 && buf.alloc(&seg1, ns * K) != hipSuccess) || buf.alloc(&lam, ns) != hipSuccess ||
      buf.alloc(&w1, ns) != hipSuccess || buf.alloc(&lam2, ns) != hipSuccess ||
      buf.alloc(&Qb, 6 * (size_t)N) != hipSuccess || buf.alloc(&tmp, ns) != hipSuccess)
    return ctx_fail(c, FV3D_ENOMEM, "adjoint: checkpoint allocation failed");
  // the two gradients, then room for one partial pair per diffusion tile
  const size_t ngrad = 2 + 2 * (size_t)N;
  HIPCHK(c, hipMalloc(&grad, ngrad * sizeof(double)));
  struct GradFree { double* p; ~GradFree() { hipFree(p); } } gf{grad};
  // every frozen stage of this run (forward pass and recomputation) is
  // scanned for densities the floor fmax(1e-6, rho) (Update.h:160) set:
  // the adjoint differentiates the map without it
  int* floor_flag = nullptr;
  HIPCHK(c, hipMalloc(&floor_flag, sizeof(int)));
  struct FloorGuard {
    fv3d_ctx* c;
    int* p;
    ~FloorGuard() { c->adj_floor = nullptr; hipFree(p); }
  } fg{c, floor_flag};
  HIPCHK(c, hipMemsetAsync(floor_flag, 0, sizeof(int), c->stream));
  c->adj_floor = floor_flag;
  HIPCHK(c, hipMemsetAsync(grad, 0, ngrad * sizeof(double), c->stream));
  HIPCHK(c, hipMemsetAsync(Qb, 0, 6 * (size_t)N * sizeof(real), c->stream));

  // adjoint of the final state, and the dt sequence on the host
  {
    std::vector<real> h(lamT.begin(), lamT.end());
    HIPCHK(c, hipMemcpyAsync(lam, h.data(), ns * sizeof(real), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  std::vector<double> hdts(nsteps);
  HIPCHK(c, hipMemcpyAsync(hdts.data(), dts_dev, nsteps * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));

  hipEvent_t e0, e1, e2;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  HIPCHK(c, hipEventCreate(&e2));
  HIPCHK(c, hipEventRecord(e0, c->stream));
  // ---- forward pass: checkpoints every K steps (the steps after the last
  // checkpoint are recomputed by the backward pass anyway, unless kept)
  HIPCHK(c, hipMemcpyAsync(ckpt, c->A, ns * sizeof(real), hipMemcpyDeviceToDevice, c->stream));
  {
    real* cur = tmp;
    HIPCHK(c, hipMemcpyAsync(cur, c->A, ns * sizeof(real), hipMemcpyDeviceToDevice, c->stream));
    real* other = lam2;  // scratch until the backward pass
    const int fwd_end = keep > 0 ? nsteps : (ncp - 1) * K;
    for (int s = 0; s < fwd_end; ++s) {
      const int cp = s / K, i = s - cp * K;
      if (cp >= first_kept) {  // a kept segment: step inside the store
        real* S0 = kept0(cp);
        if (i == 0)
          HIPCHK(c, hipMemcpyAsync(S0, cur, ns * sizeof(real), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, frozen_step<real>(c, dts_dev + s, S0 + (size_t)i * ns,
                                    rk2 ? kept1(cp) + (size_t)i * ns : nullptr,
                                    S0 + (size_t)(i + 1) * ns));
        if (i == K - 1 && s + 1 < nsteps)
          HIPCHK(c, hipMemcpyAsync(cur, S0 + (size_t)K * ns, ns * sizeof(real),
                                   hipMemcpyDeviceToDevice, c->stream));
        continue;
      }
      if (rk2) {
        HIPCHK(c, frozen_step<real>(c, dts_dev + s, cur, w1, cur));
      } else {
        HIPCHK(c, frozen_step<real>(c, dts_dev + s, cur, nullptr, other));
        std::swap(cur, other);
      }
      if ((s + 1) % K == 0)
        HIPCHK(c, hipMemcpyAsync(ckpt + (size_t)((s + 1) / K) * ns, cur, ns * sizeof(real),
                                 hipMemcpyDeviceToDevice, c->stream));
    }
  }
  HIPCHK(c, hipEventRecord(e1, c->stream));

  // ---- backward pass
  AdjLaunch<real> A;
  A.recon = c->p.reconstruction == FV3D_PLM ? FV3D_PLM : FV3D_PCM;
  A.riemann = c->p.riemann_solver;
  A.tc = c->p.thermal_conductivity_active != 0;
  A.visc = c->p.viscosity_active != 0;
  A.P = c->dp;
  A.Qb = Qb;
  A.grad = grad;
  A.simple = std::getenv("FV3D_ADJ_SIMPLE") != nullptr;
  A.colored = std::getenv("FV3D_ADJ_COLORED") != nullptr;
  A.colored_diffusion = A.colored;
  A.no_fused_finalize = std::getenv("FV3D_ADJ_UNFUSED") != nullptr;
  for (int cp = ncp - 1; cp >= 0; --cp) {
    const int s0 = cp * K, s1 = std::min(nsteps, s0 + K);
    // the segment's states: S0[i] = U^{s0+i}, S1[i] = U1 of step s0+i
    real* S0 = seg0;
    real* S1 = seg1;
    if (cp >= first_kept) {
      S0 = kept0(cp);
      S1 = kept1(cp);
    } else {  // recompute the segment from its checkpoint
      HIPCHK(c, hipMemcpyAsync(S0, ckpt + (size_t)cp * ns, ns * sizeof(real),
                               hipMemcpyDeviceToDevice, c->stream));
      for (int s = s0; s < s1 - 1; ++s) {
        const int i = s - s0;
        HIPCHK(c, frozen_step<real>(c, dts_dev + s, S0 + (size_t)i * ns,
                                    rk2 ? S1 + (size_t)i * ns : nullptr, S0 + (size_t)(i + 1) * ns));
      }
      real* last0 = S0 + (size_t)(s1 - 1 - s0) * ns;
      real* last1 = rk2 ? S1 + (size_t)(s1 - 1 - s0) * ns : nullptr;
      if (cp == ncp - 1) {
        // the run's final step in full: its output U^n is scanned for the
        // density floor like every other stage output (tmp is free now)
        HIPCHK(c, frozen_step<real>(c, dts_dev + (s1 - 1), last0, last1, tmp));
      } else if (rk2) {  // the last step's stage state (its output is not needed)
        StageLaunch<real> L = frozen_stage<real>(c, dts_dev + (s1 - 1));
        L.Uin = last0;
        L.Uout = last1;
        HIPCHK(c, launch_stage<real>(L, c->stream));
        HIPCHK(c, floor_scan<real>(c, last1));
      }
    }
    for (int s = s1 - 1; s >= s0; --s) {
      const int i = s - s0;
      A.dt = hdts[s];
      real* U0 = S0 + (size_t)i * ns;
      if (rk2) {
        A.U = S1 + (size_t)i * ns; A.v = lam; A.sigma = (real)0.5; A.extra = nullptr;
        A.alpha = (real)0; A.out = w1;
        HIPCHK(c, launch_adjoint_stage<real>(A, c->stream));
        A.U = U0; A.v = w1; A.sigma = (real)1; A.extra = lam; A.alpha = (real)0.5; A.out = lam2;
        HIPCHK(c, launch_adjoint_stage<real>(A, c->stream));
      } else {
        A.U = U0; A.v = lam; A.sigma = (real)1; A.extra = nullptr; A.alpha = (real)0; A.out = lam2;
        HIPCHK(c, launch_adjoint_stage<real>(A, c->stream));
      }
      std::swap(lam, lam2);
    }
  }
  HIPCHK(c, hipEventRecord(e2, c->stream));
  HIPCHK(c, hipEventSynchronize(e2));
  float fwd = 0.f, bwd = 0.f;
  HIPCHK(c, hipEventElapsedTime(&fwd, e0, e1));
  HIPCHK(c, hipEventElapsedTime(&bwd, e1, e2));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipEventDestroy(e2);
  c->adj_ms[0] = fwd;
  c->adj_ms[1] = bwd;
  c->adj_ms[2] = fwd + bwd;

  std::vector<real> h(ns);
  int floor_hit = 0;
  HIPCHK(c, hipMemcpyAsync(h.data(), lam, ns * sizeof(real), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(grads, grad, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&floor_hit, floor_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (floor_hit)
    return ctx_fail(c, FV3D_EFLOOR,
                    "adjoint: the density floor fmax(1e-6, rho) (Update.h:160) changed a density in "
                    "this run; the adjoint takes it as inactive, so the gradient would be wrong");
  lam0.assign(h.begin(), h.end());
  return FV3D_OK;
}

int check_adjoint_params(fv3d_ctx* c) {
  const fv3d_params& p = c->p;
  if (p.gravity) return ctx_fail(c, FV3D_EINVAL, "adjoint: gravity is not supported");
  if (p.boundary_z == FV3D_BC_C91) return ctx_fail(c, FV3D_EINVAL, "adjoint: boundaries_z=C91 is not supported");
  if (p.thermal_conductivity_active && p.thermal_conductivity_mode == FV3D_TCM_B02)
    return ctx_fail(c, FV3D_EINVAL, "adjoint: conductivity_mode=B02 is not supported");
  if (p.reconstruction == FV3D_PCM_WB || p.well_balanced_flux_at_z_bc)
    return ctx_fail(c, FV3D_EINVAL, "adjoint: well-balanced variants are not supported");
  return FV3D_OK;
}

int upload_dts(fv3d_ctx* c, int nsteps, const double* dts, double** out) {
  for (int s = 0; s < nsteps; ++s)
    if (!(dts[s] > 0.0) || !std::isfinite(dts[s]))
      return ctx_fail(c, FV3D_EINVAL, "dts must be positive and finite");
  HIPCHK(c, hipMalloc(out, nsteps * sizeof(double)));
  HIPCHK(c, hipMemcpyAsync(*out, dts, nsteps * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return FV3D_OK;
}

}  // namespace

extern "C" {

int fv3d_forward_frozen(fv3d_ctx* c, int nsteps, const double* dts, float* ms) {
  if (!c || nsteps < 0 || (nsteps > 0 && !dts)) return ctx_fail(c, FV3D_EINVAL, "bad argument");
  if (!c->have_state) return ctx_fail(c, FV3D_EINVAL, "no state set");
  if (nsteps == 0) return FV3D_OK;
  HIPCHK(c, hipSetDevice(c->device));
  double* d = nullptr;
  if (int rc = upload_dts(c, nsteps, dts, &d)) return rc;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  HIPCHK(c, hipEventRecord(e0, c->stream));
  const int rc = c->precision == FV3D_FP64 ? forward_frozen_t<double>(c, nsteps, d)
                                           : forward_frozen_t<float>(c, nsteps, d);
  HIPCHK(c, hipEventRecord(e1, c->stream));
  HIPCHK(c, hipEventSynchronize(e1));
  if (ms) HIPCHK(c, hipEventElapsedTime(ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(d);
  if (rc) return rc;
  return ctx_reset_dt(c);
}

int fv3d_adjoint_run(fv3d_ctx* c, int nsteps, int K, const double* dts, const double* dJ_dUT,
                     double* dJ_dU0, double* dJ_dkappa, double* dJ_dmu) {
  if (!c || nsteps <= 0 || K <= 0 || !dts || !dJ_dUT || !dJ_dU0)
    return ctx_fail(c, FV3D_EINVAL, "bad argument");
  if (!c->have_state) return ctx_fail(c, FV3D_EINVAL, "no state set");
  if (int rc = check_adjoint_params(c)) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<double> lamT, lam0;
  aos_to_soa_domain(c->p, dJ_dUT, 0, lamT);
  double* d = nullptr;
  if (int rc = upload_dts(c, nsteps, dts, &d)) return rc;
  double grads[2] = {0.0, 0.0};
  const int rc = c->precision == FV3D_FP64
                     ? adjoint_run_t<double>(c, nsteps, K, d, lamT, lam0, grads)
                     : adjoint_run_t<float>(c, nsteps, K, d, lamT, lam0, grads);
  hipFree(d);
  if (rc) return rc;
  soa_domain_to_aos(c->p, lam0.data(), 0, dJ_dU0);
  if (dJ_dkappa) *dJ_dkappa = grads[0];
  if (dJ_dmu) *dJ_dmu = grads[1];
  return FV3D_OK;
}

int fv3d_set_adjoint_store(fv3d_ctx* c, size_t bytes) {
  if (!c) return FV3D_EINVAL;
  c->adj_store_bytes = bytes;
  return FV3D_OK;
}

int fv3d_adjoint_times(const fv3d_ctx* c, float ms[3]) {
  if (!c || !ms) return FV3D_EINVAL;
  std::memcpy(ms, c->adj_ms, sizeof c->adj_ms);
  return FV3D_OK;
}

}  // extern "C"
