// fv3d_api.hip -- the C ABI of libfv3d_hip (declared in include/fv3d_hip.h).
//
// A context owns the device state of one box: two conserved-state buffers
// (A = U^n, B = U* of RK2 / the previous state of Euler), a three-slot ring
// of inv_dt_hyp accumulators and a health flag.  A step is:
//   RK2:   stage(A -> B)               reads slot s,   clears slot s+2
//          stage(B, rk=A -> A)         reads slot s,   reduces slot s+1
//   Euler: stage(A -> B), swap(A, B)   reads slot s,   reduces s+1, clears s+2
// The FP32 compensated RK2 adds a third state buffer C and a compensation
// array:
//          inc1(A -> U* in B)          reads slot s,   clears slot s+2
//          inc2(B, rk=A, comp -> C), swap(A, C)   reduces slot s+1
// so dt never round-trips through the host and consecutive steps need no
// synchronisation.  The slot cleared by step s was last read by step s-1 and
// is next written by step s+1, so clearing it inside step s races with
// nothing.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

#include "fv3d_ctx.h"

using namespace fv3d_hip;

static thread_local std::string g_err;

namespace fv3d_hip {
int record_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
}  // namespace fv3d_hip

int ctx_fail(fv3d_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  g_err = msg;
  return code;
}

static int fail(fv3d_ctx* c, int code, const std::string& msg) { return ctx_fail(c, code, msg); }

static bool kahan(const fv3d_ctx* c) { return ctx_kahan(c); }

extern "C" {

int fv3d_abi_version(void) { return 1; }
size_t fv3d_params_size(void) { return sizeof(fv3d_params); }

const char* fv3d_last_error(const fv3d_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_err.c_str();
}

int fv3d_create(const fv3d_params* p, int precision, int device, fv3d_ctx** out) {
  if (!p || !out) return fail(nullptr, FV3D_EINVAL, "null argument");
  *out = nullptr;
  std::string why;
  if (check_params(*p, why)) return fail(nullptr, FV3D_EINVAL, why);
  if (precision != FV3D_FP64 && precision != FV3D_FP32 && precision != FV3D_FP32_PLAIN)
    return fail(nullptr, FV3D_EINVAL, "precision not supported");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, FV3D_EHIP, "no HIP device visible");
  if (device < 0 || device >= ndev) return fail(nullptr, FV3D_EINVAL, "device index out of range");
  fv3d_ctx* c = new fv3d_ctx();
  c->p = *p;
  c->precision = precision;
  c->device = device;
  c->dp = make_devparams(*p, 0, device_cus(device), precision == FV3D_FP64);
  c->elem = precision == FV3D_FP64 ? sizeof(double) : sizeof(float);
  c->state_bytes = (size_t)5 * p->Nx * p->Ny * p->Nz * c->elem;
  auto cleanup = [&](int code, const std::string& m) {
    fv3d_destroy(c);
    return fail(nullptr, code, m);
  };
  if (hipSetDevice(device) != hipSuccess) return cleanup(FV3D_EHIP, "hipSetDevice failed");
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(FV3D_EHIP, "hipStreamCreate failed");
  if (hipMalloc(&c->A, c->state_bytes) != hipSuccess ||
      hipMalloc(&c->B, c->state_bytes) != hipSuccess)
    return cleanup(FV3D_ENOMEM, "state allocation failed");
  if (kahan(c) && (hipMalloc(&c->C, c->state_bytes) != hipSuccess ||
                   hipMalloc(&c->Cmp, c->state_bytes) != hipSuccess))
    return cleanup(FV3D_ENOMEM, "state allocation failed");
  if (p->boundary_z == FV3D_BC_C91) {
    if (hipMalloc(&c->c91g, 4 * 5 * (size_t)p->Nx * p->Ny * c->elem) != hipSuccess)
      return cleanup(FV3D_ENOMEM, "allocation failed");
    c->dp.c91g = c->c91g;
  }
  if (p->thermal_conductivity_active && p->thermal_conductivity_mode == FV3D_TCM_B02) {
    const std::vector<double> t = b02_face_table(*p);
    if (hipMalloc(&c->kztab, t.size() * sizeof(double)) != hipSuccess)
      return cleanup(FV3D_ENOMEM, "allocation failed");
    if (hipMemcpy(c->kztab, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(FV3D_EHIP, "copy failed");
    c->dp.kztab = c->kztab;
  }
  if (hipMalloc(&c->ring, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&c->health, sizeof(int)) != hipSuccess)
    return cleanup(FV3D_ENOMEM, "allocation failed");
  if (hipMemsetAsync(c->ring, 0, 4 * sizeof(unsigned long long), c->stream) != hipSuccess ||
      hipMemsetAsync(c->health, 0, sizeof(int), c->stream) != hipSuccess)
    return cleanup(FV3D_EHIP, "memset failed");
  *out = c;
  return FV3D_OK;
}

int fv3d_destroy(fv3d_ctx* c) {
  if (!c) return FV3D_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->A) hipFree(c->A);
  if (c->B) hipFree(c->B);
  if (c->C) hipFree(c->C);
  if (c->Cmp) hipFree(c->Cmp);
  if (c->ring) hipFree(c->ring);
  if (c->health) hipFree(c->health);
  if (c->qbuf) hipFree(c->qbuf);
  if (c->c91g) hipFree(c->c91g);
  if (c->kztab) hipFree(c->kztab);
  for (void* a : c->fn_arrays)
    if (a) hipFree(a);
  if (c->fn_acc) hipFree(c->fn_acc);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return FV3D_OK;
}

}  // extern "C"

static int ensure_qbuf(fv3d_ctx* c) {
  if (c->qbuf) return FV3D_OK;
  const size_t nt = (size_t)c->p.Ntx * c->p.Nty * c->p.Ntz;
  if (hipMalloc(&c->qbuf, 5 * nt * sizeof(double)) != hipSuccess)
    return fail(c, FV3D_ENOMEM, "staging allocation failed");
  return FV3D_OK;
}

// prim of a device state buffer into host memory
static int state_prim_to_host(fv3d_ctx* c, const void* U, double* Q, int with_ghosts) {
  int rc = ensure_qbuf(c);
  if (rc) return rc;
  const fv3d_params& p = c->p;
  const size_t nt = (size_t)p.Ntx * p.Nty * p.Ntz;
  if (c->precision == FV3D_FP64)
    HIPCHK(c, launch_cons_to_prim_ghosted<double>((const double*)U, c->qbuf, c->dp, p.Ng, c->stream));
  else
    HIPCHK(c, launch_cons_to_prim_ghosted<float>((const float*)U, c->qbuf, c->dp, p.Ng, c->stream));
  if (with_ghosts) HIPCHK(c, launch_fill_boundaries(c->qbuf, p, c->stream));
  std::vector<double> h(5 * nt);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->qbuf, 5 * nt * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (with_ghosts) {
    soa_ghosted_to_aos(p, h.data(), Q);
  } else {
    const long long tx = p.Ntx, ty = p.Nty;
    for (int k = 0; k < p.Nz; ++k)
      for (int j = 0; j < p.Ny; ++j)
        for (int i = 0; i < p.Nx; ++i) {
          const long long s = (i + p.Ng) + tx * ((j + p.Ng) + ty * (k + p.Ng));
          double* d = Q + (((long long)k * p.Ny + j) * p.Nx + i) * 5;
          for (int f = 0; f < 5; ++f) d[f] = h[f * nt + s];
        }
  }
  return FV3D_OK;
}

int ctx_reset_dt(fv3d_ctx* c) {
  HIPCHK(c, launch_fill_zero_u64(c->ring, 4, c->stream));
  HIPCHK(c, hipMemsetAsync(c->health, 0, sizeof(int), c->stream));
  if (c->precision == FV3D_FP64)
    HIPCHK(c, launch_dt_reduce<double>((const double*)c->A, c->ring, c->health, c->dp, c->stream));
  else
    HIPCHK(c, launch_dt_reduce<float>((const float*)c->A, c->ring, c->health, c->dp, c->stream));
  if (c->Cmp) HIPCHK(c, hipMemsetAsync(c->Cmp, 0, c->state_bytes, c->stream));
  c->step = 0;
  c->have_state = true;
  c->have_stage = false;
  return FV3D_OK;
}

extern "C" {

int fv3d_set_prim(fv3d_ctx* c, const double* Q, int with_ghosts) {
  if (!c || !Q) return fail(c, FV3D_EINVAL, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<double> h;
  aos_to_soa_domain(c->p, Q, with_ghosts, h);
  double* d = nullptr;
  HIPCHK(c, hipMalloc(&d, h.size() * sizeof(double)));
  hipError_t e = hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                                c->stream);
  if (e == hipSuccess)
    e = c->precision == FV3D_FP64 ? launch_prim_to_cons<double>(d, (double*)c->A, c->dp, c->stream)
                                  : launch_prim_to_cons<float>(d, (float*)c->A, c->dp, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(d);
  if (e != hipSuccess) return fail(c, FV3D_EHIP, std::string("set_prim: ") + hipGetErrorString(e));
  return ctx_reset_dt(c);
}

int fv3d_set_cons(fv3d_ctx* c, const double* U, int with_ghosts) {
  if (!c || !U) return fail(c, FV3D_EINVAL, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<double> h;
  aos_to_soa_domain(c->p, U, with_ghosts, h);
  std::vector<float> hf;
  const void* src = h.data();
  if (c->precision != FV3D_FP64) {
    hf.assign(h.begin(), h.end());
    src = hf.data();
  }
  HIPCHK(c, hipMemcpyAsync(c->A, src, c->state_bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ctx_reset_dt(c);
}

int fv3d_get_cons(fv3d_ctx* c, double* U, int with_ghosts) {
  if (!c || !U) return fail(c, FV3D_EINVAL, "null argument");
  if (!c->have_state) return fail(c, FV3D_EINVAL, "no state set");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t n = c->state_bytes / c->elem;
  std::vector<double> h(n);
  if (c->precision == FV3D_FP64) {
    HIPCHK(c, hipMemcpyAsync(h.data(), c->A, c->state_bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> hf(n);
    HIPCHK(c, hipMemcpyAsync(hf.data(), c->A, c->state_bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    h.assign(hf.begin(), hf.end());
  }
  soa_domain_to_aos(c->p, h.data(), with_ghosts, U);
  return FV3D_OK;
}

int fv3d_get_prim(fv3d_ctx* c, double* Q, int with_ghosts) {
  if (!c || !Q) return fail(c, FV3D_EINVAL, "null argument");
  if (!c->have_state) return fail(c, FV3D_EINVAL, "no state set");
  HIPCHK(c, hipSetDevice(c->device));
  return state_prim_to_host(c, c->A, Q, with_ghosts);
}

int fv3d_get_stage_prim(fv3d_ctx* c, double* Q, int with_ghosts) {
  if (!c || !Q) return fail(c, FV3D_EINVAL, "null argument");
  if (!c->have_stage) return fail(c, FV3D_EINVAL, "no step taken yet");
  HIPCHK(c, hipSetDevice(c->device));
  return state_prim_to_host(c, c->B, Q, with_ghosts);
}

int fv3d_compute_dt(fv3d_ctx* c, double inv[3], double* dt) {
  if (!c) return fail(c, FV3D_EINVAL, "null argument");
  if (!c->have_state) return fail(c, FV3D_EINVAL, "no state set");
  HIPCHK(c, hipSetDevice(c->device));
  unsigned long long bits = 0;
  int health = 0;
  HIPCHK(c, hipMemcpyAsync(&bits, c->ring + (c->step % 3), sizeof bits, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(&health, c->health, sizeof health, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double ih;
  std::memcpy(&ih, &bits, sizeof ih);
  double m = ih;
  if (m < c->dp.inv_tc) m = c->dp.inv_tc;
  if (m < c->dp.inv_visc) m = c->dp.inv_visc;
  if (inv) { inv[0] = ih; inv[1] = c->dp.inv_tc; inv[2] = c->dp.inv_visc; }
  if (dt) *dt = c->p.CFL / m;
  if (health) return fail(c, FV3D_ENONFINITE, "non-finite inverse time step in the state");
  return FV3D_OK;
}

}  // extern "C"

// dt_slot: nullptr takes dt from the ring (CFL / max of the inverse time
// steps of the current state); otherwise the device word holds dt itself
template <typename real>
static int launch_step_t(fv3d_ctx* c, hipEvent_t* ev, const unsigned long long* dt_slot) {
  const long s = c->step;
  unsigned long long* cur = c->ring + (s % 3);
  unsigned long long* nxt = c->ring + ((s + 1) % 3);
  unsigned long long* zer = c->ring + ((s + 2) % 3);
  StageLaunch<real> L;
  L.recon = c->p.reconstruction == FV3D_PLM ? FV3D_PLM : FV3D_PCM;
  L.riemann = c->p.riemann_solver;
  L.tc = c->p.thermal_conductivity_active != 0;
  L.visc = c->p.viscosity_active != 0;
  L.health = c->health;
  L.P = c->dp;
  L.inv_cur = cur;
  if (dt_slot) {
    L.inv_cur = dt_slot;
    L.P.dt_direct = 1;
  }
  real* A = (real*)c->A;
  real* B = (real*)c->B;
  if (kahan(c)) {
    L.mode = 1;  // STAGE_INC1: U* into B
    L.Uin = A; L.Urk = nullptr; L.Uout = B; L.inv_next = nullptr; L.inv_zero = zer;
    if (ev) HIPCHK(c, hipEventRecord(ev[0], c->stream));
    HIPCHK(c, launch_stage<real>(L, c->stream));
    if (ev) HIPCHK(c, hipEventRecord(ev[1], c->stream));
    L.mode = 2;  // STAGE_INC2: from U*, compensated U^{n+1} into C
    L.Uin = B; L.Urk = A; L.Cio = (real*)c->Cmp; L.Uout = (real*)c->C;
    L.inv_next = nxt; L.inv_zero = nullptr;
    HIPCHK(c, launch_stage<real>(L, c->stream));
    if (ev) HIPCHK(c, hipEventRecord(ev[2], c->stream));
    std::swap(c->A, c->C);
  } else if (c->p.time_stepping == FV3D_TS_RK2) {
    L.Uin = A; L.Urk = nullptr; L.Uout = B; L.inv_next = nullptr; L.inv_zero = zer;
    if (ev) HIPCHK(c, hipEventRecord(ev[0], c->stream));
    HIPCHK(c, launch_stage<real>(L, c->stream));
    if (ev) HIPCHK(c, hipEventRecord(ev[1], c->stream));
    L.Uin = B; L.Urk = A; L.Uout = A; L.inv_next = nxt; L.inv_zero = nullptr;
    HIPCHK(c, launch_stage<real>(L, c->stream));
    if (ev) HIPCHK(c, hipEventRecord(ev[2], c->stream));
  } else {
    L.Uin = A; L.Urk = nullptr; L.Uout = B; L.inv_next = nxt; L.inv_zero = zer;
    if (ev) HIPCHK(c, hipEventRecord(ev[0], c->stream));
    HIPCHK(c, launch_stage<real>(L, c->stream));
    if (ev) HIPCHK(c, hipEventRecord(ev[1], c->stream));
    std::swap(c->A, c->B);
  }
  c->step = s + 1;
  c->have_stage = true;
  return FV3D_OK;
}

static int launch_step(fv3d_ctx* c, hipEvent_t* ev, const unsigned long long* dt_slot = nullptr) {
  return c->precision == FV3D_FP64 ? launch_step_t<double>(c, ev, dt_slot)
                                   : launch_step_t<float>(c, ev, dt_slot);
}

extern "C" {

int fv3d_step(fv3d_ctx* c) {
  if (!c) return fail(c, FV3D_EINVAL, "null argument");
  if (!c->have_state) return fail(c, FV3D_EINVAL, "no state set");
  HIPCHK(c, hipSetDevice(c->device));
  return launch_step(c, nullptr);
}

int fv3d_step_dt(fv3d_ctx* c, double dt) {
  if (!c) return fail(c, FV3D_EINVAL, "null argument");
  if (!c->have_state) return fail(c, FV3D_EINVAL, "no state set");
  if (!std::isfinite(dt)) return fail(c, FV3D_EINVAL, "dt must be finite");
  HIPCHK(c, hipSetDevice(c->device));
  // ring slot 3 is not part of the three-slot inv_dt_hyp ring: it carries
  // the caller's dt to both stages, stream-ordered before them
  unsigned long long bits;
  std::memcpy(&bits, &dt, sizeof bits);
  HIPCHK(c, launch_put_u64(c->ring + 3, bits, c->stream));
  return launch_step(c, nullptr, c->ring + 3);
}

int fv3d_steps(fv3d_ctx* c, int nsteps) {
  if (!c) return fail(c, FV3D_EINVAL, "null argument");
  if (!c->have_state) return fail(c, FV3D_EINVAL, "no state set");
  HIPCHK(c, hipSetDevice(c->device));
  for (int i = 0; i < nsteps; ++i) {
    int rc = launch_step(c, nullptr);
    if (rc) return rc;
  }
  return FV3D_OK;
}

long fv3d_step_count(const fv3d_ctx* c) { return c ? c->step : -1; }

int fv3d_synchronize(fv3d_ctx* c) {
  if (!c) return fail(c, FV3D_EINVAL, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FV3D_OK;
}

int fv3d_time_steps(fv3d_ctx* c, int nsteps, float* total_ms, float* stage_ms) {
  if (!c || nsteps <= 0) return fail(c, FV3D_EINVAL, "bad argument");
  if (!c->have_state) return fail(c, FV3D_EINVAL, "no state set");
  HIPCHK(c, hipSetDevice(c->device));
  const int per = c->p.time_stepping == FV3D_TS_RK2 ? 3 : 2;
  std::vector<hipEvent_t> ev(stage_ms ? (size_t)per * nsteps : 0);
  hipEvent_t t0, t1;
  HIPCHK(c, hipEventCreate(&t0));
  HIPCHK(c, hipEventCreate(&t1));
  for (auto& e : ev) HIPCHK(c, hipEventCreate(&e));
  HIPCHK(c, hipEventRecord(t0, c->stream));
  int rc = FV3D_OK;
  for (int i = 0; i < nsteps && rc == FV3D_OK; ++i)
    rc = launch_step(c, stage_ms ? &ev[(size_t)per * i] : nullptr);
  HIPCHK(c, hipEventRecord(t1, c->stream));
  HIPCHK(c, hipEventSynchronize(t1));
  if (rc == FV3D_OK) {
    if (total_ms) HIPCHK(c, hipEventElapsedTime(total_ms, t0, t1));
    if (stage_ms) {
      int o = 0;
      for (int i = 0; i < nsteps; ++i)
        for (int s = 0; s + 1 < per; ++s)
          HIPCHK(c, hipEventElapsedTime(&stage_ms[o++], ev[(size_t)per * i + s],
                                        ev[(size_t)per * i + s + 1]));
    }
  }
  for (auto& e : ev) hipEventDestroy(e);
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  return rc;
}

int fv3d_fill_boundaries(const fv3d_params* p, double* Q) {
  if (!p || !Q) return fail(nullptr, FV3D_EINVAL, "null argument");
  const size_t nt = (size_t)p->Ntx * p->Nty * p->Ntz;
  std::vector<double> h(5 * nt);
  aos_to_soa_ghosted(*p, Q, h.data());
  double* d = nullptr;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    return fail(nullptr, FV3D_EHIP, "hipStreamCreate failed");
  hipError_t e = hipMalloc(&d, 5 * nt * sizeof(double));
  if (e == hipSuccess) e = hipMemcpyAsync(d, h.data(), 5 * nt * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = launch_fill_boundaries(d, *p, s);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, 5 * nt * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (d) hipFree(d);
  hipStreamDestroy(s);
  if (e != hipSuccess) return fail(nullptr, FV3D_EHIP, std::string("fill_boundaries: ") + hipGetErrorString(e));
  soa_ghosted_to_aos(*p, h.data(), Q);
  return FV3D_OK;
}

int fv3d_stream_copy(size_t nbytes, int reps, float* ms_per_copy) {
  if (nbytes < 16 || reps <= 0) return fail(nullptr, FV3D_EINVAL, "bad argument");
  const long long n = (long long)(nbytes / 16) * 2;
  double *a = nullptr, *b = nullptr;
  hipStream_t s;
  hipEvent_t e0, e1;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    return fail(nullptr, FV3D_EHIP, "hipStreamCreate failed");
  hipError_t e = hipMalloc(&a, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&b, n * sizeof(double));
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, n * sizeof(double), s);
  if (e == hipSuccess) e = launch_copy(a, b, n, s);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy(a, b, n, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (ms_per_copy) *ms_per_copy = ms / reps;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (a) hipFree(a);
  if (b) hipFree(b);
  hipStreamDestroy(s);
  if (e != hipSuccess) return fail(nullptr, FV3D_EHIP, std::string("stream_copy: ") + hipGetErrorString(e));
  return FV3D_OK;
}

int fv3d_riemann(int solver, double gamma0, long n, const double* qL, const double* qR,
                 double* flux, double* pout) {
  if (n < 0 || (n > 0 && (!qL || !qR || !flux || !pout)))
    return fail(nullptr, FV3D_EINVAL, "bad argument");
  if (solver != FV3D_HLL && solver != FV3D_HLLC) return fail(nullptr, FV3D_EINVAL, "solver");
  if (n == 0) return FV3D_OK;
  double *dl = nullptr, *dr = nullptr, *df = nullptr, *dp = nullptr;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    return fail(nullptr, FV3D_EHIP, "hipStreamCreate failed");
  const size_t b5 = (size_t)n * 5 * sizeof(double), b1 = (size_t)n * sizeof(double);
  hipError_t e = hipMalloc(&dl, b5);
  if (e == hipSuccess) e = hipMalloc(&dr, b5);
  if (e == hipSuccess) e = hipMalloc(&df, b5);
  if (e == hipSuccess) e = hipMalloc(&dp, b1);
  if (e == hipSuccess) e = hipMemcpyAsync(dl, qL, b5, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dr, qR, b5, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = launch_riemann_batch(solver, gamma0, n, dl, dr, df, dp, s);
  if (e == hipSuccess) e = hipMemcpyAsync(flux, df, b5, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(pout, dp, b1, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  for (double* x : {dl, dr, df, dp})
    if (x) hipFree(x);
  hipStreamDestroy(s);
  if (e != hipSuccess) return fail(nullptr, FV3D_EHIP, std::string("riemann: ") + hipGetErrorString(e));
  return FV3D_OK;
}

}  // extern "C"
