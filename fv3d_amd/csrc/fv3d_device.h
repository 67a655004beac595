// fv3d_device.h -- device-side state algebra, Riemann solvers and ghost maps
// for the CDNA4 kernels of libfv3d_hip.
//
// Every expression keeps the evaluation order of the reference (C++ left to
// right, no contraction: the library is built with -ffp-contract=off), so the
// FP64 results are the reference's bit for bit.  The references next to each
// function give the file:line whose arithmetic it restates.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fv3d_params.h"

namespace fv3d_hip {

// Kernel-constant parameters, passed by value (kernarg segment).
struct DevParams {
  int Nx, Ny, Nz;
  long long sy, sz, sf;         // element strides of the SoA state: x=1, y, z, field
  int bcx, bcy, bcz;            // BoundaryType per axis
  int pcm_wb;                   // PCM_WB pressure shift on z faces (Update.h:22-28)
  int wb_flux_zbc;              // well_balanced_flux_at_z_bc (Update.h:132-137)
  int c91_flux_zbc;             // boundary_z == BC_C91 flux override (Update.h:138-143)
  int gravity;
  int pow2x, pow2y, pow2z;      // d is a power of two: x/d == x*(1/d) exactly
  double gamma, gm1, entho;     // gamma0, gamma0-1.0, 1.0/(gamma0-1.0)
  double dx, dy, dz, idx, idy, idz;
  double g;
  double kface;                 // 0.5*(kappa+kappa): constant-mode face conductivity
  int bctc_zmin, bctc_zmax;
  double bctc_zmin_value, bctc_zmax_value;
  double mu;
  double CFL, inv_tc, inv_visc; // dt inputs; inv_tc/inv_visc are constants of the run
  int twx, tty;                 // stage tile: columns (64 or 32) and rows
  int tiles_x, tiles_y, tiles_z, nzc;
  int dt_direct;                // the dt slot holds dt itself (a frozen dt sequence)
  // boundary_z == BC_C91: primitives of the four z-ghost planes (k = -2, -1,
  // Nz, Nz+1) as [plane][5][Ny][Nx], written before each stage
  void* c91g;
  double zmin, theta1;
  // conductivity_mode == B02: face conductivities per z level, [3][Nz] =
  // (x and y faces, lower z face, upper z face), from the host
  const double* kztab;
};

template <typename real>
struct Prim {
  real r, u, v, w, p;
};

// ---------------------------------------------------------------------------
// state algebra
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Quotients and square roots of the stage arithmetic.  FP64 keeps the IEEE
// results, the reference's correctly rounded bits.  FP32 takes the hardware
// reciprocal (1 ulp) times the numerator, within 2 ulp of the quotient, in 2
// instructions where the range-safe FP32 division takes 6 to 9.
// (Sharing one FP64 reciprocal between quotients with a common denominator,
// behind a wave-uniform check that keeps the IEEE bits, measured 9% slower
// on the 256^3 blast: the branch splits the schedule of the Riemann solver.)
// ---------------------------------------------------------------------------
template <typename real>
__device__ __forceinline__ real qdiv(real a, real b) {
  return a / b;
}
template <>
__device__ __forceinline__ float qdiv<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}

// FP64 reciprocal and square root to about an ulp (not correctly rounded),
// for FP64 arithmetic on FP32 data where the last bit does not matter
__device__ __forceinline__ double recip_f64(double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
  return __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
}
__device__ __forceinline__ double sqrt_f64(double x) {
  const double g = __builtin_amdgcn_rsq(x);
  double s = x * g, h = 0.5 * g;
  const double e = __builtin_fma(-h, s, 0.5);
  s = __builtin_fma(s, e, s);
  h = __builtin_fma(h, e, h);
  s = __builtin_fma(__builtin_fma(-s, s, x), h, s);
  // rsq(0) is inf: keep sqrt(0) = 0 and sqrt(x < 0) = NaN
  return x > 0.0 ? s : (x == 0.0 ? 0.0 : __builtin_nan(""));
}

// ---------------------------------------------------------------------------
// FP64 quotients and square roots without the range-scaling steps (the
// FAST template switch below).  The compiler's correctly rounded FP64
// division is v_div_scale (twice), v_rcp, four FMAs refining the
// reciprocal, a product, an FMA for the remainder, v_div_fmas and
// v_div_fixup.  v_div_scale returns its operand unchanged (and v_div_fmas
// is then a plain FMA) unless the numerator is below about 2^-960 in
// magnitude, the denominator is denormal or the quotient is denormal; zero,
// infinite and NaN operands are settled by v_div_fixup, which both forms
// end with.  Outside those cases the sequence without v_div_scale gives the
// same bits (measured on MI355X over 2.7e8 random pairs per exponent range
// and every pair of special values), and quotients with one denominator can
// share its refined reciprocal.  The square root likewise drops its scaling
// and its zero and infinity fix for positive finite arguments of at least
// 2^-767.  The caller guarantees the operand ranges (the stage kernel tests
// every cell it loads, cell_in_fast_range).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rcp_refined(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  return __builtin_fma(r, e, r);
}
__device__ __forceinline__ double div_refined(double n, double d, double r) {
  const double q = n * r;
  const double rem = __builtin_fma(-d, q, n);
  return __builtin_amdgcn_div_fixup(__builtin_fma(rem, r, q), d, n);
}
__device__ __forceinline__ double sqrt_unscaled(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = y * 0.5;
  const double e = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  double t = __builtin_fma(-g, g, x);
  g = __builtin_fma(t, h, g);
  t = __builtin_fma(-g, g, x);
  return __builtin_fma(t, h, g);
}
// quotient, shared reciprocal and square root of the stage arithmetic:
// FAST (FP64 only) takes the forms above, otherwise qdiv and sqrt
template <typename real, bool FAST>
struct Ar {
  static __device__ __forceinline__ real rcp(real b) { return b; }
  static __device__ __forceinline__ real divr(real a, real b, real) { return qdiv(a, b); }
  static __device__ __forceinline__ real div(real a, real b) { return qdiv(a, b); }
  static __device__ __forceinline__ real sqr(real x) { return sqrt(x); }
};
template <>
struct Ar<double, true> {
  static __device__ __forceinline__ double rcp(double b) { return rcp_refined(b); }
  static __device__ __forceinline__ double divr(double a, double b, double r) {
    return div_refined(a, b, r);
  }
  static __device__ __forceinline__ double div(double a, double b) {
    return div_refined(a, b, rcp_refined(b));
  }
  static __device__ __forceinline__ double sqr(double x) { return sqrt_unscaled(x); }
};

// t / h for a cell width h (1/h = ih): the product when h is a power of two
// (the same bits) and in FP32, else the quotient
template <typename real>
__device__ __forceinline__ real over_h(real t, bool pow2, double ih, double h) {
  if (sizeof(real) == 4 || pow2) return t * (real)ih;
  return t / (real)h;
}

// States.h:34-45 (consToPrim); gm1 = gamma0-1.0
template <typename real>
__device__ __forceinline__ Prim<real> cons_to_prim(real U0, real U1, real U2, real U3, real U4,
                                                   real gm1) {
  Prim<real> q;
  q.r = U0;
  q.u = qdiv(U1, U0);
  q.v = qdiv(U2, U0);
  q.w = qdiv(U3, U0);
  const real Ek = (real)0.5 * q.r * (q.u * q.u + q.v * q.v + q.w * q.w);
  q.p = (U4 - Ek) * gm1;
  return q;
}

// States.h:20-31 (primToCons)
template <typename real>
__device__ __forceinline__ void prim_to_cons(const Prim<real>& q, real gm1, real U[5]) {
  U[0] = q.r;
  U[1] = q.r * q.u;
  U[2] = q.r * q.v;
  U[3] = q.r * q.w;
  const real Ek = (real)0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / q.r;
  U[4] = (Ek + q.p / gm1);
}

// States.h:47-50
template <typename real, bool FAST = false>
__device__ __forceinline__ real sound_speed(real r, real p, real gamma) {
  using A = Ar<real, FAST>;
  return A::sqr(A::div(p * gamma, r));
}

// FAST operand ranges of one cell (FP64): rho and p in [2^-40, 2^41), each
// momentum zero or of magnitude in [2^-40, 2^41).  PLM-minmod face states
// lie between a cell's value and the mean with its neighbour, so face
// densities and pressures stay within a factor 2 of that range and face
// velocities are zero or of magnitude in [2^-134, 2^82].  Every sum,
// difference and product of HLL is then zero or at least the ulp of its
// smallest operand, which bounds its numerators (zero or above 2^-760,
// against the 2^-960 where v_div_scale starts to scale), its quotients
// (normal) and the sound speed's argument (positive, finite and above
// 2^-767), so the FAST forms change no bit.  A positive value in the range
// has its high word in [0x3D700000, 0x42800000).
__device__ __forceinline__ bool cell_in_fast_range(double U0, double U1, double U2, double U3,
                                                   double p) {
  auto hi = [](double x) { return (unsigned)(__double_as_longlong(x) >> 32); };
  auto pos_ok = [&](double x) { return hi(x) - 0x3D700000u < 0x42800000u - 0x3D700000u; };
  auto mom_ok = [&](double x) {
    return ((hi(x) & 0x7FFFFFFFu) - 0x3D700000u < 0x42800000u - 0x3D700000u) || x == 0.0;
  };
  return pos_ok(U0) && pos_ok(p) && mom_ok(U1) && mom_ok(U2) && mom_ok(U3);
}

// Update.h:73-80: minmod(a, b) = a*b<0 ? 0 : (|a|<|b| ? a : b).  The product
// (not a sign test) decides, as in the reference.
template <typename real>
__device__ __forceinline__ real minmod(real a, real b) {
  return (a * b < (real)0) ? (real)0 : ((fabs(a) < fabs(b)) ? a : b);
}
// FP32: the median of (a, b, 0) is minmod(a, b) (same signs: the smaller
// magnitude, ties equal; opposite signs or a zero: 0), one v_med3_f32 for
// five instructions.  It differs from the product test only where a*b
// underflows, which the FP32 form does not need to reproduce.
template <>
__device__ __forceinline__ float minmod<float>(float a, float b) {
  return __builtin_amdgcn_fmed3f(a, b, 0.0f);
}

// ---------------------------------------------------------------------------
// Riemann solvers.  Inputs are face states already rotated into the face
// frame (swap_component, States.h:103-111): slot u is the normal velocity.
// ---------------------------------------------------------------------------

// RiemannSolvers.h:38-112.  Only the selected star side is evaluated; each
// value that is used is produced by the reference's own expression.
template <typename real>
__device__ __forceinline__ void hllc(const Prim<real> L, const Prim<real> R, real gamma,
                                     real entho, real F[5], real& pout) {
  // states by value and selects on scalars: a select between the addresses
  // of two stack objects would force them into scratch memory
  const real rL = L.r, uL = L.u, vL = L.v, wL = L.w, pL = L.p;
  const real rR = R.r, uR = R.u, vR = R.v, wR = R.w, pR = R.p;
  const real ekL = (real)0.5 * rL * (uL * uL + vL * vL);  // drops the 2nd transverse velocity
  const real EL = ekL + pL * entho;
  const real ekR = (real)0.5 * rR * (uR * uR + vR * vR);
  const real ER = ekR + pR * entho;
  const real cL = sound_speed(rL, pL, gamma);
  const real cR = sound_speed(rR, pR, gamma);
  const real cmax = fmax(cL, cR);
  const real SL = fmin(uL, uR) - cmax;
  const real SR = fmax(uL, uR) + cmax;
  const real rcL = rL * (uL - SL);
  const real rcR = rR * (SR - uR);
  const real uS = qdiv(rcR * uR + rcL * uL + (pL - pR), rcR + rcL);
  const real pS = qdiv(rcR * pL + rcL * pR + rcL * rcR * (uL - uR), rcR + rcL);
  // star side K: left when uS > 0, else right
  const bool kl = uS > (real)0;
  const real rK = kl ? rL : rR;
  const real uK = kl ? uL : uR;
  const real pK = kl ? pL : pR;
  const real EK = kl ? EL : ER;
  const real SK = kl ? SL : SR;
  const real rS = qdiv(rK * (SK - uK), SK - uS);
  const real ES = qdiv((SK - uK) * EK - pK * uK + pS * uS, SK - uS);
  const bool left = SL > (real)0;
  const bool star = !left && (kl || SR > (real)0);
  // non-star side: left if SL > 0 else right
  const real sr = star ? rS : (left ? rL : rR);
  const real su = star ? uS : (left ? uL : uR);
  const real sp = star ? pS : (left ? pL : pR);
  const real E = star ? ES : (left ? EL : ER);
  const bool vw_left = star ? kl : left;
  const real sv = vw_left ? vL : vR;
  const real sw = vw_left ? wL : wR;
  pout = sp;
  F[0] = sr * su;
  F[1] = sr * su * su + sp;
  F[2] = F[0] * sv;
  F[3] = F[0] * sw;
  F[4] = (E + sp) * su;
}

// RiemannSolvers.h:5-36 with computeFlux (States.h:113-127) and primToCons
// (States.h:20-31).  p/(gamma0-1) appears in both and is computed once.
template <typename real, bool FAST = false>
__device__ __forceinline__ void hll(const Prim<real> L, const Prim<real> R, real gamma, real gm1,
                                    real F[5], real& pout) {
  using A = Ar<real, FAST>;
  const real aL = sound_speed<real, FAST>(L.r, L.p, gamma);
  const real aR = sound_speed<real, FAST>(R.r, R.p, gamma);
  const real SL = fmin(L.u - aL, R.u - aR);
  const real SR = fmax(L.u + aL, R.u + aR);
  const real ig = A::rcp(gm1);
  const real peL = A::divr(L.p, gm1, ig), peR = A::divr(R.p, gm1, ig);
  const real EkL = (real)0.5 * L.r * (L.u * L.u + L.v * L.v + L.w * L.w);
  const real EkR = (real)0.5 * R.r * (R.u * R.u + R.v * R.v + R.w * R.w);
  const real EfL = (peL + EkL), EfR = (peR + EkR);
  real FL[5], FR[5];
  FL[0] = L.r * L.u; FL[1] = L.r * L.u * L.u + L.p; FL[2] = L.r * L.u * L.v;
  FL[3] = L.r * L.u * L.w; FL[4] = (L.p + EfL) * L.u;
  FR[0] = R.r * R.u; FR[1] = R.r * R.u * R.u + R.p; FR[2] = R.r * R.u * R.v;
  FR[3] = R.r * R.u * R.w; FR[4] = (R.p + EfR) * R.u;
  if (SL >= (real)0) {
#pragma unroll
    for (int n = 0; n < 5; ++n) F[n] = FL[n];
    pout = L.p;
  } else if (SR <= (real)0) {
#pragma unroll
    for (int n = 0; n < 5; ++n) F[n] = FR[n];
    pout = R.p;
  } else {
    real UL[5], UR[5];
    UL[0] = L.r; UL[1] = L.r * L.u; UL[2] = L.r * L.v; UL[3] = L.r * L.w;
    UL[4] = (A::div((real)0.5 * (UL[1] * UL[1] + UL[2] * UL[2] + UL[3] * UL[3]), L.r) + peL);
    UR[0] = R.r; UR[1] = R.r * R.u; UR[2] = R.r * R.v; UR[3] = R.r * R.w;
    UR[4] = (A::div((real)0.5 * (UR[1] * UR[1] + UR[2] * UR[2] + UR[3] * UR[3]), R.r) + peR);
    pout = (real)0.5 * (L.p + R.p);
    const real SLSR = SL * SR, dS = SR - SL, iS = A::rcp(dS);
#pragma unroll
    for (int n = 0; n < 5; ++n)
      F[n] = A::divr(FL[n] * SR - FR[n] * SL + (UR[n] - UL[n]) * SLSR, dS, iS);
  }
}

template <typename real, int RS, bool FAST = false>
__device__ __forceinline__ void riemann(const Prim<real> L, const Prim<real> R, real gamma,
                                        real gm1, real entho, real F[5], real& pout) {
  if (RS == FV3D_HLL) hll<real, FAST>(L, R, gamma, gm1, F, pout);
  else hllc(L, R, gamma, entho, F, pout);  // FAST applies to HLL only
}

// swap_component (States.h:103-111) on a state in the face frame
template <typename real, int DIR>
__device__ __forceinline__ Prim<real> rotate(const Prim<real>& q) {
  if (DIR == FV3D_IX) return q;
  if (DIR == FV3D_IY) return Prim<real>{q.r, q.v, q.u, q.w, q.p};
  return Prim<real>{q.r, q.w, q.v, q.u, q.p};
}

template <typename real, int DIR>
__device__ __forceinline__ void unrotate_flux(real F[5]) {
  if (DIR == FV3D_IY) { real t = F[1]; F[1] = F[2]; F[2] = t; }
  if (DIR == FV3D_IZ) { real t = F[1]; F[1] = F[3]; F[3] = t; }
}

// ---------------------------------------------------------------------------
// ghost maps.  fillBoundaries (BoundaryConditions.h:146-215) runs X, Y, Z
// passes; for absorbing/reflecting/periodic each pass changes one index and
// at most negates one velocity, so the value of any ghost (edges and corners
// included) is the cell reached by mapping each coordinate on its own, with
// the normal velocity of every reflected axis negated.  C91 in x or y takes
// the periodic (default) branch of the X and Y passes (BoundaryConditions.h:165,186).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int map_coord(int i, int N, int bc, bool& flip) {
  flip = false;
  int s = i;
  if (i < 0) {
    if (bc == FV3D_BC_ABSORBING) s = 0;
    else if (bc == FV3D_BC_REFLECTING) { s = -i - 1; flip = true; }
    else s = i + N;
  } else if (i >= N) {
    if (bc == FV3D_BC_ABSORBING) s = N - 1;
    else if (bc == FV3D_BC_REFLECTING) { s = 2 * N - i - 1; flip = true; }
    else s = i - N;
  }
  // cells further out than the 2-deep ghost shell only occur in the overhang
  // of partial tiles and are never used; keep their address in range
  s = s < 0 ? 0 : (s >= N ? N - 1 : s);
  return s;
}

__device__ __forceinline__ double u64_as_double(unsigned long long b) {
  return __longlong_as_double((long long)b);
}
__device__ __forceinline__ unsigned long long double_as_u64(double d) {
  return (unsigned long long)__double_as_longlong(d);
}

// dt = CFL / max(inv_dt_hyp, inv_dt_par_tc, inv_dt_par_visc)  (ComputeDt.h:62)
__device__ __forceinline__ double dt_from_inv(double ih, const DevParams& P) {
  if (P.dt_direct) return ih;
  double m = ih;
  if (m < P.inv_tc) m = P.inv_tc;
  if (m < P.inv_visc) m = P.inv_visc;
  return P.CFL / m;
}

}  // namespace fv3d_hip
