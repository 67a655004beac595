// fv3d_adjoint_impl.h -- discrete adjoint of one stage of the fv3d update.
//
// A stage is S(U) = U + dt R(prim(U)) with dt frozen (taken from the
// forward run), R the flux divergence of the reconstructed Riemann fluxes
// (Update.h:90-162) plus conduction (ThermalConduction.h:24-98) and
// viscosity (Viscosity.h:14-155).  Given the adjoint v of the stage output
// these kernels form
//     out = sigma*v + J_S(U)^T (sigma*v) - sigma*v + alpha*extra
//         = sigma*v + dt J_R(U)^T (sigma*v) + alpha*extra
// in three steps:
//   k_adj_faces      one thread per face instance: recomputes the face
//                    (slopes with their minmod branches, face states, the
//                    Riemann branch) and scatters the vector-Jacobian
//                    product into a primitive-space accumulator Qb
//                    [6][N] = (rho, u, v, w, p, T) at the four stencil cells,
//                    through the ghost maps (reflected ghosts negate the
//                    normal velocity, so does their adjoint);
//   k_adj_diffusion  one thread per cell: the transposes of the conduction
//                    (into the T slot) and viscosity stencils, including the
//                    reference's viscosity index quirks, and the parameter
//                    gradients dS/dkappa, dS/dmu dotted with sigma*v;
//   k_adj_finalize   one thread per cell: Qb -> conserved through the
//                    transpose of consToPrim (T = p/rho first), plus the
//                    identity and the RK2 weights; clears Qb.
// Non-smooth points keep the branch the forward evaluation takes (minmod
// selection, HLL/HLLC wave region, fmin/fmax operands).  The density floor
// (Update.h:160) is taken as inactive.  Accumulation into Qb uses atomics,
// so sums may associate differently from run to run.
#pragma once

#include <algorithm>
#include <atomic>
#include <type_traits>

#include "fv3d_device.h"
#include "fv3d_kernels.h"

namespace fv3d_hip {

// ---------------------------------------------------------------------------
// reverse Riemann solvers (face frame, arrays r,u,v,w,p); Lb/Rb accumulate
// ---------------------------------------------------------------------------
template <typename real>
__device__ __forceinline__ void hllc_vjp(const real L[5], const real R[5], real gamma, real entho,
                                         const real Fb[5], real Lb[5], real Rb[5]) {
  const real rL = L[0], uL = L[1], vL = L[2], pL = L[4];
  const real rR = R[0], uR = R[1], vR = R[2], pR = R[4];
  const real h = (real)0.5;
  const real EL = h * rL * (uL * uL + vL * vL) + pL * entho;
  const real ER = h * rR * (uR * uR + vR * vR) + pR * entho;
  const real cL = sqrt(pL * gamma / rL), cR = sqrt(pR * gamma / rR);
  const real cmax = fmax(cL, cR);
  const real SL = fmin(uL, uR) - cmax, SR = fmax(uL, uR) + cmax;
  const real rcL = rL * (uL - SL), rcR = rR * (SR - uR);
  // Quotients over den and D go through one reciprocal each: within an ulp
  // of the forward values, and the branch tests (signs, SL, SR) are exact.
  const real den = rcR + rcL, iden = (real)1 / den;
  const real uS = (rcR * uR + rcL * uL + (pL - pR)) * iden;
  const real pS = (rcR * pL + rcL * pR + rcL * rcR * (uL - uR)) * iden;
  const bool kl = uS > (real)0;
  const real rK = kl ? rL : rR, uK = kl ? uL : uR, pK = kl ? pL : pR, EK = kl ? EL : ER,
             SK = kl ? SL : SR;
  const real A = SK - uK, D = SK - uS, iD = (real)1 / D;
  const real rS = rK * A * iD;
  const real ES = (A * EK - pK * uK + pS * uS) * iD;
  const bool left = SL > (real)0;
  const bool star = !left && (kl || SR > (real)0);
  const real sr = star ? rS : (left ? rL : rR);
  const real su = star ? uS : (left ? uL : uR);
  const real sp = star ? pS : (left ? pL : pR);
  const real E = star ? ES : (left ? EL : ER);
  const bool vw_left = star ? kl : left;
  const real sv = vw_left ? vL : R[2];
  const real sw = vw_left ? L[3] : R[3];
  const real F0 = sr * su;

  // F = (sr su, sr su^2 + sp, F0 sv, F0 sw, (E + sp) su)
  const real tv = Fb[2] * sv + Fb[3] * sw;
  const real srb = Fb[0] * su + Fb[1] * su * su + tv * su;
  const real sub = Fb[0] * sr + Fb[1] * (real)2 * sr * su + tv * sr + Fb[4] * (E + sp);
  const real spb = Fb[1] + Fb[4] * su;
  const real Eb = Fb[4] * su;
  // Every side choice below is a select on values: a branch that adds to Lb
  // on one side and Rb on the other becomes a select between the two
  // arrays' addresses, which puts both in scratch memory.
  const real z = (real)0;
  const real t2 = Fb[2] * F0, t3 = Fb[3] * F0;
  Lb[2] += vw_left ? t2 : z; Lb[3] += vw_left ? t3 : z;
  Rb[2] += vw_left ? z : t2; Rb[3] += vw_left ? z : t3;

  // star region: the adjoints of rS = rK A / D and ES = (A EK - pK uK +
  // pS uS) / D with A = SK - uK, D = SK - uS (evaluated in every region,
  // used only in the star region)
  real Ab = srb * rK * iD, Db = -srb * rS * iD;
  const real rKb = srb * A * iD;
  const real NEb = Eb * iD;
  Db -= Eb * ES * iD;
  Ab += NEb * EK;
  const real EKb = NEb * A;
  const real pKb = -NEb * uK;
  const real uKb = -NEb * pK - Ab;
  const real SKb = Ab + Db;
  const real uSb = star ? sub + NEb * pS - Db : z;
  const real pSb = star ? spb + NEb * uS : z;
  // the side that receives the state adjoints: K in the star region, else
  // the upwind side
  const bool toL = star ? kl : left;
  const real c0 = star ? rKb : srb, c1 = star ? uKb : sub, c4 = star ? pKb : spb;
  const real cE = star ? EKb : Eb, cS = star ? SKb : z;
  Lb[0] += toL ? c0 : z; Lb[1] += toL ? c1 : z; Lb[4] += toL ? c4 : z;
  Rb[0] += toL ? z : c0; Rb[1] += toL ? z : c1; Rb[4] += toL ? z : c4;
  const real ELb = toL ? cE : z, ERb = toL ? z : cE;
  real SLb = toL ? cS : z, SRb = toL ? z : cS;
  const real NUb = uSb * iden, NPb = pSb * iden;
  const real denb = -(uSb * uS + pSb * pS) * iden;
  real rcRb = NUb * uR + denb, rcLb = NUb * uL + denb;
  Rb[1] += NUb * rcR;
  Lb[1] += NUb * rcL;
  Lb[4] += NUb;
  Rb[4] -= NUb;
  const real du = uL - uR;
  rcRb += NPb * (pL + rcL * du);
  rcLb += NPb * (pR + rcR * du);
  Lb[4] += NPb * rcR;
  Rb[4] += NPb * rcL;
  Lb[1] += NPb * rcL * rcR;
  Rb[1] -= NPb * rcL * rcR;
  Lb[0] += rcLb * (uL - SL);
  Lb[1] += rcLb * rL;
  SLb -= rcLb * rL;
  Rb[0] += rcRb * (SR - uR);
  SRb += rcRb * rR;
  Rb[1] -= rcRb * rR;
  const real cmaxb = SRb - SLb;
  Lb[1] += (uL < uR ? SLb : z) + (uL > uR ? SRb : z);
  Rb[1] += (uL < uR ? z : SLb) + (uL > uR ? z : SRb);
  // cmax = c of one side M: dc/dp = gamma / (2 r c), dc/dr = -c / (2 r)
  const bool cl = cL >= cR;
  const real i2r = h / (cl ? rL : rR), i2rc = i2r / cmax;
  const real b4 = cmaxb * gamma * i2rc, b0 = cmaxb * cmax * i2r;
  Lb[4] += cl ? b4 : z; Lb[0] -= cl ? b0 : z;
  Rb[4] += cl ? z : b4; Rb[0] -= cl ? z : b0;
  // E = 0.5 r (u^2 + v^2) + p entho: the second transverse velocity is
  // dropped (RiemannSolvers.h:52-55), so it has no adjoint here
  Lb[0] += ELb * h * (uL * uL + vL * vL); Lb[1] += ELb * rL * uL; Lb[2] += ELb * rL * vL;
  Lb[4] += ELb * entho;
  Rb[0] += ERb * h * (uR * uR + vR * vR); Rb[1] += ERb * rR * uR; Rb[2] += ERb * rR * vR;
  Rb[4] += ERb * entho;
}

// Euler flux of a primitive state (States.h:113-127) reversed
template <typename real>
__device__ __forceinline__ void euler_flux_vjp(const real q[5], real gm1, const real Fb[5],
                                               real qb[5]) {
  const real r = q[0], u = q[1], v = q[2], w = q[3], p = q[4];
  const real k2 = u * u + v * v + w * w;
  const real Ef = p / gm1 + (real)0.5 * r * k2;
  qb[0] += Fb[0] * u + Fb[1] * u * u + Fb[2] * u * v + Fb[3] * u * w + Fb[4] * u * (real)0.5 * k2;
  qb[1] += Fb[0] * r + Fb[1] * (real)2 * r * u + Fb[2] * r * v + Fb[3] * r * w +
           Fb[4] * ((p + Ef) + u * r * u);
  qb[2] += Fb[2] * r * u + Fb[4] * u * r * v;
  qb[3] += Fb[3] * r * u + Fb[4] * u * r * w;
  qb[4] += Fb[1] + Fb[4] * u * ((real)1 + (real)1 / gm1);
}

// conserved variables of a primitive state (States.h:20-31) reversed
template <typename real>
__device__ __forceinline__ void prim_to_cons_vjp(const real q[5], real gm1, const real Ub[5],
                                                 real qb[5]) {
  const real r = q[0], u = q[1], v = q[2], w = q[3];
  qb[0] += Ub[0] + Ub[1] * u + Ub[2] * v + Ub[3] * w + Ub[4] * (real)0.5 * (u * u + v * v + w * w);
  qb[1] += Ub[1] * r + Ub[4] * r * u;
  qb[2] += Ub[2] * r + Ub[4] * r * v;
  qb[3] += Ub[3] * r + Ub[4] * r * w;
  qb[4] += Ub[4] / gm1;
}

template <typename real>
__device__ __forceinline__ void hll_vjp(const real L[5], const real R[5], real gamma, real gm1,
                                        const real Fb[5], real Lb[5], real Rb[5]) {
  const real aL = sqrt(L[4] * gamma / L[0]), aR = sqrt(R[4] * gamma / R[0]);
  const real xL = L[1] - aL, xR = R[1] - aR, yL = L[1] + aL, yR = R[1] + aR;
  const real SL = fmin(xL, xR), SR = fmax(yL, yR);
  if (SL >= (real)0) { euler_flux_vjp(L, gm1, Fb, Lb); return; }
  if (SR <= (real)0) { euler_flux_vjp(R, gm1, Fb, Rb); return; }
  real FL[5], FR[5], UL[5], UR[5];
  auto flux = [&](const real q[5], real F[5], real U[5]) {
    const real r = q[0], u = q[1], v = q[2], w = q[3], p = q[4];
    const real Ef = p / gm1 + (real)0.5 * r * (u * u + v * v + w * w);
    F[0] = r * u; F[1] = r * u * u + p; F[2] = r * u * v; F[3] = r * u * w; F[4] = (p + Ef) * u;
    U[0] = r; U[1] = r * u; U[2] = r * v; U[3] = r * w; U[4] = Ef;
  };
  flux(L, FL, UL);
  flux(R, FR, UR);
  const real dS = SR - SL;
  real FLb[5], FRb[5], ULb[5], URb[5];
  real SLb = 0, SRb = 0;
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const real dU = UR[n] - UL[n];
    const real F = (FL[n] * SR - FR[n] * SL + dU * SL * SR) / dS;
    FLb[n] = Fb[n] * SR / dS;
    FRb[n] = -Fb[n] * SL / dS;
    URb[n] = Fb[n] * SL * SR / dS;
    ULb[n] = -URb[n];
    SRb += Fb[n] * (FL[n] + dU * SL - F) / dS;
    SLb += Fb[n] * (-FR[n] + dU * SR + F) / dS;
  }
  euler_flux_vjp(L, gm1, FLb, Lb);
  euler_flux_vjp(R, gm1, FRb, Rb);
  prim_to_cons_vjp(L, gm1, ULb, Lb);
  prim_to_cons_vjp(R, gm1, URb, Rb);
  real aLb = 0, aRb = 0;
  if (xL < xR) { Lb[1] += SLb; aLb -= SLb; } else { Rb[1] += SLb; aRb -= SLb; }
  if (yL > yR) { Lb[1] += SRb; aLb += SRb; } else { Rb[1] += SRb; aRb += SRb; }
  Lb[4] += aLb * gamma / ((real)2 * L[0] * aL);
  Lb[0] -= aLb * aL / ((real)2 * L[0]);
  Rb[4] += aRb * gamma / ((real)2 * R[0] * aR);
  Rb[0] -= aRb * aR / ((real)2 * R[0]);
}

// minmod(a, b) branch: 0 -> zero, 1 -> a, 2 -> b (same tests as minmod)
template <typename real>
__device__ __forceinline__ int minmod_branch(real a, real b) {
  return (a * b < (real)0) ? 0 : ((fabs(a) < fabs(b)) ? 1 : 2);
}
template <typename real>
__device__ __forceinline__ real minmod_sel(real a, real b, int br) {
  return br == 0 ? (real)0 : (br == 1 ? a : b);
}

// Slope at b from a, b, c (x, y: minmod(back, fwd); z: minmod(fwd, back)),
// with the selection of every variable: +1 backward, -1 forward, 0 zero.
template <typename real, bool ZORD>
__device__ __forceinline__ void slope_sel(const real a[5], const real b[5], const real c[5],
                                          real s[5], int sel[5]) {
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const real bk = b[n] - a[n], fw = c[n] - b[n];
    if (!ZORD) {
      const int br = minmod_branch(bk, fw);
      s[n] = minmod_sel(bk, fw, br);
      sel[n] = br == 0 ? 0 : (br == 1 ? 1 : -1);
    } else {
      const int br = minmod_branch(fw, bk);
      s[n] = minmod_sel(fw, bk, br);
      sel[n] = br == 0 ? 0 : (br == 1 ? -1 : 1);
    }
  }
}
template <typename real>
__device__ __forceinline__ void slope_vjp(const int sel[5], const real sb[5], real ab[5],
                                          real bb[5], real cb[5]) {
  // selects on values: branches that update different arrays would put
  // the arrays in scratch memory
  const real z = (real)0;
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const real x = sb[n];
    const bool bw = sel[n] > 0, fw = sel[n] < 0;
    ab[n] -= bw ? x : z;
    bb[n] += bw ? x : (fw ? -x : z);
    cb[n] += fw ? x : z;
  }
}

template <int DIR, typename real>
__device__ __forceinline__ void swap_dir(real x[5]) {
  if (DIR == FV3D_IY) { const real t = x[1]; x[1] = x[2]; x[2] = t; }
  if (DIR == FV3D_IZ) { const real t = x[1]; x[1] = x[3]; x[3] = t; }
}

template <typename real>
__device__ __forceinline__ void load_prim_arr(const real* __restrict__ U, long long cell,
                                              long long sf, real gm1, real q[5]) {
  const Prim<real> p = cons_to_prim(U[cell], U[cell + sf], U[cell + 2 * sf], U[cell + 3 * sf],
                                    U[cell + 4 * sf], gm1);
  q[0] = p.r; q[1] = p.u; q[2] = p.v; q[3] = p.w; q[4] = p.p;
}

template <typename real>
__device__ __forceinline__ void atomic_add(real* p, real v) {
  if (v != (real)0) unsafeAtomicAdd(p, v);
}

// ---------------------------------------------------------------------------
// faces.  Instance m+1/2 along DIR, m in [-1, N-1], stencil m-1..m+2;
// its flux enters cell m (if inside) with -dt/d and cell m+1 (if inside)
// with +dt/d.
// ---------------------------------------------------------------------------
template <typename real, int DIR, int RECON, int RS>
__global__ void __launch_bounds__(256)
k_adj_faces(const real* __restrict__ U, const real* __restrict__ v, real sigma, real* Qb,
            double dt, DevParams P) {
  const int Nd = DIR == FV3D_IX ? P.Nx : (DIR == FV3D_IY ? P.Ny : P.Nz);
  const long long nf = (long long)(P.Nx + (DIR == FV3D_IX)) * (P.Ny + (DIR == FV3D_IY)) *
                       (P.Nz + (DIR == FV3D_IZ));
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= nf) return;
  const int ex = P.Nx + (DIR == FV3D_IX), ey = P.Ny + (DIR == FV3D_IY);
  const int fi = (int)(t % ex), fj = (int)((t / ex) % ey), fk = (int)(t / ((long long)ex * ey));
  const int m = (DIR == FV3D_IX ? fi : (DIR == FV3D_IY ? fj : fk)) - 1;
  const int bc = DIR == FV3D_IX ? P.bcx : (DIR == FV3D_IY ? P.bcy : P.bcz);
  const long long stride = DIR == FV3D_IX ? 1 : (DIR == FV3D_IY ? P.sy : P.sz);
  const long long base = (DIR == FV3D_IX ? 0 : fi) + (DIR == FV3D_IY ? 0 : (long long)fj * P.sy) +
                         (DIR == FV3D_IZ ? 0 : (long long)fk * P.sz);
  const real gm1 = (real)P.gm1;
  real q[4][5];
  long long src[4];
  bool flip[4];
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    const int s = map_coord(m - 1 + o, Nd, bc, flip[o]);
    src[o] = base + (long long)s * stride;
    load_prim_arr(U, src[o], P.sf, gm1, q[o]);
    if (flip[o]) q[o][1 + DIR] = -q[o][1 + DIR];
  }
  // adjoint of the face flux (lab frame)
  const double d = DIR == FV3D_IX ? P.dx : (DIR == FV3D_IY ? P.dy : P.dz);
  const real coef = (real)((double)sigma * dt / d);
  real Fb[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const real vc = (m + 1 < Nd) ? v[src[2] + n * P.sf] : (real)0;
    const real vb = (m >= 0) ? v[src[1] + n * P.sf] : (real)0;
    Fb[n] = coef * (vc - vb);
  }
  real qL[5], qR[5], sb[5], sc[5];
  int selb[5] = {0, 0, 0, 0, 0}, selc[5] = {0, 0, 0, 0, 0};
  if (RECON == FV3D_PLM) {
    slope_sel<real, DIR == FV3D_IZ>(q[0], q[1], q[2], sb, selb);
    slope_sel<real, DIR == FV3D_IZ>(q[1], q[2], q[3], sc, selc);
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      qL[n] = q[1][n] + sb[n] * (real)0.5;
      qR[n] = q[2][n] - sc[n] * (real)0.5;
    }
  } else {
#pragma unroll
    for (int n = 0; n < 5; ++n) { qL[n] = q[1][n]; qR[n] = q[2][n]; }
  }
  swap_dir<DIR>(qL);
  swap_dir<DIR>(qR);
  swap_dir<DIR>(Fb);
  real Lb[5] = {0, 0, 0, 0, 0}, Rb[5] = {0, 0, 0, 0, 0};
  if (RS == FV3D_HLL) hll_vjp(qL, qR, (real)P.gamma, gm1, Fb, Lb, Rb);
  else hllc_vjp(qL, qR, (real)P.gamma, (real)P.entho, Fb, Lb, Rb);
  swap_dir<DIR>(Lb);
  swap_dir<DIR>(Rb);
  real qb[4][5];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    qb[0][n] = 0; qb[1][n] = Lb[n]; qb[2][n] = Rb[n]; qb[3][n] = 0;
  }
  if (RECON == FV3D_PLM) {
    real ssb[5], ssc[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) { ssb[n] = (real)0.5 * Lb[n]; ssc[n] = (real)-0.5 * Rb[n]; }
    slope_vjp(selb, ssb, qb[0], qb[1], qb[2]);
    slope_vjp(selc, ssc, qb[1], qb[2], qb[3]);
  }
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    if (flip[o]) qb[o][1 + DIR] = -qb[o][1 + DIR];
#pragma unroll
    for (int n = 0; n < 5; ++n) atomic_add(&Qb[src[o] + n * P.sf], qb[o][n]);
  }
}

// ---------------------------------------------------------------------------
// conduction and viscosity (one thread per cell)
// ---------------------------------------------------------------------------

// The viscous update of one cell, Viscosity.h:28-153 with mu = 1, as a
// function of its 3x3x3 velocity stencil S(dk, dj, di, c) (c = u, v, w):
// out[0..3] = increments of rho u, rho v, rho w, E (before * mu * dt).
// In reverse mode (REV) it hands d(out . g)/dS to SB(dk, dj, di, c, value)
// instead, one term at a time.
// One side s (1: lower, 2: upper) of one direction BLK (0: x, 1: y, 2: z) of
// the viscous update of a cell (Viscosity.h:28-153, mu = 1): adds
// sgn(s) * (flux of that side) to out[0..3] (rho u, rho v, rho w, E), or in
// reverse mode (REV) hands d(g . sgn(s) flux)/dS to SB(dk, dj, di, c,
// value), one term at a time.  S(dk, dj, di, c) is the cell's 3x3x3
// velocity stencil (c = u, v, w).  The x side s=2 of cell i and the x side
// s=1 of cell i+1 are the same flux (likewise in y); the z sides carry the
// reference's index and component mix-ups and are not shared.
template <typename real, bool REV, int BLK, class SF, class SBF>
__device__ __forceinline__ void visc_side(SF&& S, const int s, real oodx, real oody, real oodz,
                                          const real g[4], real out[4], SBF&& SB) {
  const real f43 = (real)(4.0 / 3.0), f23 = (real)(2.0 / 3.0), h = (real)0.5, q4 = (real)0.25;
  const real sgn = s == 1 ? (real)-1 : (real)1;
  if (BLK == 0) {
  // ---- X: S(1,1,s) and S(1,1,s-1), y and z differences around them
  {
    const real qiu = h * (S(1, 1, s, 0) + S(1, 1, s - 1, 0));
    const real qiv = h * (S(1, 1, s, 1) + S(1, 1, s - 1, 1));
    const real qiw = h * (S(1, 1, s, 2) + S(1, 1, s - 1, 2));
    const real dudx = oodx * (S(1, 1, s, 0) - S(1, 1, s - 1, 0));
    const real dvdx = oodx * (S(1, 1, s, 1) - S(1, 1, s - 1, 1));
    const real dwdx = oodx * (S(1, 1, s, 2) - S(1, 1, s - 1, 2));
    const real dudy = q4 * oody * (S(1, 2, s, 0) - S(1, 0, s, 0) + S(1, 2, s - 1, 0) - S(1, 0, s - 1, 0));
    const real dvdy = q4 * oody * (S(1, 2, s, 1) - S(1, 0, s, 1) + S(1, 2, s - 1, 1) - S(1, 0, s - 1, 1));
    const real dudz = q4 * oodz * (S(2, 1, s, 0) - S(0, 1, s, 0) + S(2, 1, s - 1, 0) - S(0, 1, s - 1, 0));
    const real dwdz = q4 * oodz * (S(2, 1, s, 2) - S(0, 1, s, 2) + S(2, 1, s - 1, 2) - S(0, 1, s - 1, 2));
    const real txx = f43 * dudx - f23 * (dvdy + dwdz);
    const real txy = dvdx + dudy;
    const real txz = dwdx + dudz;
    out[0] += sgn * txx; out[1] += sgn * txy; out[2] += sgn * txz;
    out[3] += sgn * (txx * qiu + txy * qiv + txz * qiw);
    if (REV) {
      const real txxb = sgn * (g[0] + g[3] * qiu), txyb = sgn * (g[1] + g[3] * qiv),
                 txzb = sgn * (g[2] + g[3] * qiw);
      const real qiub = sgn * g[3] * txx * h, qivb = sgn * g[3] * txy * h, qiwb = sgn * g[3] * txz * h;
      SB(1, 1, s, 0, qiub); SB(1, 1, s - 1, 0, qiub);
      SB(1, 1, s, 1, qivb); SB(1, 1, s - 1, 1, qivb);
      SB(1, 1, s, 2, qiwb); SB(1, 1, s - 1, 2, qiwb);
      const real dudxb = f43 * txxb * oodx, dvdyb = -f23 * txxb * q4 * oody, dwdzb = -f23 * txxb * q4 * oodz;
      const real dvdxb = txyb * oodx, dudyb = txyb * q4 * oody;
      const real dwdxb = txzb * oodx, dudzb = txzb * q4 * oodz;
      SB(1, 1, s, 0, dudxb); SB(1, 1, s - 1, 0, -(dudxb));
      SB(1, 1, s, 1, dvdxb); SB(1, 1, s - 1, 1, -(dvdxb));
      SB(1, 1, s, 2, dwdxb); SB(1, 1, s - 1, 2, -(dwdxb));
      SB(1, 2, s, 0, dudyb); SB(1, 0, s, 0, -(dudyb)); SB(1, 2, s - 1, 0, dudyb); SB(1, 0, s - 1, 0, -(dudyb));
      SB(1, 2, s, 1, dvdyb); SB(1, 0, s, 1, -(dvdyb)); SB(1, 2, s - 1, 1, dvdyb); SB(1, 0, s - 1, 1, -(dvdyb));
      SB(2, 1, s, 0, dudzb); SB(0, 1, s, 0, -(dudzb)); SB(2, 1, s - 1, 0, dudzb); SB(0, 1, s - 1, 0, -(dudzb));
      SB(2, 1, s, 2, dwdzb); SB(0, 1, s, 2, -(dwdzb)); SB(2, 1, s - 1, 2, dwdzb); SB(0, 1, s - 1, 2, -(dwdzb));
    }
  }
  } else if (BLK == 1) {
  // ---- Y: S(1,s,1) and S(1,s-1,1)
  {
    const real qiu = h * (S(1, s, 1, 0) + S(1, s - 1, 1, 0));
    const real qiv = h * (S(1, s, 1, 1) + S(1, s - 1, 1, 1));
    const real qiw = h * (S(1, s, 1, 2) + S(1, s - 1, 1, 2));
    const real dudy = oody * (S(1, s, 1, 0) - S(1, s - 1, 1, 0));
    const real dvdy = oody * (S(1, s, 1, 1) - S(1, s - 1, 1, 1));
    const real dwdy = oody * (S(1, s, 1, 2) - S(1, s - 1, 1, 2));
    const real dudx = q4 * oodx * (S(1, s, 2, 0) - S(1, s, 0, 0) + S(1, s - 1, 2, 0) - S(1, s - 1, 0, 0));
    const real dvdx = q4 * oodx * (S(1, s, 2, 1) - S(1, s, 0, 1) + S(1, s - 1, 2, 1) - S(1, s - 1, 0, 1));
    const real dwdz = q4 * oodz * (S(2, s, 1, 2) - S(0, s, 1, 2) + S(2, s - 1, 1, 2) - S(0, s - 1, 1, 2));
    const real dvdz = q4 * oodz * (S(2, s, 1, 1) - S(0, s, 1, 1) + S(2, s - 1, 1, 1) - S(0, s - 1, 1, 1));
    const real tyy = f43 * dvdy - f23 * (dudx + dwdz);
    const real txy = dvdx + dudy;
    const real tyz = dvdz + dwdy;
    out[0] += sgn * txy; out[1] += sgn * tyy; out[2] += sgn * tyz;
    out[3] += sgn * (txy * qiu + tyy * qiv + tyz * qiw);
    if (REV) {
      const real txyb = sgn * (g[0] + g[3] * qiu), tyyb = sgn * (g[1] + g[3] * qiv),
                 tyzb = sgn * (g[2] + g[3] * qiw);
      const real qiub = sgn * g[3] * txy * h, qivb = sgn * g[3] * tyy * h, qiwb = sgn * g[3] * tyz * h;
      SB(1, s, 1, 0, qiub); SB(1, s - 1, 1, 0, qiub);
      SB(1, s, 1, 1, qivb); SB(1, s - 1, 1, 1, qivb);
      SB(1, s, 1, 2, qiwb); SB(1, s - 1, 1, 2, qiwb);
      const real dvdyb = f43 * tyyb * oody, dudxb = -f23 * tyyb * q4 * oodx, dwdzb = -f23 * tyyb * q4 * oodz;
      const real dvdxb = txyb * q4 * oodx, dudyb = txyb * oody;
      const real dvdzb = tyzb * q4 * oodz, dwdyb = tyzb * oody;
      SB(1, s, 1, 0, dudyb); SB(1, s - 1, 1, 0, -(dudyb));
      SB(1, s, 1, 1, dvdyb); SB(1, s - 1, 1, 1, -(dvdyb));
      SB(1, s, 1, 2, dwdyb); SB(1, s - 1, 1, 2, -(dwdyb));
      SB(1, s, 2, 0, dudxb); SB(1, s, 0, 0, -(dudxb)); SB(1, s - 1, 2, 0, dudxb); SB(1, s - 1, 0, 0, -(dudxb));
      SB(1, s, 2, 1, dvdxb); SB(1, s, 0, 1, -(dvdxb)); SB(1, s - 1, 2, 1, dvdxb); SB(1, s - 1, 0, 1, -(dvdxb));
      SB(2, s, 1, 2, dwdzb); SB(0, s, 1, 2, -(dwdzb)); SB(2, s - 1, 1, 2, dwdzb); SB(0, s - 1, 1, 2, -(dwdzb));
      SB(2, s, 1, 1, dvdzb); SB(0, s, 1, 1, -(dvdzb)); SB(2, s - 1, 1, 1, dvdzb); SB(0, s - 1, 1, 1, -(dvdzb));
    }
  }
  } else {
  // ---- Z, with the reference's index and component mix-ups
  // (Viscosity.h:122-133): S(s,1,1) and S(s-1,1,1)
  {
    const real qiu = h * (S(s, 1, 1, 0) + S(s - 1, 1, 1, 0));
    const real qiv = h * (S(s, 1, 1, 1) + S(s - 1, 1, 1, 1));
    const real qiw = h * (S(s, 1, 1, 2) + S(s - 1, 1, 1, 2));
    const real dudz = oodz * (S(s, 1, 1, 0) - S(s - 1, 1, 1, 0));
    const real dvdz = oodz * (S(s, 1, 1, 1) - S(s - 1, 1, 1, 1));
    const real dwdz = oodz * (S(s, 1, 1, 2) - S(s - 1, 1, 1, 2));
    const real dudx = q4 * oodx * (S(s, 1, 2, 0) - S(1, s, 0, 0) + S(s - 1, 1, 2, 0) - S(1, s - 1, 0, 0));
    const real dwdx = q4 * oodx * (S(s, 1, 2, 1) - S(1, s, 0, 2) + S(s - 1, 1, 2, 1) - S(1, s - 1, 0, 2));
    const real dvdy = q4 * oodz * (S(s, 2, 1, 2) - S(s, 0, 1, 1) + S(s - 1, 2, 1, 2) - S(s - 1, 0, 1, 1));
    const real dwdy = q4 * oodz * (S(s, 2, 1, 1) - S(s, 0, 1, 2) + S(s - 1, 2, 1, 1) - S(s - 1, 0, 1, 2));
    const real tzz = f43 * dwdz - f23 * (dudx + dvdy);
    const real txz = dudz + dwdx;
    const real tyz = dvdz + dwdy;
    out[0] += sgn * txz; out[1] += sgn * tyz; out[2] += sgn * tzz;
    out[3] += sgn * (txz * qiu + tyz * qiv + tzz * qiw);
    if (REV) {
      const real txzb = sgn * (g[0] + g[3] * qiu), tyzb = sgn * (g[1] + g[3] * qiv),
                 tzzb = sgn * (g[2] + g[3] * qiw);
      const real qiub = sgn * g[3] * txz * h, qivb = sgn * g[3] * tyz * h, qiwb = sgn * g[3] * tzz * h;
      SB(s, 1, 1, 0, qiub); SB(s - 1, 1, 1, 0, qiub);
      SB(s, 1, 1, 1, qivb); SB(s - 1, 1, 1, 1, qivb);
      SB(s, 1, 1, 2, qiwb); SB(s - 1, 1, 1, 2, qiwb);
      const real dwdzb = f43 * tzzb * oodz, dudxb = -f23 * tzzb * q4 * oodx, dvdyb = -f23 * tzzb * q4 * oodz;
      const real dudzb = txzb * oodz, dwdxb = txzb * q4 * oodx;
      const real dvdzb = tyzb * oodz, dwdyb = tyzb * q4 * oodz;
      SB(s, 1, 1, 0, dudzb); SB(s - 1, 1, 1, 0, -(dudzb));
      SB(s, 1, 1, 1, dvdzb); SB(s - 1, 1, 1, 1, -(dvdzb));
      SB(s, 1, 1, 2, dwdzb); SB(s - 1, 1, 1, 2, -(dwdzb));
      SB(s, 1, 2, 0, dudxb); SB(1, s, 0, 0, -(dudxb)); SB(s - 1, 1, 2, 0, dudxb); SB(1, s - 1, 0, 0, -(dudxb));
      SB(s, 1, 2, 1, dwdxb); SB(1, s, 0, 2, -(dwdxb)); SB(s - 1, 1, 2, 1, dwdxb); SB(1, s - 1, 0, 2, -(dwdxb));
      SB(s, 2, 1, 2, dvdyb); SB(s, 0, 1, 1, -(dvdyb)); SB(s - 1, 2, 1, 2, dvdyb); SB(s - 1, 0, 1, 1, -(dvdyb));
      SB(s, 2, 1, 1, dwdyb); SB(s, 0, 1, 2, -(dwdyb)); SB(s - 1, 2, 1, 1, dwdyb); SB(s - 1, 0, 1, 2, -(dwdyb));
    }
  }
  }
}

// The viscous update of one cell, Viscosity.h:28-153 with mu = 1, as a
// function of its 3x3x3 velocity stencil S(dk, dj, di, c) (c = u, v, w):
// out[0..3] = increments of rho u, rho v, rho w, E (before * mu * dt).
// In reverse mode (REV) it hands d(out . g)/dS to SB(dk, dj, di, c, value)
// instead, one term at a time.
template <typename real, bool REV, class SF, class SBF>
__device__ __forceinline__ void visc_cell(SF&& S, real oodx, real oody, real oodz, const real g[4],
                                          real out[4], SBF&& SB) {
#pragma unroll
  for (int n = 0; n < 4; ++n) out[n] = 0;
#pragma unroll
  for (int s = 1; s < 3; ++s) {
    visc_side<real, REV, 0>(S, s, oodx, oody, oodz, g, out, SB);
    visc_side<real, REV, 1>(S, s, oodx, oody, oodz, g, out, SB);
    visc_side<real, REV, 2>(S, s, oodx, oody, oodz, g, out, SB);
  }
}

template <typename real, bool TC, bool VISC>
__global__ void __launch_bounds__(256)
k_adj_diffusion(const real* __restrict__ U, const real* __restrict__ v, real sigma, real* Qb,
                double dt, DevParams P, double* grad) {
  const long long n = P.sf;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  double gk = 0.0, gm = 0.0;
  if (t < n) {
    const int i = (int)(t % P.Nx), j = (int)((t / P.Nx) % P.Ny), k = (int)(t / P.sz);
    const real gm1 = (real)P.gm1;
    const real rdt = (real)dt;
    if (TC) {
      // E += dt/dx (FR - FL) + dt/dy (FD - FU) + dt/dz (FB - FF), F = kf dT / d
      // (ThermalConduction.h:47-95); all terms are linear in kf
      const real Eb = sigma * v[t + 4 * P.sf];
      real T[7];
      long long src[7];
      const int off[7][3] = {{0, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
#pragma unroll
      for (int o = 0; o < 7; ++o) {
        bool f;
        const int si = map_coord(i + off[o][0], P.Nx, P.bcx, f);
        const int sj = map_coord(j + off[o][1], P.Ny, P.bcy, f);
        const int sk = map_coord(k + off[o][2], P.Nz, P.bcz, f);
        src[o] = si + (long long)sj * P.sy + (long long)sk * P.sz;
        real q[5];
        load_prim_arr(U, src[o], P.sf, gm1, q);
        T[o] = q[4] / q[0];
      }
      const real ax = rdt / (real)(P.dx * P.dx), ay = rdt / (real)(P.dy * P.dy),
                 az = rdt / (real)(P.dz * P.dz);
      // coefficients of dE/dT per stencil entry, per unit kf
      real c[7] = {-(real)2 * (ax + ay + az), ax, ax, ay, ay, az, az};
      real fixedE = 0;  // kf-proportional part that does not depend on T
      if (k == 0 && P.bctc_zmin != FV3D_BCTC_NONE) {
        // FF is replaced: the -FF term (coefficient -dt/dz) changes
        c[0] += az;  // undo the default FF = (TC - TF)/dz
        c[5] -= az;
        if (P.bctc_zmin == FV3D_BCTC_FIXED_TEMPERATURE) {
          c[0] -= (real)2 * az;
          fixedE += (real)2 * az * (real)P.bctc_zmin_value;
        } else if (P.bctc_zmin == FV3D_BCTC_FIXED_GRADIENT) {
          fixedE -= rdt / (real)P.dz * (real)P.bctc_zmin_value;
        }
      }
      if (k == P.Nz - 1 && P.bctc_zmax != FV3D_BCTC_NONE) {
        c[0] += az;
        c[6] -= az;
        if (P.bctc_zmax == FV3D_BCTC_FIXED_TEMPERATURE) {
          c[0] -= (real)2 * az;
          fixedE += (real)2 * az * (real)P.bctc_zmax_value;
        } else if (P.bctc_zmax == FV3D_BCTC_FIXED_GRADIENT) {
          fixedE += rdt / (real)P.dz * (real)P.bctc_zmax_value;
        }
      }
      real dE1 = fixedE;
#pragma unroll
      for (int o = 0; o < 7; ++o) dE1 += c[o] * T[o];
      gk = (double)Eb * (double)dE1;
      const real kf = (real)P.kface;
#pragma unroll
      for (int o = 0; o < 7; ++o) atomic_add(&Qb[src[o] + 5 * P.sf], Eb * kf * c[o]);
    }
    if (VISC) {
      real S[3][3][3][3], Sb[3][3][3][3];
      long long src[3][3][3];
      bool fl[3][3][3][3];
#pragma unroll
      for (int dk = 0; dk < 3; ++dk)
#pragma unroll
        for (int dj = 0; dj < 3; ++dj)
#pragma unroll
          for (int di = 0; di < 3; ++di) {
            bool fx, fy, fz;
            const int si = map_coord(i + di - 1, P.Nx, P.bcx, fx);
            const int sj = map_coord(j + dj - 1, P.Ny, P.bcy, fy);
            const int sk = map_coord(k + dk - 1, P.Nz, P.bcz, fz);
            const long long s = si + (long long)sj * P.sy + (long long)sk * P.sz;
            src[dk][dj][di] = s;
            const real r = U[s];
            S[dk][dj][di][0] = U[s + P.sf] / r;
            S[dk][dj][di][1] = U[s + 2 * P.sf] / r;
            S[dk][dj][di][2] = U[s + 3 * P.sf] / r;
            fl[dk][dj][di][0] = fx; fl[dk][dj][di][1] = fy; fl[dk][dj][di][2] = fz;
            if (fx) S[dk][dj][di][0] = -S[dk][dj][di][0];
            if (fy) S[dk][dj][di][1] = -S[dk][dj][di][1];
            if (fz) S[dk][dj][di][2] = -S[dk][dj][di][2];
#pragma unroll
            for (int c = 0; c < 3; ++c) Sb[dk][dj][di][c] = 0;
          }
      const real oodx = (real)1 / (real)P.dx, oody = (real)1 / (real)P.dy, oodz = (real)1 / (real)P.dz;
      real vb[4], out[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) vb[c] = sigma * v[t + (c + 1) * P.sf];
      auto Sget = [&](int dk, int dj, int di, int c) -> real { return S[dk][dj][di][c]; };
      auto Sacc = [&](int dk, int dj, int di, int c, real x) { Sb[dk][dj][di][c] += x; };
      visc_cell<real, false>(Sget, oodx, oody, oodz, vb, out, Sacc);
      double gmu = 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) gmu += (double)vb[c] * (double)out[c];
      gm = gmu * dt;
      real g[4];
      const real mdt = (real)P.mu * rdt;
#pragma unroll
      for (int c = 0; c < 4; ++c) g[c] = vb[c] * mdt;
      visc_cell<real, true>(Sget, oodx, oody, oodz, g, out, Sacc);
#pragma unroll
      for (int dk = 0; dk < 3; ++dk)
#pragma unroll
        for (int dj = 0; dj < 3; ++dj)
#pragma unroll
          for (int di = 0; di < 3; ++di)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const real b = fl[dk][dj][di][c] ? -Sb[dk][dj][di][c] : Sb[dk][dj][di][c];
              atomic_add(&Qb[src[dk][dj][di] + (c + 1) * P.sf], b);
            }
    }
  }
  // parameter gradients: wave sums, then one atomic per wave
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    gk += __shfl_xor(gk, off);
    gm += __shfl_xor(gm, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (TC && gk != 0.0) unsafeAtomicAdd(&grad[0], gk);
    if (VISC && gm != 0.0) unsafeAtomicAdd(&grad[1], gm);
  }
}

// ---------------------------------------------------------------------------
// LDS-tiled forms of the two scatter kernels above.  A workgroup stages the
// primitives of its window (tile plus stencil halo) in LDS once, sums the
// contributions of its faces or cells into an LDS accumulator of the same
// window, and adds the window to Qb with one global atomic per value.
// ---------------------------------------------------------------------------
// Tiles are swept in colour passes: along each axis alternate tiles take
// colours 0 and 1, and the last of an odd count takes colour 2.  A window
// reaches at most half a tile beyond its tile, so the windows of one pass
// are disjoint and the two ends of a periodic axis never run together: each
// pass adds its window sums to Qb with plain read-modify-writes, and only
// ghost cells (whose source may be another window's cell) use atomics, after
// a barrier.
inline int tile_ncolors(int n) { return n == 1 ? 1 : ((n & 1) ? 3 : 2); }
// tiles of colour c along an axis of n tiles, and the b-th of them
inline int tile_count(int c, int n) {
  if (n == 1) return 1;
  if (c == 2) return 1;
  return (n & 1) ? (n - 1) / 2 : n / 2;
}
__device__ __forceinline__ int tile_of(int b, int c, int n) {
  return c < 0 ? b : (c == 2 ? n - 1 : 2 * b + c);  // c < 0: every tile, atomics only
}

constexpr int AF_W = 64;  // faces (x) or lanes (y, z) per tile row
constexpr int AF_R = 4;   // rows (x: y rows; y, z: face rows) per tile

template <typename real, int DIR>
struct FaceTile {
  // window extents: DIR x: AF_R rows of AF_W+4 cells; y, z: AF_R+4 rows of AF_W
  static constexpr int NWR = DIR == FV3D_IX ? AF_R : AF_R + 4;
  static constexpr int NWC = DIR == FV3D_IX ? AF_W + 4 : AF_W;
  static constexpr int NWIN = NWR * NWC;
};

template <typename real, int DIR, int RECON, int RS>
__global__ void __launch_bounds__(256)
k_adj_faces_tiled(const real* __restrict__ U, const real* __restrict__ v, real sigma, real* Qb,
                  double dt, DevParams P, int color, int ntiles) {
  using T = FaceTile<real, DIR>;
  __shared__ real sq[5][T::NWIN];
  // adjoint accumulator in FP64 for both precisions: FP32 LDS atomic adds
  // (ds_add_f32) measured 4 to 9 times slower than ds_add_f64 in these kernels
  __shared__ double sa[5][T::NWIN];
  const int tid = threadIdx.x;
  const int Nd = DIR == FV3D_IX ? P.Nx : (DIR == FV3D_IY ? P.Ny : P.Nz);
  const int bc = DIR == FV3D_IX ? P.bcx : (DIR == FV3D_IY ? P.bcy : P.bcz);
  const long long stride = DIR == FV3D_IX ? 1 : (DIR == FV3D_IY ? P.sy : P.sz);
  // tile origin: x0 (lanes or faces), d0 (first face along DIR), fixed coords
  // the grid covers the tiles of one colour along DIR (all tiles across it)
  int i0, j0, k0;
  if (DIR == FV3D_IX) { i0 = tile_of(blockIdx.x, color, ntiles) * AF_W; j0 = blockIdx.y * AF_R; k0 = blockIdx.z; }
  else if (DIR == FV3D_IY) { i0 = blockIdx.x * AF_W; j0 = tile_of(blockIdx.y, color, ntiles) * AF_R; k0 = blockIdx.z; }
  else { i0 = blockIdx.x * AF_W; j0 = blockIdx.y; k0 = tile_of(blockIdx.z, color, ntiles) * AF_R; }
  const int d0 = DIR == FV3D_IX ? i0 : (DIR == FV3D_IY ? j0 : k0);
  // window cell w -> (coordinate along DIR, base offset of the other two)
  auto win = [&](int w, int& g, bool& inside, long long& base) {
    const int r = w / T::NWC, c = w % T::NWC;
    int i, j, k;
    if (DIR == FV3D_IX) { i = i0 - 2 + c; j = j0 + r; k = k0; g = i; inside = j < P.Ny; }
    else if (DIR == FV3D_IY) { i = i0 + c; j = j0 - 2 + r; k = k0; g = j; inside = i < P.Nx; }
    else { i = i0 + c; j = j0; k = k0 - 2 + r; g = k; inside = i < P.Nx; }
    base = (DIR == FV3D_IX ? 0 : i) + (DIR == FV3D_IY ? 0 : (long long)j * P.sy) +
           (DIR == FV3D_IZ ? 0 : (long long)k * P.sz);
  };
  const real gm1 = (real)P.gm1;
  for (int w = tid; w < T::NWIN; w += blockDim.x) {
    int g; bool inside; long long base;
    win(w, g, inside, base);
    real q[5] = {0, 0, 0, 0, 0};
    if (inside) {
      bool fl;
      const int s = map_coord(g, Nd, bc, fl);
      load_prim_arr(U, base + (long long)s * stride, P.sf, gm1, q);
      if (fl) q[1 + DIR] = -q[1 + DIR];
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) { sq[n][w] = q[n]; sa[n][w] = 0; }
  }
  __syncthreads();

  const double d = DIR == FV3D_IX ? P.dx : (DIR == FV3D_IY ? P.dy : P.dz);
  const real coef = (real)((double)sigma * dt / d);
  // face fm (relative to d0, -1..AF_*-1) in tile row/lane t: stencil window
  // cells along DIR fm+1 .. fm+4
  // computes face fm of row/lane t; fills its contributions qb to the
  // window cells wc (the stencil a, b, c, d); false if there is no such face
  auto face = [&](int t, int fm, real (&qb)[4][5], int (&wc)[4]) -> bool {
    const int m = d0 + fm;
    if (m < -1 || m > Nd - 1) return false;
#pragma unroll
    for (int o = 0; o < 4; ++o)
      wc[o] = DIR == FV3D_IX ? t * T::NWC + (fm + 1 + o) : (fm + 1 + o) * T::NWC + t;
    // transverse coordinate inside the domain?
    {
      int g; bool inside; long long base;
      win(wc[1], g, inside, base);
      if (!inside) return false;
      real q[4][5];
#pragma unroll
      for (int o = 0; o < 4; ++o)
#pragma unroll
        for (int n = 0; n < 5; ++n) q[o][n] = sq[n][wc[o]];
      real Fb[5];
      bool fl;
      const long long cb = base + (long long)map_coord(m, Nd, bc, fl) * stride;
      const long long cc = base + (long long)map_coord(m + 1, Nd, bc, fl) * stride;
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        const real vc = (m + 1 < Nd) ? v[cc + n * P.sf] : (real)0;
        const real vb = (m >= 0) ? v[cb + n * P.sf] : (real)0;
        Fb[n] = coef * (vc - vb);
      }
      real qL[5], qR[5], sb[5], sc[5];
      int selb[5] = {0, 0, 0, 0, 0}, selc[5] = {0, 0, 0, 0, 0};
      if (RECON == FV3D_PLM) {
        slope_sel<real, DIR == FV3D_IZ>(q[0], q[1], q[2], sb, selb);
        slope_sel<real, DIR == FV3D_IZ>(q[1], q[2], q[3], sc, selc);
#pragma unroll
        for (int n = 0; n < 5; ++n) {
          qL[n] = q[1][n] + sb[n] * (real)0.5;
          qR[n] = q[2][n] - sc[n] * (real)0.5;
        }
      } else {
#pragma unroll
        for (int n = 0; n < 5; ++n) { qL[n] = q[1][n]; qR[n] = q[2][n]; }
      }
      swap_dir<DIR>(qL);
      swap_dir<DIR>(qR);
      swap_dir<DIR>(Fb);
      real Lb[5] = {0, 0, 0, 0, 0}, Rb[5] = {0, 0, 0, 0, 0};
      if (RS == FV3D_HLL) hll_vjp(qL, qR, (real)P.gamma, gm1, Fb, Lb, Rb);
      else hllc_vjp(qL, qR, (real)P.gamma, (real)P.entho, Fb, Lb, Rb);
      swap_dir<DIR>(Lb);
      swap_dir<DIR>(Rb);
#pragma unroll
      for (int n = 0; n < 5; ++n) { qb[0][n] = 0; qb[1][n] = Lb[n]; qb[2][n] = Rb[n]; qb[3][n] = 0; }
      if (RECON == FV3D_PLM) {
        real ssb[5], ssc[5];
#pragma unroll
        for (int n = 0; n < 5; ++n) { ssb[n] = (real)0.5 * Lb[n]; ssc[n] = (real)-0.5 * Rb[n]; }
        slope_vjp(selb, ssb, qb[0], qb[1], qb[2]);
        slope_vjp(selc, ssc, qb[1], qb[2], qb[3]);
      }
    }
    return true;
  };
  // AF_W x AF_R faces, one per thread, summed into the window with LDS
  // atomic adds (ds_add_f64 / ds_add_f32).  The tile at the low domain edge
  // also owns face -1/2 of every row.
  {
    real qb[4][5];
    int wc[4];
    const bool ok = DIR == FV3D_IX ? face(tid / AF_W, tid % AF_W, qb, wc)
                                   : face(tid % AF_W, tid / AF_W, qb, wc);
    if (ok)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int n = 0; n < 5; ++n) atomicAdd(&sa[n][wc[r]], (double)qb[r][n]);
  }
  if (d0 == 0) {
    real qb[4][5];
    int wc[4];
    const int nx = DIR == FV3D_IX ? AF_R : AF_W;
    const bool ok = tid < nx && face(tid, -1, qb, wc);
    if (ok)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int n = 0; n < 5; ++n) atomicAdd(&sa[n][wc[r]], (double)qb[r][n]);
  }
  __syncthreads();

  // one pass of atomic flushes, or (colour passes) the plain updates of the
  // tile's own cells and then the atomic ghost updates
  for (int ghosts = 0; ghosts < (color < 0 ? 1 : 2); ++ghosts) {
    if (ghosts) __syncthreads();  // this block's own updates land first
    for (int w = tid; w < T::NWIN; w += blockDim.x) {
      int g; bool inside; long long base;
      win(w, g, inside, base);
      if (!inside || g < -2 || g > Nd + 1) continue;
      const bool ghost = g < 0 || g >= Nd;
      if (color < 0 ? ghosts == 1 : ghost != (ghosts == 1)) continue;
      bool fl;
      const long long src = base + (long long)map_coord(g, Nd, bc, fl) * stride;
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        real x = (real)sa[n][w];
        if (fl && n == 1 + DIR) x = -x;
        if (ghost || color < 0) atomic_add(&Qb[src + n * P.sf], x);
        else if (x != (real)0) Qb[src + n * P.sf] += x;
      }
    }
  }
}

// Faces along y or z, marched by one thread per (x, transverse) line and
// chunk of faces.  The four-cell stencil window, the slope of its second
// cell and the adjoints of the four cells slide along the line in
// registers: a face costs one new cell (its conversion and its loads of U
// and v) and one minmod selection, and a cell leaves the registers once per
// chunk, with one global atomic add per value.  Lanes run along x, so every
// load and atomic is coalesced.  Faces m+1/2 with m in [m0, m1), m0 =
// -1 + chunk*F; they reach cells m0-1 .. m1+1.
constexpr int AL_T = 4;  // lines (transverse coordinates) per 256-thread workgroup

// The last step of an adjoint stage for one cell: out = sigma*v + J_c2p^T qb
// + alpha*extra, with qb the cell's primitive-space adjoint (rho, u, v, w,
// p, T) and q its primitives.
template <typename real>
__device__ __forceinline__ void finalize_cell(const real q[5], real qb[5], real Tb, const real vc[5],
                                              real sigma, const real* __restrict__ extra, real alpha,
                                              real* out, long long t, long long sf, real gm1) {
  // T = p / rho
  qb[4] += Tb / q[0];
  qb[0] -= Tb * q[4] / (q[0] * q[0]);
  // consToPrim (States.h:34-45) reversed
  const real r = q[0], u = q[1], vv = q[2], w = q[3];
  const real Ub1 = qb[1] / r - qb[4] * gm1 * u;
  const real Ub2 = qb[2] / r - qb[4] * gm1 * vv;
  const real Ub3 = qb[3] / r - qb[4] * gm1 * w;
  const real Ub4 = qb[4] * gm1;
  const real Ub0 = qb[0] - (qb[1] * u + qb[2] * vv + qb[3] * w) / r +
                   qb[4] * gm1 * (real)0.5 * (u * u + vv * vv + w * w);
  const real Ub[5] = {Ub0, Ub1, Ub2, Ub3, Ub4};
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    real o = sigma * vc[n] + Ub[n];
    if (extra) o += alpha * extra[t + n * sf];
    out[t + n * sf] = o;
  }
}

// z-line chunks of F faces: cell k is reached by one chunk only, and by no
// ghost image, when it lies 2 or more cells inside its chunk's face range
// and 2 or more cells inside the domain.  The z-face kernel finalizes those
// cells itself; k_adj_finalize does the rest.
__device__ __forceinline__ bool z_exclusive(int k, int Nz, int F) {
  if (k < 2 || k > Nz - 3) return false;
  const int m0 = -1 + ((k + 1) / F) * F;
  const int m1 = min(m0 + F, Nz);
  return k >= m0 + 2 && k <= m1 - 2;
}

template <typename real, int DIR, int RECON, int RS, bool FIN = false>
__global__ void __launch_bounds__(64 * AL_T)
k_adj_faces_line(const real* __restrict__ U, const real* __restrict__ v, real sigma, real* Qb,
                 double dt, DevParams P, int F, const real* __restrict__ extra = nullptr,
                 real alpha = 0, real* out = nullptr) {
  static_assert(!FIN || DIR == FV3D_IZ, "only the z faces, launched last, finalize");
  static_assert(DIR == FV3D_IY || DIR == FV3D_IZ, "x faces use the tiled kernel");
  constexpr bool ZORD = DIR == FV3D_IZ;
  const int Nd = DIR == FV3D_IY ? P.Ny : P.Nz;
  const int Nt = DIR == FV3D_IY ? P.Nz : P.Ny;
  const int bc = DIR == FV3D_IY ? P.bcy : P.bcz;
  const long long stride = DIR == FV3D_IY ? P.sy : P.sz;
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int t = blockIdx.y * AL_T + (threadIdx.x >> 6);
  const int m0 = -1 + (int)blockIdx.z * F;
  const int m1 = min(m0 + F, Nd);
  if (i >= P.Nx || t >= Nt || m0 >= m1) return;
  const long long base = i + (long long)t * (DIR == FV3D_IY ? P.sz : P.sy);
  const real gm1 = (real)P.gm1;
  const real coef = (real)((double)sigma * dt / (DIR == FV3D_IY ? P.dy : P.dz));

  real q[4][5];  // primitives of cells m-1 .. m+2 (ghosts mirrored)
  long long src[4];
  bool flip[4];
  auto load = [&](int o, int cm) {
    bool f;
    const int sidx = map_coord(cm, Nd, bc, f);
    src[o] = base + (long long)sidx * stride;
    flip[o] = f;
    load_prim_arr(U, src[o], P.sf, gm1, q[o]);
    if (f) q[o][1 + DIR] = -q[o][1 + DIR];
  };
  real acc[4][5];
  auto flush = [&](int o) {
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      const real x = (flip[o] && n == 1 + DIR) ? -acc[o][n] : acc[o][n];
      atomic_add(&Qb[src[o] + n * P.sf], x);
    }
  };
  real va[5] = {0, 0, 0, 0, 0};  // v of cell m-1 (FIN)
  load(0, m0 - 1);
  load(1, m0);
  load(2, m0 + 1);
#pragma unroll
  for (int o = 0; o < 4; ++o)
#pragma unroll
    for (int n = 0; n < 5; ++n) acc[o][n] = (real)0;
  real vb[5];  // v of cell m (zero outside the domain)
#pragma unroll
  for (int n = 0; n < 5; ++n) vb[n] = m0 >= 0 ? v[src[1] + n * P.sf] : (real)0;
  real sb[5] = {0, 0, 0, 0, 0};  // slope of cell m, carried from the previous face
  int selb[5] = {0, 0, 0, 0, 0};
  if (RECON == FV3D_PLM) slope_sel<real, ZORD>(q[0], q[1], q[2], sb, selb);

  for (int m = m0; m < m1; ++m) {
    load(3, m + 2);
    real Fb[5], vc[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      vc[n] = m + 1 < Nd ? v[src[2] + n * P.sf] : (real)0;
      Fb[n] = coef * (vc[n] - vb[n]);
    }
    real qL[5], qR[5], sc[5] = {0, 0, 0, 0, 0};
    int selc[5] = {0, 0, 0, 0, 0};
    if (RECON == FV3D_PLM) {
      slope_sel<real, ZORD>(q[1], q[2], q[3], sc, selc);
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        qL[n] = q[1][n] + sb[n] * (real)0.5;
        qR[n] = q[2][n] - sc[n] * (real)0.5;
      }
    } else {
#pragma unroll
      for (int n = 0; n < 5; ++n) { qL[n] = q[1][n]; qR[n] = q[2][n]; }
    }
    swap_dir<DIR>(qL);
    swap_dir<DIR>(qR);
    swap_dir<DIR>(Fb);
    real Lb[5] = {0, 0, 0, 0, 0}, Rb[5] = {0, 0, 0, 0, 0};
    if (RS == FV3D_HLL) hll_vjp(qL, qR, (real)P.gamma, gm1, Fb, Lb, Rb);
    else hllc_vjp(qL, qR, (real)P.gamma, (real)P.entho, Fb, Lb, Rb);
    swap_dir<DIR>(Lb);
    swap_dir<DIR>(Rb);
#pragma unroll
    for (int n = 0; n < 5; ++n) { acc[1][n] += Lb[n]; acc[2][n] += Rb[n]; }
    if (RECON == FV3D_PLM) {
      real ssb[5], ssc[5];
#pragma unroll
      for (int n = 0; n < 5; ++n) { ssb[n] = (real)0.5 * Lb[n]; ssc[n] = (real)-0.5 * Rb[n]; }
      slope_vjp(selb, ssb, acc[0], acc[1], acc[2]);
      slope_vjp(selc, ssc, acc[1], acc[2], acc[3]);
    }
    // cell m-1: no later face of the chunk reaches it
    if (FIN && z_exclusive(m - 1, Nd, F)) {
      // every other term of its adjoint is in Qb already, and no other
      // thread adds to it: finalize it here (same sums as a flush followed
      // by k_adj_finalize) and clear its Qb
      const long long c0 = src[0];
      real qb[5];
#pragma unroll
      for (int n = 0; n < 5; ++n) qb[n] = Qb[c0 + n * P.sf] + acc[0][n];
      const real Tb = Qb[c0 + 5 * P.sf];
      finalize_cell(q[0], qb, Tb, va, sigma, extra, alpha, out, c0, P.sf, gm1);
#pragma unroll
      for (int n = 0; n < 6; ++n) Qb[c0 + n * P.sf] = (real)0;
    } else {
      flush(0);
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) va[n] = vb[n];
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      src[o] = src[o + 1];
      flip[o] = flip[o + 1];
#pragma unroll
      for (int n = 0; n < 5; ++n) { q[o][n] = q[o + 1][n]; acc[o][n] = acc[o + 1][n]; }
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      acc[3][n] = (real)0;
      vb[n] = vc[n];
      sb[n] = sc[n];
      selb[n] = selc[n];
    }
  }
  flush(0);
  flush(1);
  flush(2);
}

// The values of v that the diffusion adjoint of a cell reads (fields 1-4,
// momenta and energy): at the cell (c) and, with viscosity, at its +x (x)
// and +y (y) neighbours inside the domain, fetched ahead of the arithmetic
// (fetch_diff_v) so that a caller can issue the loads early.
template <typename real>
struct DiffVPre {
  real c[4], x[4], y[4];
};
template <typename real, bool VISC>
__device__ __forceinline__ DiffVPre<real> fetch_diff_v(const real* __restrict__ v, const DevParams& P,
                                                       int i, int j, long long cell) {
  DiffVPre<real> V;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    V.c[c] = v[cell + (c + 1) * P.sf];
    V.x[c] = (VISC && i + 1 < P.Nx) ? v[cell + 1 + (c + 1) * P.sf] : (real)0;
    V.y[c] = (VISC && j + 1 < P.Ny) ? v[cell + P.sy + (c + 1) * P.sf] : (real)0;
  }
  return V;
}

// The diffusion adjoint of one cell (i, j, k): the transposes of its
// conduction update (ThermalConduction.h:47-95) and of its viscous sides
// (Viscosity.h:28-153; x and y faces once, as side 2 of the lower cell with
// both cells' adjoints, side 1 only at the low domain edge; z sides per
// cell), and the cell's dS/dkappa, dS/dmu dotted with sigma*v.  PRE: the
// values of v come in pre (fetch_diff_v), else they are read from v where
// the arithmetic uses them.  S(dk, dj, di, c) reads the window value c (u,
// v, w, T) at offset (di-1, dj-1, dk-1), mirrored for ghosts; A(dk, dj, di,
// c, x) adds x to its adjoint.
template <typename real, bool TC, bool VISC, bool PRE = false, typename GET, typename ADD>
__device__ __forceinline__ void diffusion_cell_adjoint(const real* __restrict__ v,
                                                       const DiffVPre<real>& pre, real sigma,
                                                       double dt, const DevParams& P, int i,
                                                       int j, int k, long long cell, GET&& S,
                                                       ADD&& A, double& gk, double& gm) {
  const real rdt = (real)dt;
  // The conduction terms and those of the four viscous sides every cell has
  // are summed per (offset, component) in registers first: their 119 terms
  // land on 51 distinct values, each added to the window once.  Entries no
  // term reaches fold to the constant 0 and are skipped.
  real acc[3][3][3][4];
#pragma unroll
  for (int e = 0; e < 108; ++e) (&acc[0][0][0][0])[e] = (real)0;
  auto flush_acc = [&](auto& a, int nc) {
#pragma unroll
    for (int dk = 0; dk < 3; ++dk)
#pragma unroll
      for (int dj = 0; dj < 3; ++dj)
#pragma unroll
        for (int di = 0; di < 3; ++di)
#pragma unroll
          for (int c = 0; c < 4; ++c)
            if (c < nc && !(__builtin_constant_p(a[dk][dj][di][c]) && a[dk][dj][di][c] == (real)0))
              A(dk, dj, di, c, a[dk][dj][di][c]);
  };
  if (TC) {
    const real Eb = sigma * (PRE ? pre.c[3] : v[cell + 4 * P.sf]);
    // stencil offsets (dk, dj, di): centre, x-, x+, y-, y+, z-, z+
    constexpr int ok[7] = {1, 1, 1, 1, 1, 0, 2}, oj[7] = {1, 1, 1, 0, 2, 1, 1},
                  oi[7] = {1, 0, 2, 1, 1, 1, 1};
    const real ax = rdt / (real)(P.dx * P.dx), ay = rdt / (real)(P.dy * P.dy),
               az = rdt / (real)(P.dz * P.dz);
    real c[7] = {-(real)2 * (ax + ay + az), ax, ax, ay, ay, az, az};
    real fixedE = 0;
    if (k == 0 && P.bctc_zmin != FV3D_BCTC_NONE) {
      c[0] += az; c[5] -= az;
      if (P.bctc_zmin == FV3D_BCTC_FIXED_TEMPERATURE) {
        c[0] -= (real)2 * az;
        fixedE += (real)2 * az * (real)P.bctc_zmin_value;
      } else if (P.bctc_zmin == FV3D_BCTC_FIXED_GRADIENT) {
        fixedE -= rdt / (real)P.dz * (real)P.bctc_zmin_value;
      }
    }
    if (k == P.Nz - 1 && P.bctc_zmax != FV3D_BCTC_NONE) {
      c[0] += az; c[6] -= az;
      if (P.bctc_zmax == FV3D_BCTC_FIXED_TEMPERATURE) {
        c[0] -= (real)2 * az;
        fixedE += (real)2 * az * (real)P.bctc_zmax_value;
      } else if (P.bctc_zmax == FV3D_BCTC_FIXED_GRADIENT) {
        fixedE += rdt / (real)P.dz * (real)P.bctc_zmax_value;
      }
    }
    real dE1 = fixedE;
#pragma unroll
    for (int o = 0; o < 7; ++o) dE1 += c[o] * S(ok[o], oj[o], oi[o], 3);
    gk = (double)Eb * (double)dE1;
    const real kf = (real)P.kface;
#pragma unroll
    for (int o = 0; o < 7; ++o) acc[ok[o]][oj[o]][oi[o]][3] += Eb * kf * c[o];
  }
  if (VISC) {
    // per side: x and y faces once (side 2 of the lower cell, carrying
    // both cells' adjoints; side 1 only at the low domain edge), z sides
    // per cell
    const real oodx = (real)1 / (real)P.dx, oody = (real)1 / (real)P.dy, oodz = (real)1 / (real)P.dz;
    const real sdt = sigma * rdt, mu = (real)P.mu;
    real gi[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) gi[c] = sdt * (PRE ? pre.c[c] : v[cell + (c + 1) * P.sf]);
    auto Sget = [&](int dk, int dj, int di, int c) -> real { return S(dk, dj, di, c); };
    double gmu = 0.0;
    // one pass gives the side's flux (for dS/dmu) and its transpose, summed into a
    auto side = [&](auto blk, int sd, const real gw[4], auto& a) {
      constexpr int B = decltype(blk)::value;
      real out[4] = {0, 0, 0, 0}, gm4[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) gm4[c] = gw[c] * mu;
      auto Sloc = [&](int dk, int dj, int di, int c, real x) { a[dk][dj][di][c] += x; };
      visc_side<real, true, B>(Sget, sd, oodx, oody, oodz, gm4, out, Sloc);
#pragma unroll
      for (int c = 0; c < 4; ++c) gmu += (double)gw[c] * (double)out[c];
    };
    // the low-edge sides are summed and flushed on their own, so that the
    // cell-wide sum keeps its compile-time zeros
    auto edge_side = [&](auto blk, const real gw[4]) {
      real e[3][3][3][4];
#pragma unroll
      for (int n = 0; n < 108; ++n) (&e[0][0][0][0])[n] = (real)0;
      side(blk, 1, gw, e);
      flush_acc(e, 3);
    };
    using BX = std::integral_constant<int, 0>;
    using BY = std::integral_constant<int, 1>;
    using BZ = std::integral_constant<int, 2>;
    real gf[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) gf[c] = gi[c] - (i + 1 < P.Nx ? sdt * (PRE ? pre.x[c] : v[cell + 1 + (c + 1) * P.sf]) : (real)0);
    side(BX{}, 2, gf, acc);
    if (i == 0) edge_side(BX{}, gi);
#pragma unroll
    for (int c = 0; c < 4; ++c) gf[c] = gi[c] - (j + 1 < P.Ny ? sdt * (PRE ? pre.y[c] : v[cell + P.sy + (c + 1) * P.sf]) : (real)0);
    side(BY{}, 2, gf, acc);
    if (j == 0) edge_side(BY{}, gi);
    side(BZ{}, 1, gi, acc);
    side(BZ{}, 2, gi, acc);
    gm = gmu;
  }
  flush_acc(acc, 4);
}

constexpr int AD_X = 32, AD_Y = 4, AD_Z = 2;  // diffusion tile (256 threads: 3 workgroups per CU)

template <typename real, bool TC, bool VISC>
__global__ void __launch_bounds__(AD_X * AD_Y * AD_Z)
k_adj_diffusion_tiled(const real* __restrict__ U, const real* __restrict__ v, real sigma,
                      real* Qb, double dt, DevParams P, double* grad, int3 color, int3 ntiles) {
  constexpr int WX = AD_X + 2, WY = AD_Y + 2, WZ = AD_Z + 2, NW = WX * WY * WZ;
  __shared__ real sq[4][NW];  // u, v, w, T of the window (ghost-mapped)
  __shared__ double sa[4][NW];  // their adjoints (FP64 in both precisions)
  const int tid = threadIdx.x;
  // the grid covers the tiles of one colour
  const int i0 = tile_of(blockIdx.x, color.x, ntiles.x) * AD_X;
  const int j0 = tile_of(blockIdx.y, color.y, ntiles.y) * AD_Y;
  const int k0 = tile_of(blockIdx.z, color.z, ntiles.z) * AD_Z;
  const real gm1 = (real)P.gm1;
  auto wcoord = [&](int w, int& i, int& j, int& k) {
    i = i0 - 1 + w % WX;
    j = j0 - 1 + (w / WX) % WY;
    k = k0 - 1 + w / (WX * WY);
  };
  for (int w = tid; w < NW; w += blockDim.x) {
    int i, j, k;
    wcoord(w, i, j, k);
    bool fx, fy, fz;
    const int si = map_coord(i, P.Nx, P.bcx, fx), sj = map_coord(j, P.Ny, P.bcy, fy),
              sk = map_coord(k, P.Nz, P.bcz, fz);
    real q[5];
    load_prim_arr(U, si + (long long)sj * P.sy + (long long)sk * P.sz, P.sf, gm1, q);
    sq[0][w] = fx ? -q[1] : q[1];
    sq[1][w] = fy ? -q[2] : q[2];
    sq[2][w] = fz ? -q[3] : q[3];
    sq[3][w] = q[4] / q[0];
#pragma unroll
    for (int c = 0; c < 4; ++c) sa[c][w] = 0;
  }
  __syncthreads();

  const int ti = tid % AD_X, tj = (tid / AD_X) % AD_Y, tk = tid / (AD_X * AD_Y);
  const int i = i0 + ti, j = j0 + tj, k = k0 + tk;
  double gk = 0.0, gm = 0.0;
  // this cell's contributions to its 3x3x3 neighbourhood go straight into
  // the window accumulator with LDS atomic adds
  if (i < P.Nx && j < P.Ny && k < P.Nz) {
    const long long cell = i + (long long)j * P.sy + (long long)k * P.sz;
    auto W = [&](int dk, int dj, int di) {  // window index of offset (di-1, dj-1, dk-1)
      return (ti + di) + WX * ((tj + dj) + WY * (tk + dk));
    };
    diffusion_cell_adjoint<real, TC, VISC>(
        v, DiffVPre<real>{}, sigma, dt, P, i, j, k, cell,
        [&](int dk, int dj, int di, int c) -> real { return sq[c][W(dk, dj, di)]; },
        [&](int dk, int dj, int di, int c, real x) { atomicAdd(&sa[c][W(dk, dj, di)], (double)x); }, gk, gm);
  }
  __syncthreads();
  for (int ghosts = 0; ghosts < (color.x < 0 ? 1 : 2); ++ghosts) {
    if (ghosts) __syncthreads();  // this block's own updates land first
    for (int w = tid; w < NW; w += blockDim.x) {
      int wi, wj, wk;
      wcoord(w, wi, wj, wk);
      if (wi < -1 || wi > P.Nx || wj < -1 || wj > P.Ny || wk < -1 || wk > P.Nz) continue;
      const bool ghost = wi < 0 || wi >= P.Nx || wj < 0 || wj >= P.Ny || wk < 0 || wk >= P.Nz;
      if (ghost != (ghosts == 1) && color.x >= 0) continue;
      if (color.x < 0 && ghosts == 1) continue;
      bool fx, fy, fz;
      const int si = map_coord(wi, P.Nx, P.bcx, fx), sj = map_coord(wj, P.Ny, P.bcy, fy),
                sk = map_coord(wk, P.Nz, P.bcz, fz);
      const long long src = si + (long long)sj * P.sy + (long long)sk * P.sz;
      auto put = [&](long long a, real x) {
        if (ghost || color.x < 0) atomic_add(&Qb[a], x);
        else if (x != (real)0) Qb[a] += x;
      };
      if (VISC) {
        put(src + 1 * P.sf, (real)(fx ? -sa[0][w] : sa[0][w]));
        put(src + 2 * P.sf, (real)(fy ? -sa[1][w] : sa[1][w]));
        put(src + 3 * P.sf, (real)(fz ? -sa[2][w] : sa[2][w]));
      }
      if (TC) put(src + 5 * P.sf, (real)sa[3][w]);
    }
  }
  // parameter gradients: one partial per tile (summed by k_adj_reduce; a
  // global atomic per wave on one address would serialise every wave)
  __shared__ double sred[2][AD_X * AD_Y * AD_Z / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    gk += __shfl_xor(gk, off);
    gm += __shfl_xor(gm, off);
  }
  if ((tid & 63) == 0) { sred[0][tid >> 6] = gk; sred[1][tid >> 6] = gm; }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < AD_X * AD_Y * AD_Z / 64; ++w) { a += sred[0][w]; b += sred[1][w]; }
    const long long t = (i0 / AD_X) + (long long)ntiles.x * ((j0 / AD_Y) + (long long)ntiles.y * (k0 / AD_Z));
    grad[2 + 2 * t] = a;
    grad[3 + 2 * t] = b;
  }
}

// The diffusion adjoint marched along z: a workgroup owns a 64 x 4 column
// of a z chunk and keeps three planes of its (66 x 6) window, primitives
// and adjoints, in LDS rings.  Each window cell is converted once per chunk
// (plus two halo planes), and a plane's adjoints leave LDS once, with one
// global atomic add per value, when the march has passed it.  (The FP64
// kernel; FP32 runs k_adj_diffusion_pipe.)
constexpr int AZ_X = 64, AZ_Y = 4;

template <typename real, bool TC, bool VISC>
__global__ void __launch_bounds__(AZ_X * AZ_Y)
k_adj_diffusion_march(const real* __restrict__ U, const real* __restrict__ v, real sigma,
                      real* Qb, double dt, DevParams P, double* grad, int nzc) {
  constexpr int WX = AZ_X + 2, WY = AZ_Y + 2, NWP = WX * WY;
  __shared__ real sq[3][4][NWP];  // u, v, w, T of three window planes (ghost-mapped)
  __shared__ double sa[3][4][NWP];  // their adjoints (FP64 in both precisions: ds_add_f64)
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * AZ_X, j0 = blockIdx.y * AZ_Y, k0 = blockIdx.z * nzc;
  const int k1 = min(k0 + nzc, P.Nz);
  const real gm1 = (real)P.gm1;
  auto slot = [](int kk) { return (kk + 3) % 3; };  // kk >= -1
  auto load_plane = [&](int kk) {
    const int sl = slot(kk);
    bool fz;
    const int sk = map_coord(kk, P.Nz, P.bcz, fz);
    for (int w = tid; w < NWP; w += AZ_X * AZ_Y) {
      bool fx, fy;
      const int si = map_coord(i0 - 1 + w % WX, P.Nx, P.bcx, fx);
      const int sj = map_coord(j0 - 1 + w / WX, P.Ny, P.bcy, fy);
      real q[5];
      load_prim_arr(U, si + (long long)sj * P.sy + (long long)sk * P.sz, P.sf, gm1, q);
      sq[sl][0][w] = fx ? -q[1] : q[1];
      sq[sl][1][w] = fy ? -q[2] : q[2];
      sq[sl][2][w] = fz ? -q[3] : q[3];
      sq[sl][3][w] = q[4] / q[0];
#pragma unroll
      for (int c = 0; c < 4; ++c) sa[sl][c][w] = 0;
    }
  };
  auto flush_plane = [&](int kk) {
    if (kk < -1 || kk > P.Nz) return;
    const int sl = slot(kk);
    bool fz;
    const int sk = map_coord(kk, P.Nz, P.bcz, fz);
    for (int w = tid; w < NWP; w += AZ_X * AZ_Y) {
      const int wi = i0 - 1 + w % WX, wj = j0 - 1 + w / WX;
      if (wi > P.Nx || wj > P.Ny) continue;
      bool fx, fy;
      const int si = map_coord(wi, P.Nx, P.bcx, fx), sj = map_coord(wj, P.Ny, P.bcy, fy);
      const long long src = si + (long long)sj * P.sy + (long long)sk * P.sz;
      if (VISC) {
        atomic_add(&Qb[src + 1 * P.sf], (real)(fx ? -sa[sl][0][w] : sa[sl][0][w]));
        atomic_add(&Qb[src + 2 * P.sf], (real)(fy ? -sa[sl][1][w] : sa[sl][1][w]));
        atomic_add(&Qb[src + 3 * P.sf], (real)(fz ? -sa[sl][2][w] : sa[sl][2][w]));
      }
      if (TC) atomic_add(&Qb[src + 5 * P.sf], (real)sa[sl][3][w]);
    }
  };

  const int ti = tid % AZ_X, tj = tid / AZ_X;
  const int i = i0 + ti, j = j0 + tj;
  double gk = 0.0, gm = 0.0;
  load_plane(k0 - 1);
  load_plane(k0);
  for (int k = k0; k < k1; ++k) {
    load_plane(k + 1);
    __syncthreads();
    if (i < P.Nx && j < P.Ny) {
      const long long cell = i + (long long)j * P.sy + (long long)k * P.sz;
      const int sl[3] = {slot(k - 1), slot(k), slot(k + 1)};
      auto W = [&](int dj, int di) { return (ti + di) + WX * (tj + dj); };
      double gkc = 0.0, gmc = 0.0;
      diffusion_cell_adjoint<real, TC, VISC>(
          v, DiffVPre<real>{}, sigma, dt, P, i, j, k, cell,
          [&](int dk, int dj, int di, int c) -> real { return sq[sl[dk]][c][W(dj, di)]; },
          [&](int dk, int dj, int di, int c, real x) { atomicAdd(&sa[sl[dk]][c][W(dj, di)], (double)x); },
          gkc, gmc);
      gk += gkc;
      gm += gmc;
    }
    __syncthreads();
    flush_plane(k - 1);  // no later cell of the chunk reaches plane k-1
    __syncthreads();     // before the next plane overwrites its slot
  }
  flush_plane(k1 - 1);
  flush_plane(k1);
  // parameter gradients: one partial per workgroup
  __shared__ double sred[2][AZ_X * AZ_Y / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    gk += __shfl_xor(gk, off);
    gm += __shfl_xor(gm, off);
  }
  if ((tid & 63) == 0) { sred[0][tid >> 6] = gk; sred[1][tid >> 6] = gm; }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < AZ_X * AZ_Y / 64; ++w) { a += sred[0][w]; b += sred[1][w]; }
    const long long t = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    grad[2 + 2 * t] = a;
    grad[3 + 2 * t] = b;
  }
}

// k_adj_diffusion_march in a pipelined order, for FP32: each step first
// issues its global loads (the next window plane, the cell's values of v)
// and flushes the plane completed one step earlier while they are in
// flight (measured 19% faster on the 128^3 box).  In FP64 the same order,
// or the same loads, lift the kernel to 255 VGPRs and measured 23-65%
// slower, so FP64 keeps the order above.
template <typename real, bool TC, bool VISC>
__global__ void __launch_bounds__(AZ_X * AZ_Y)
k_adj_diffusion_pipe(const real* __restrict__ U, const real* __restrict__ v, real sigma,
                      real* Qb, double dt, DevParams P, double* grad, int nzc) {
  constexpr int WX = AZ_X + 2, WY = AZ_Y + 2, NWP = WX * WY;
  constexpr int NT = AZ_X * AZ_Y;
  constexpr int NPT = (NWP + NT - 1) / NT;  // window cells per thread
  __shared__ real sq[3][4][NWP];  // u, v, w, T of three window planes (ghost-mapped)
  __shared__ double sa[3][4][NWP];  // their adjoints (FP64 in both precisions: ds_add_f64)
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * AZ_X, j0 = blockIdx.y * AZ_Y, k0 = blockIdx.z * nzc;
  const int k1 = min(k0 + nzc, P.Nz);
  const real gm1 = (real)P.gm1;
  auto slot = [](int kk) { return (kk + 3) % 3; };  // kk >= -1
  // raw conserved values of this thread's window cells (w = tid + n * NT)
  // of one plane, and their x/y reflection flags
  real raw[NPT][5];
  bool rfx[NPT], rfy[NPT];
  auto fetch_plane = [&](int kk) {
    bool fz;
    const int sk = map_coord(kk, P.Nz, P.bcz, fz);
#pragma unroll
    for (int n = 0; n < NPT; ++n) {
      const int w = tid + n * NT;
      if (w < NWP) {
        const int si = map_coord(i0 - 1 + w % WX, P.Nx, P.bcx, rfx[n]);
        const int sj = map_coord(j0 - 1 + w / WX, P.Ny, P.bcy, rfy[n]);
        const long long base = si + (long long)sj * P.sy + (long long)sk * P.sz;
#pragma unroll
        for (int f = 0; f < 5; ++f) raw[n][f] = U[base + f * P.sf];
      }
    }
  };
  auto store_plane = [&](int kk) {
    const int sl = slot(kk);
    bool fz;
    map_coord(kk, P.Nz, P.bcz, fz);
#pragma unroll
    for (int n = 0; n < NPT; ++n) {
      const int w = tid + n * NT;
      if (w < NWP) {
        const Prim<real> q = cons_to_prim(raw[n][0], raw[n][1], raw[n][2], raw[n][3], raw[n][4], gm1);
        sq[sl][0][w] = rfx[n] ? -q.u : q.u;
        sq[sl][1][w] = rfy[n] ? -q.v : q.v;
        sq[sl][2][w] = fz ? -q.w : q.w;
        sq[sl][3][w] = q.p / q.r;
#pragma unroll
        for (int c = 0; c < 4; ++c) sa[sl][c][w] = 0;
      }
    }
  };
  auto flush_plane = [&](int kk) {
    if (kk < -1 || kk > P.Nz) return;
    const int sl = slot(kk);
    bool fz;
    const int sk = map_coord(kk, P.Nz, P.bcz, fz);
    for (int w = tid; w < NWP; w += NT) {
      const int wi = i0 - 1 + w % WX, wj = j0 - 1 + w / WX;
      if (wi > P.Nx || wj > P.Ny) continue;
      bool fx, fy;
      const int si = map_coord(wi, P.Nx, P.bcx, fx), sj = map_coord(wj, P.Ny, P.bcy, fy);
      const long long src = si + (long long)sj * P.sy + (long long)sk * P.sz;
      if (VISC) {
        atomic_add(&Qb[src + 1 * P.sf], (real)(fx ? -sa[sl][0][w] : sa[sl][0][w]));
        atomic_add(&Qb[src + 2 * P.sf], (real)(fy ? -sa[sl][1][w] : sa[sl][1][w]));
        atomic_add(&Qb[src + 3 * P.sf], (real)(fz ? -sa[sl][2][w] : sa[sl][2][w]));
      }
      if (TC) atomic_add(&Qb[src + 5 * P.sf], (real)sa[sl][3][w]);
    }
  };

  const int ti = tid % AZ_X, tj = tid / AZ_X;
  const int i = i0 + ti, j = j0 + tj;
  const bool active = i < P.Nx && j < P.Ny;
  double gk = 0.0, gm = 0.0;
  fetch_plane(k0 - 1);
  store_plane(k0 - 1);
  fetch_plane(k0);
  store_plane(k0);
  for (int k = k0; k < k1; ++k) {
    const long long cell = i + (long long)j * P.sy + (long long)k * P.sz;
    fetch_plane(k + 1);
    DiffVPre<real> V{};
    if (active) V = fetch_diff_v<real, VISC>(v, P, i, j, cell);
    // plane k-2 is complete (its last adds came from the cells of plane
    // k-1), and its ring slot takes plane k+1
    if (k - 2 >= k0 - 1) flush_plane(k - 2);
    __syncthreads();
    store_plane(k + 1);
    __syncthreads();
    if (active) {
      const int sl[3] = {slot(k - 1), slot(k), slot(k + 1)};
      auto W = [&](int dj, int di) { return (ti + di) + WX * (tj + dj); };
      double gkc = 0.0, gmc = 0.0;
      auto S = [&](int dk, int dj, int di, int c) -> real { return sq[sl[dk]][c][W(dj, di)]; };
      auto A = [&](int dk, int dj, int di, int c, real x) {
        atomicAdd(&sa[sl[dk]][c][W(dj, di)], (double)x);
      };
      diffusion_cell_adjoint<real, TC, VISC, true>(v, V, sigma, dt, P, i, j, k, cell, S, A, gkc, gmc);
      gk += gkc;
      gm += gmc;
    }
    __syncthreads();  // this plane's adds are done before the next step flushes them
  }
  flush_plane(k1 - 2);
  flush_plane(k1 - 1);
  flush_plane(k1);
  // parameter gradients: one partial per workgroup
  __shared__ double sred[2][AZ_X * AZ_Y / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    gk += __shfl_xor(gk, off);
    gm += __shfl_xor(gm, off);
  }
  if ((tid & 63) == 0) { sred[0][tid >> 6] = gk; sred[1][tid >> 6] = gm; }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < AZ_X * AZ_Y / 64; ++w) { a += sred[0][w]; b += sred[1][w]; }
    const long long t = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    grad[2 + 2 * t] = a;
    grad[3 + 2 * t] = b;
  }
}

// grad[0..1] += the per-tile partials grad[2 + 2t], grad[3 + 2t]
template <typename real>  // one instance per precision's translation unit
__global__ void __launch_bounds__(256) k_adj_reduce(double* grad, long long ntiles) {
  __shared__ double s[2][256];
  double a = 0.0, b = 0.0;
  for (long long t = threadIdx.x; t < ntiles; t += 256) { a += grad[2 + 2 * t]; b += grad[3 + 2 * t]; }
  s[0][threadIdx.x] = a;
  s[1][threadIdx.x] = b;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s[0][threadIdx.x] += s[0][threadIdx.x + h];
      s[1][threadIdx.x] += s[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { grad[0] += s[0][0]; grad[1] += s[1][0]; }
}

// out = sigma*v + J_c2p(U)^T Qb + alpha*extra; Qb cleared.  Fz > 0: the
// z-face kernel (chunks of Fz faces) has finalized the cells z_exclusive
// picks; they are skipped.
template <typename real>
__global__ void __launch_bounds__(256)
k_adj_finalize(const real* __restrict__ U, const real* __restrict__ v, real sigma,
               const real* __restrict__ extra, real alpha, real* Qb, real* out, DevParams P,
               int Fz) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= P.sf) return;
  if (Fz > 0 && z_exclusive((int)(t / P.sz), P.Nz, Fz)) return;
  const long long sf = P.sf;
  const real gm1 = (real)P.gm1;
  real q[5];
  load_prim_arr(U, t, sf, gm1, q);
  real qb[5], vc[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    qb[n] = Qb[t + n * sf];
    vc[n] = v[t + n * sf];
  }
  const real Tb = Qb[t + 5 * sf];
  finalize_cell(q, qb, Tb, vc, sigma, extra, alpha, out, t, sf, gm1);
#pragma unroll
  for (int n = 0; n < 6; ++n) Qb[t + n * sf] = (real)0;
}

// resident workgroups of a line kernel instance on the whole current
// device, queried once per device (threads that race here compute the same
// value; the cache is atomic)
template <typename real, int DIR, int RECON, int RS, bool FIN>
static long long line_round_capacity() {
  constexpr int kMaxDev = 64;
  static std::atomic<long long> caps[kMaxDev];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  long long cap = dev >= 0 && dev < kMaxDev ? caps[dev].load(std::memory_order_relaxed) : 0;
  if (cap == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&k_adj_faces_line<real, DIR, RECON, RS, FIN>),
            64 * AL_T, 0) != hipSuccess || per_cu < 1)
      per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    cap = (long long)per_cu * cus;
    if (dev >= 0 && dev < kMaxDev) caps[dev].store(cap, std::memory_order_relaxed);
  }
  return cap;
}

// Fz (z faces only): set to the chunk length when the z-face kernel also
// finalizes the cells only it reaches (the default path; launched last)
template <typename real, int DIR, int RECON, int RS>
static hipError_t adj_faces_dir(const AdjLaunch<real>& L, hipStream_t s, int* Fz = nullptr) {
  const DevParams& P = L.P;
  if (L.simple) {
    const long long nf = (long long)(P.Nx + (DIR == FV3D_IX)) * (P.Ny + (DIR == FV3D_IY)) *
                         (P.Nz + (DIR == FV3D_IZ));
    const int bs = 256;
    hipLaunchKernelGGL((k_adj_faces<real, DIR, RECON, RS>), dim3((unsigned)((nf + bs - 1) / bs)),
                       dim3(bs), 0, s, L.U, L.v, L.sigma, L.Qb, L.dt, P);
    return hipGetLastError();
  }
  if (DIR != FV3D_IX && !L.colored) {
    const int Nd = DIR == FV3D_IY ? P.Ny : P.Nz, Nt = DIR == FV3D_IY ? P.Nz : P.Ny;
    // chunks of 15 to 32 faces, the count chosen so that the workgroups fill
    // whole rounds of the device (resident workgroups per CU x CUs): the
    // cost model is rounds x (faces per chunk + 3 for the warm-up cells).
    // At 128^3 FP64 (two workgroups per CU) that is 8 chunks of 17 faces in
    // one round instead of 9 of 15 in two (-2.5% on the backward pass);
    // shorter chunks measured slower in FP32.
    const bool fin = DIR == FV3D_IZ && Fz;
    const long long per_round = fin ? line_round_capacity<real, FV3D_IZ, RECON, RS, true>()
                                    : line_round_capacity<real, DIR == FV3D_IX ? FV3D_IY : DIR, RECON, RS, false>();
    const long long lines = (long long)((P.Nx + 63) / 64) * ((Nt + AL_T - 1) / AL_T);
    const int cmin = std::max(1, (Nd + 1) / 32), cmax = std::max(cmin, (Nd + 1 + 14) / 15);
    int nch = cmin;
    double best = 0.0;
    for (int c = cmin; c <= cmax; ++c) {
      const int f = (Nd + 1 + c - 1) / c;
      const long long rounds = (lines * c + per_round - 1) / per_round;
      const double cost = (double)rounds * (f + 3);
      if (c == cmin || cost < best) { best = cost; nch = c; }
    }
    const int F = (Nd + 1 + nch - 1) / nch;
    const dim3 g((P.Nx + 63) / 64, (Nt + AL_T - 1) / AL_T, (Nd + 1 + F - 1) / F);
    if (DIR == FV3D_IZ && Fz) {
      hipLaunchKernelGGL((k_adj_faces_line<real, FV3D_IZ, RECON, RS, true>), g, dim3(64 * AL_T), 0, s,
                         L.U, L.v, L.sigma, L.Qb, L.dt, P, F, L.extra, L.alpha, L.out);
      *Fz = F;
    } else {
      hipLaunchKernelGGL((k_adj_faces_line<real, DIR == FV3D_IX ? FV3D_IY : DIR, RECON, RS>), g,
                         dim3(64 * AL_T), 0, s, L.U, L.v, L.sigma, L.Qb, L.dt, P, F);
    }
    return hipGetLastError();
  }
  dim3 grid;
  if (DIR == FV3D_IX) grid = dim3((P.Nx + AF_W - 1) / AF_W, (P.Ny + AF_R - 1) / AF_R, P.Nz);
  else if (DIR == FV3D_IY) grid = dim3((P.Nx + AF_W - 1) / AF_W, (P.Ny + AF_R - 1) / AF_R, P.Nz);
  else grid = dim3((P.Nx + AF_W - 1) / AF_W, P.Ny, (P.Nz + AF_R - 1) / AF_R);
  const int nt = DIR == FV3D_IX ? grid.x : (DIR == FV3D_IY ? grid.y : grid.z);
  if (!L.colored) {
    hipLaunchKernelGGL((k_adj_faces_tiled<real, DIR, RECON, RS>), grid, dim3(AF_W * AF_R), 0, s, L.U,
                       L.v, L.sigma, L.Qb, L.dt, P, -1, nt);
    return hipGetLastError();
  }
  for (int c = 0; c < tile_ncolors(nt); ++c) {
    dim3 g = grid;
    (DIR == FV3D_IX ? g.x : (DIR == FV3D_IY ? g.y : g.z)) = tile_count(c, nt);
    hipLaunchKernelGGL((k_adj_faces_tiled<real, DIR, RECON, RS>), g, dim3(AF_W * AF_R), 0, s, L.U,
                       L.v, L.sigma, L.Qb, L.dt, P, c, nt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename real, int RECON, int RS>
static hipError_t adj_faces_xy(const AdjLaunch<real>& L, hipStream_t s) {
  hipError_t e = adj_faces_dir<real, FV3D_IX, RECON, RS>(L, s);
  if (e == hipSuccess) e = adj_faces_dir<real, FV3D_IY, RECON, RS>(L, s);
  return e;
}

template <typename real, bool TC, bool VISC>
static void adj_diffusion(const AdjLaunch<real>& L, hipStream_t s) {
  const DevParams& P = L.P;
  if (L.simple) {
    const int bs = 256;
    hipLaunchKernelGGL((k_adj_diffusion<real, TC, VISC>), dim3((unsigned)((P.sf + bs - 1) / bs)),
                       dim3(bs), 0, s, L.U, L.v, L.sigma, L.Qb, L.dt, P, L.grad);
    return;
  }
  if (!L.colored) {
    // z chunks of ~16 planes, at least ~2 workgroups per CU in all
    const int tx = (P.Nx + AZ_X - 1) / AZ_X, ty = (P.Ny + AZ_Y - 1) / AZ_Y;
    int nch = std::max(1, std::min((P.Nz + 7) / 8, (512 + tx * ty - 1) / (tx * ty)));
    const int nzc = (P.Nz + nch - 1) / nch;
    nch = (P.Nz + nzc - 1) / nzc;
    if (sizeof(real) == 4)
      hipLaunchKernelGGL((k_adj_diffusion_pipe<real, TC, VISC>), dim3(tx, ty, nch), dim3(AZ_X * AZ_Y),
                         0, s, L.U, L.v, L.sigma, L.Qb, L.dt, P, L.grad, nzc);
    else
      hipLaunchKernelGGL((k_adj_diffusion_march<real, TC, VISC>), dim3(tx, ty, nch),
                         dim3(AZ_X * AZ_Y), 0, s, L.U, L.v, L.sigma, L.Qb, L.dt, P, L.grad, nzc);
    hipLaunchKernelGGL((k_adj_reduce<real>), dim3(1), dim3(256), 0, s, L.grad,
                       (long long)tx * ty * nch);
    return;
  }
  const dim3 grid((P.Nx + AD_X - 1) / AD_X, (P.Ny + AD_Y - 1) / AD_Y, (P.Nz + AD_Z - 1) / AD_Z);
  const int3 nt = make_int3(grid.x, grid.y, grid.z);
  const long long ntl = (long long)nt.x * nt.y * nt.z;
  if (!L.colored_diffusion) {
    hipLaunchKernelGGL((k_adj_diffusion_tiled<real, TC, VISC>), grid, dim3(AD_X * AD_Y * AD_Z), 0, s,
                       L.U, L.v, L.sigma, L.Qb, L.dt, P, L.grad, make_int3(-1, -1, -1), nt);
    hipLaunchKernelGGL((k_adj_reduce<real>), dim3(1), dim3(256), 0, s, L.grad, ntl);
    return;
  }
  for (int cz = 0; cz < tile_ncolors(nt.z); ++cz)
    for (int cy = 0; cy < tile_ncolors(nt.y); ++cy)
      for (int cx = 0; cx < tile_ncolors(nt.x); ++cx) {
        const dim3 g(tile_count(cx, nt.x), tile_count(cy, nt.y), tile_count(cz, nt.z));
        hipLaunchKernelGGL((k_adj_diffusion_tiled<real, TC, VISC>), g, dim3(AD_X * AD_Y * AD_Z), 0, s,
                           L.U, L.v, L.sigma, L.Qb, L.dt, P, L.grad, make_int3(cx, cy, cz), nt);
      }
  hipLaunchKernelGGL((k_adj_reduce<real>), dim3(1), dim3(256), 0, s, L.grad, ntl);
}

// one adjoint stage: Qb must be zero on entry and is zero on return.  The
// z faces go last, after the diffusion transpose, so that in the default
// path they can finalize the cells only they still add to.
template <typename real>
static hipError_t launch_adjoint_stage_t(const AdjLaunch<real>& L, hipStream_t s) {
  const bool plm = L.recon == FV3D_PLM, hll = L.riemann == FV3D_HLL;
  hipError_t e;
  if (plm && hll) e = adj_faces_xy<real, FV3D_PLM, FV3D_HLL>(L, s);
  else if (plm) e = adj_faces_xy<real, FV3D_PLM, FV3D_HLLC>(L, s);
  else if (hll) e = adj_faces_xy<real, FV3D_PCM, FV3D_HLL>(L, s);
  else e = adj_faces_xy<real, FV3D_PCM, FV3D_HLLC>(L, s);
  if (e != hipSuccess) return e;
  if (L.tc && L.visc) adj_diffusion<real, true, true>(L, s);
  else if (L.tc) adj_diffusion<real, true, false>(L, s);
  else if (L.visc) adj_diffusion<real, false, true>(L, s);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  int Fz = 0;
  int* fz = (L.simple || L.colored || L.no_fused_finalize) ? nullptr : &Fz;
  if (plm && hll) e = adj_faces_dir<real, FV3D_IZ, FV3D_PLM, FV3D_HLL>(L, s, fz);
  else if (plm) e = adj_faces_dir<real, FV3D_IZ, FV3D_PLM, FV3D_HLLC>(L, s, fz);
  else if (hll) e = adj_faces_dir<real, FV3D_IZ, FV3D_PCM, FV3D_HLL>(L, s, fz);
  else e = adj_faces_dir<real, FV3D_IZ, FV3D_PCM, FV3D_HLLC>(L, s, fz);
  if (e != hipSuccess) return e;
  const int bs = 256;
  const unsigned nb = (unsigned)((L.P.sf + bs - 1) / bs);
  hipLaunchKernelGGL((k_adj_finalize<real>), dim3(nb), dim3(bs), 0, s, L.U, L.v, L.sigma,
                     L.extra, L.alpha, L.Qb, L.out, L.P, Fz);
  return hipGetLastError();
}

}  // namespace fv3d_hip
