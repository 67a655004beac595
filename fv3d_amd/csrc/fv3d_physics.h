// fv3d_physics.h -- the diffusive terms of one cell: thermal conduction's
// energy increment (ThermalConduction.h:47-95) and the viscous flux sums
// (Viscosity.h:38-147).  The fused stage kernel reads the stencil from LDS
// and registers, the per-functor kernels from a ghosted array; both call
// these, so the two paths share one statement of the reference's arithmetic
// (same expression trees, no contraction).
#pragma once
#include "fv3d_device.h"

namespace fv3d_hip {

// ThermalConduction.h:61-95 for one cell: T of the cell (Tc) and of its six
// neighbours (left/right in x, up/down in y, front/back in z), the face
// conductivities (x and y faces kxy, lower z face kzF, upper kzB), and
// whether the cell is on the lower / upper z wall.  Returns the increment
// dt/dx (FR-FL) + dt/dy (FD-FU) + dt/dz (FB-FF) that is added to E.
template <typename real>
__device__ __forceinline__ real conduction_increment(real Tc, real TL, real TR, real TU, real TD,
                                                     real TF, real TB, real kxy, real kzF,
                                                     real kzB, bool zmin_wall, bool zmax_wall,
                                                     const DevParams& P, bool p2x, bool p2y,
                                                     bool p2z, real dtdx, real dtdy, real dtdz) {
  auto ddx = [&](real v) { return over_h(v, p2x, P.idx, P.dx); };
  auto ddy = [&](real v) { return over_h(v, p2y, P.idy, P.dy); };
  auto ddz = [&](real v) { return over_h(v, p2z, P.idz, P.dz); };
  const real FL = ddx(kxy * (Tc - TL));
  const real FR = ddx(kxy * (TR - Tc));
  const real FU = ddy(kxy * (Tc - TU));
  const real FD = ddy(kxy * (TD - Tc));
  real FF = ddz(kzF * (Tc - TF));
  real FB = ddz(kzB * (TB - Tc));
  if (zmin_wall && P.bctc_zmin != FV3D_BCTC_NONE) {
    if (P.bctc_zmin == FV3D_BCTC_FIXED_TEMPERATURE)
      FF = ddz(kzF * (real)2.0 * (Tc - (real)P.bctc_zmin_value));
    else if (P.bctc_zmin == FV3D_BCTC_FIXED_GRADIENT)
      FF = kzF * (real)P.bctc_zmin_value;
  }
  if (zmax_wall && P.bctc_zmax != FV3D_BCTC_NONE) {
    if (P.bctc_zmax == FV3D_BCTC_FIXED_TEMPERATURE)
      FB = ddz(kzB * (real)2.0 * ((real)P.bctc_zmax_value - Tc));
    else if (P.bctc_zmax == FV3D_BCTC_FIXED_GRADIENT)
      FB = kzB * (real)P.bctc_zmax_value;
  }
  return dtdx * (FR - FL) + dtdy * (FD - FU) + dtdz * (FB - FF);
}

// Viscosity.h:38-147.  S(dk, dj, di, v) is the primitive v (1 = u, 2 = v,
// 3 = w) of the cell at offset (dk-1, dj-1, di-1) in z, y, x from the cell
// whose side s (1: lower, 2: upper) is evaluated.  Each side function gives
// the stress terms t[1..4] of that side; the reference adds sign * mu * t
// (sign -1 on side 1, +1 on side 2) to the cell's flux sum of the direction.
//
// An x side is the same face for its two cells, evaluated alike: side 2 of
// cell i and side 1 of cell i+1 read the same values in the same order, so
// one evaluation per x face (and per y face) gives both cells' bits.  The
// z sides carry the reference's index and component mix-ups
// (Viscosity.h:122-129) and differ between a face's two cells.
template <typename real, class Stencil>
__device__ __forceinline__ void visc_x_side(const Stencil& S, int s, real oodx, real oody,
                                            real oodz, real t[5]) {
  const real four_thirds = (real)(4.0 / 3.0);
  const real two_thirds = (real)(2.0 / 3.0);
  const real qiu = (real)0.5 * (S(1, 1, s, 1) + S(1, 1, s - 1, 1));
  const real qiv = (real)0.5 * (S(1, 1, s, 2) + S(1, 1, s - 1, 2));
  const real qiw = (real)0.5 * (S(1, 1, s, 3) + S(1, 1, s - 1, 3));
  const real dudx = oodx * (S(1, 1, s, 1) - S(1, 1, s - 1, 1));
  const real dvdx = oodx * (S(1, 1, s, 2) - S(1, 1, s - 1, 2));
  const real dwdx = oodx * (S(1, 1, s, 3) - S(1, 1, s - 1, 3));
  const real dudy = (real)0.25 * oody * (S(1, 2, s, 1) - S(1, 0, s, 1)
                                         + S(1, 2, s - 1, 1) - S(1, 0, s - 1, 1));
  const real dvdy = (real)0.25 * oody * (S(1, 2, s, 2) - S(1, 0, s, 2)
                                         + S(1, 2, s - 1, 2) - S(1, 0, s - 1, 2));
  const real dudz = (real)0.25 * oodz * (S(2, 1, s, 1) - S(0, 1, s, 1)
                                         + S(2, 1, s - 1, 1) - S(0, 1, s - 1, 1));
  const real dwdz = (real)0.25 * oodz * (S(2, 1, s, 3) - S(0, 1, s, 3)
                                         + S(2, 1, s - 1, 3) - S(0, 1, s - 1, 3));
  const real txx = four_thirds * dudx - two_thirds * (dvdy + dwdz);
  const real txy = dvdx + dudy;
  const real txz = dwdx + dudz;
  t[1] = txx;
  t[2] = txy;
  t[3] = txz;
  t[4] = txx * qiu + txy * qiv + txz * qiw;
}

template <typename real, class Stencil>
__device__ __forceinline__ void visc_y_side(const Stencil& S, int s, real oodx, real oody,
                                            real oodz, real t[5]) {
  const real four_thirds = (real)(4.0 / 3.0);
  const real two_thirds = (real)(2.0 / 3.0);
  const real qiu = (real)0.5 * (S(1, s, 1, 1) + S(1, s - 1, 1, 1));
  const real qiv = (real)0.5 * (S(1, s, 1, 2) + S(1, s - 1, 1, 2));
  const real qiw = (real)0.5 * (S(1, s, 1, 3) + S(1, s - 1, 1, 3));
  const real dudy = oody * (S(1, s, 1, 1) - S(1, s - 1, 1, 1));
  const real dvdy = oody * (S(1, s, 1, 2) - S(1, s - 1, 1, 2));
  const real dwdy = oody * (S(1, s, 1, 3) - S(1, s - 1, 1, 3));
  const real dudx = (real)0.25 * oodx * (S(1, s, 2, 1) - S(1, s, 0, 1)
                                         + S(1, s - 1, 2, 1) - S(1, s - 1, 0, 1));
  const real dvdx = (real)0.25 * oodx * (S(1, s, 2, 2) - S(1, s, 0, 2)
                                         + S(1, s - 1, 2, 2) - S(1, s - 1, 0, 2));
  const real dwdz = (real)0.25 * oodz * (S(2, s, 1, 3) - S(0, s, 1, 3)
                                         + S(2, s - 1, 1, 3) - S(0, s - 1, 1, 3));
  const real dvdz = (real)0.25 * oodz * (S(2, s, 1, 2) - S(0, s, 1, 2)
                                         + S(2, s - 1, 1, 2) - S(0, s - 1, 1, 2));
  const real tyy = four_thirds * dvdy - two_thirds * (dudx + dwdz);
  const real txy = dvdx + dudy;
  const real tyz = dvdz + dwdy;
  t[1] = txy;
  t[2] = tyy;
  t[3] = tyz;
  t[4] = txy * qiu + tyy * qiv + tyz * qiw;
}

// the z side, with the reference's index and component mix-ups
// (Viscosity.h:122-129)
template <typename real, class Stencil>
__device__ __forceinline__ void visc_z_side(const Stencil& S, int s, real oodx, real oody,
                                            real oodz, real t[5]) {
  (void)oody;
  const real four_thirds = (real)(4.0 / 3.0);
  const real two_thirds = (real)(2.0 / 3.0);
  const real qiu = (real)0.5 * (S(s, 1, 1, 1) + S(s - 1, 1, 1, 1));
  const real qiv = (real)0.5 * (S(s, 1, 1, 2) + S(s - 1, 1, 1, 2));
  const real qiw = (real)0.5 * (S(s, 1, 1, 3) + S(s - 1, 1, 1, 3));
  const real dudz = oodz * (S(s, 1, 1, 1) - S(s - 1, 1, 1, 1));
  const real dvdz = oodz * (S(s, 1, 1, 2) - S(s - 1, 1, 1, 2));
  const real dwdz = oodz * (S(s, 1, 1, 3) - S(s - 1, 1, 1, 3));
  const real dudx = (real)0.25 * oodx * (S(s, 1, 2, 1) - S(1, s, 0, 1)
                                         + S(s - 1, 1, 2, 1) - S(1, s - 1, 0, 1));
  const real dwdx = (real)0.25 * oodx * (S(s, 1, 2, 2) - S(1, s, 0, 3)
                                         + S(s - 1, 1, 2, 2) - S(1, s - 1, 0, 3));
  const real dvdy = (real)0.25 * oodz * (S(s, 2, 1, 3) - S(s, 0, 1, 2)
                                         + S(s - 1, 2, 1, 3) - S(s - 1, 0, 1, 2));
  const real dwdy = (real)0.25 * oodz * (S(s, 2, 1, 2) - S(s, 0, 1, 3)
                                         + S(s - 1, 2, 1, 2) - S(s - 1, 0, 1, 3));
  const real tzz = four_thirds * dwdz - two_thirds * (dudx + dvdy);
  const real txz = dudz + dwdx;
  const real tyz = dvdz + dwdy;
  t[1] = txz;
  t[2] = tyz;
  t[3] = tzz;
  t[4] = txz * qiu + tyz * qiv + tzz * qiw;
}

// sum += sign * mu * t over components 1..4 (Viscosity.h:84-87 and alike)
template <typename real>
__device__ __forceinline__ void visc_accumulate(real sum[5], int s, real mu, const real t[5]) {
  const real sgn = (s == 1 ? (real)-1.0 : (real)1.0);
#pragma unroll
  for (int n = 1; n < 5; ++n) sum[n] += sgn * mu * t[n];
}

// The flux sums vx, vy, vz (components 1..4; the caller zeroes them) of
// one cell, side by side as the reference evaluates them.  The update is
// Unew += dt (vx + vy + vz), with no 1/d (Viscosity.h:151).
template <typename real, class Stencil>
__device__ __forceinline__ void viscous_fluxes_by_side(const Stencil& S, real oodx, real oody, real oodz,
                                               real mu, real vx[5], real vy[5], real vz[5]) {
#pragma unroll
  for (int s = 1; s < 3; ++s) {
    real t[5];
    visc_x_side<real>(S, s, oodx, oody, oodz, t);
    visc_accumulate(vx, s, mu, t);
    visc_y_side<real>(S, s, oodx, oody, oodz, t);
    visc_accumulate(vy, s, mu, t);
    visc_z_side<real>(S, s, oodx, oody, oodz, t);
    visc_accumulate(vz, s, mu, t);
  }
}

// Viscosity.h:38-147 for one cell in one block: the flux sums vx, vy, vz
// (components 1..4; the caller zeroes them) from the cell's 3x3x3 stencil,
// with the z branch's mix-ups (Viscosity.h:122-129).  The same arithmetic
// as the side functions above, side by side in the reference's order; the
// fused FP64 kernel keeps this form because built from the side functions
// its register allocation spills 8 VGPRs instead of 4 (GPU tests check
// both forms bit for bit against the oracle: this one through the fused
// kernel, the side functions through the per-pass kernel).
template <typename real, class Stencil>
__device__ __forceinline__ void viscous_fluxes(const Stencil& S, real oodx, real oody, real oodz,
                                               real mu, real vx[5], real vy[5], real vz[5]) {
  const real four_thirds = (real)(4.0 / 3.0);
  const real two_thirds = (real)(2.0 / 3.0);
#pragma unroll
  for (int s = 1; s < 3; ++s) {
    const real sgn = (s == 1 ? (real)-1.0 : (real)1.0);
    {  // X
      const real qiu = (real)0.5 * (S(1, 1, s, 1) + S(1, 1, s - 1, 1));
      const real qiv = (real)0.5 * (S(1, 1, s, 2) + S(1, 1, s - 1, 2));
      const real qiw = (real)0.5 * (S(1, 1, s, 3) + S(1, 1, s - 1, 3));
      const real dudx = oodx * (S(1, 1, s, 1) - S(1, 1, s - 1, 1));
      const real dvdx = oodx * (S(1, 1, s, 2) - S(1, 1, s - 1, 2));
      const real dwdx = oodx * (S(1, 1, s, 3) - S(1, 1, s - 1, 3));
      const real dudy = (real)0.25 * oody * (S(1, 2, s, 1) - S(1, 0, s, 1)
                                             + S(1, 2, s - 1, 1) - S(1, 0, s - 1, 1));
      const real dvdy = (real)0.25 * oody * (S(1, 2, s, 2) - S(1, 0, s, 2)
                                             + S(1, 2, s - 1, 2) - S(1, 0, s - 1, 2));
      const real dudz = (real)0.25 * oodz * (S(2, 1, s, 1) - S(0, 1, s, 1)
                                             + S(2, 1, s - 1, 1) - S(0, 1, s - 1, 1));
      const real dwdz = (real)0.25 * oodz * (S(2, 1, s, 3) - S(0, 1, s, 3)
                                             + S(2, 1, s - 1, 3) - S(0, 1, s - 1, 3));
      const real txx = four_thirds * dudx - two_thirds * (dvdy + dwdz);
      const real txy = dvdx + dudy;
      const real txz = dwdx + dudz;
      vx[1] += sgn * mu * txx;
      vx[2] += sgn * mu * txy;
      vx[3] += sgn * mu * txz;
      vx[4] += sgn * mu * (txx * qiu + txy * qiv + txz * qiw);
    }
    {  // Y
      const real qiu = (real)0.5 * (S(1, s, 1, 1) + S(1, s - 1, 1, 1));
      const real qiv = (real)0.5 * (S(1, s, 1, 2) + S(1, s - 1, 1, 2));
      const real qiw = (real)0.5 * (S(1, s, 1, 3) + S(1, s - 1, 1, 3));
      const real dudy = oody * (S(1, s, 1, 1) - S(1, s - 1, 1, 1));
      const real dvdy = oody * (S(1, s, 1, 2) - S(1, s - 1, 1, 2));
      const real dwdy = oody * (S(1, s, 1, 3) - S(1, s - 1, 1, 3));
      const real dudx = (real)0.25 * oodx * (S(1, s, 2, 1) - S(1, s, 0, 1)
                                             + S(1, s - 1, 2, 1) - S(1, s - 1, 0, 1));
      const real dvdx = (real)0.25 * oodx * (S(1, s, 2, 2) - S(1, s, 0, 2)
                                             + S(1, s - 1, 2, 2) - S(1, s - 1, 0, 2));
      const real dwdz = (real)0.25 * oodz * (S(2, s, 1, 3) - S(0, s, 1, 3)
                                             + S(2, s - 1, 1, 3) - S(0, s - 1, 1, 3));
      const real dvdz = (real)0.25 * oodz * (S(2, s, 1, 2) - S(0, s, 1, 2)
                                             + S(2, s - 1, 1, 2) - S(0, s - 1, 1, 2));
      const real tyy = four_thirds * dvdy - two_thirds * (dudx + dwdz);
      const real txy = dvdx + dudy;
      const real tyz = dvdz + dwdy;
      vy[1] += sgn * mu * txy;
      vy[2] += sgn * mu * tyy;
      vy[3] += sgn * mu * tyz;
      vy[4] += sgn * mu * (txy * qiu + tyy * qiv + tyz * qiw);
    }
    {  // Z, with the reference's index and component mix-ups (Viscosity.h:122-129)
      const real qiu = (real)0.5 * (S(s, 1, 1, 1) + S(s - 1, 1, 1, 1));
      const real qiv = (real)0.5 * (S(s, 1, 1, 2) + S(s - 1, 1, 1, 2));
      const real qiw = (real)0.5 * (S(s, 1, 1, 3) + S(s - 1, 1, 1, 3));
      const real dudz = oodz * (S(s, 1, 1, 1) - S(s - 1, 1, 1, 1));
      const real dvdz = oodz * (S(s, 1, 1, 2) - S(s - 1, 1, 1, 2));
      const real dwdz = oodz * (S(s, 1, 1, 3) - S(s - 1, 1, 1, 3));
      const real dudx = (real)0.25 * oodx * (S(s, 1, 2, 1) - S(1, s, 0, 1)
                                             + S(s - 1, 1, 2, 1) - S(1, s - 1, 0, 1));
      const real dwdx = (real)0.25 * oodx * (S(s, 1, 2, 2) - S(1, s, 0, 3)
                                             + S(s - 1, 1, 2, 2) - S(1, s - 1, 0, 3));
      const real dvdy = (real)0.25 * oodz * (S(s, 2, 1, 3) - S(s, 0, 1, 2)
                                             + S(s - 1, 2, 1, 3) - S(s - 1, 0, 1, 2));
      const real dwdy = (real)0.25 * oodz * (S(s, 2, 1, 2) - S(s, 0, 1, 3)
                                             + S(s - 1, 2, 1, 2) - S(s - 1, 0, 1, 3));
      const real tzz = four_thirds * dwdz - two_thirds * (dudx + dvdy);
      const real txz = dudz + dwdx;
      const real tyz = dvdz + dwdy;
      vz[1] += sgn * mu * txz;
      vz[2] += sgn * mu * tyz;
      vz[3] += sgn * mu * tzz;
      vz[4] += sgn * mu * (txz * qiu + tyz * qiv + tzz * qiw);
    }
  }
}

}  // namespace fv3d_hip
