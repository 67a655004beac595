// fv3d_stage_impl.h -- the fused finite-volume stage kernel for CDNA4 (gfx950).
//
// One launch advances every interior cell by one Euler stage
//     Unew = U + dt*L(prim(U))                          (Update.h:164-178)
// and, when Urk is given, applies the SSP-RK2 combine 0.5*(U0 + Unew)
// (Update.h:198-204) before the single store.  When inv_next is given the
// epilogue also reduces the next step's inv_dt_hyp (ComputeDt.h:28-54) from
// the stored state, so the step never leaves the device for dt.
//
// It replaces, in one pass over HBM, the reference kernels K2-K12:
// ghost fill (as index maps), cons->prim, minmod slopes, reconstruction,
// HLL/HLLC, flux divergence with gravity and the density floor, thermal
// conduction and viscosity.
//
// Work layout.  A workgroup owns a tile of WX (x, 64 or 32) by TY (y)
// columns and marches through a chunk of z planes.  Its window is (TY+4)
// rows by WX+4 columns (2 halo cells each side).  Every window row is owned
// by WX lanes (x = lane, coalesced loads): one row per wave for 64-column
// tiles, two rows per wave (lanes 0-31, 32-63) for 32-column tiles.  The
// waves of the halo rows also own the 4 halo columns of every row.  Every
// row owner keeps its cell's primitives for planes k, k+1, k+2 in
// registers, so z neighbours, z slopes and z faces never touch LDS.  Twelve
// waves (64 x 8) leave 168 VGPRs per lane (3 waves per SIMD), which the FP64
// solver needs without spilling.  Per plane the owners publish plane k to
// LDS, waves compute each x and y face once (the reference solves each face
// twice, Update.h:125-126, with identical inputs), and the interior rows
// update their cells.  The 32-column tile wastes no lanes on 32- and
// 96-cell-wide boxes and re-converts fewer halo cells per interior cell
// (36 x 20 window cells for 32 x 16, against 68 x 12 for 64 x 8).
#pragma once
#include <type_traits>
#include "fv3d_device.h"
#include "fv3d_kernels.h"
#include "fv3d_physics.h"

namespace fv3d_hip {

// Stage modes.  STAGE_PLAIN: Unew = U + dt L(U) (+ RK2 combine with Urk).
// FP32 SSP-RK2 with a compensated accumulation uses two further modes:
// STAGE_INC1 is the first stage (U* = U^n + dt L(U^n), stored as is, into a
// buffer of its own); STAGE_INC2 evaluates U** = U* + dt L(U*) from U* and
// adds 0.5 (U** - U^n) to U^n (Urk) with Kahan compensation (array Cio).
// The two differ from STAGE_PLAIN in the buffers they take in the batched
// kernel (the compensated state alternates between two arrays).
enum { STAGE_PLAIN = 0, STAGE_INC1 = 1, STAGE_INC2 = 2 };

// consToPrim of a loaded cell; CHECK (the FP64 kernels with FAST faces)
// also tells whether the cell is outside the FAST range (bad).  The
// conversion keeps its correctly rounded quotients: a FAST form would need
// a per-lane fallback branch, which cost more than it saved.
template <typename real, bool CHECK>
__device__ __forceinline__ Prim<real> cons_to_prim_checked(real u0, real u1, real u2, real u3,
                                                           real u4, real gm1, bool& bad) {
  const Prim<real> q = cons_to_prim(u0, u1, u2, u3, u4, gm1);
  bad = CHECK && sizeof(real) == 8 &&
        !cell_in_fast_range((double)u0, (double)u1, (double)u2, (double)u3, (double)q.p);
  return q;
}

template <typename real, bool CHECK = false>
__device__ __forceinline__ Prim<real> load_prim(const real* __restrict__ U, long long base,
                                                long long sf, real gm1, bool fx, bool fy,
                                                bool fz, bool& bad) {
  const real u0 = U[base], u1 = U[base + sf], u2 = U[base + 2 * sf], u3 = U[base + 3 * sf],
             u4 = U[base + 4 * sf];
  Prim<real> q = cons_to_prim_checked<real, CHECK>(u0, u1, u2, u3, u4, gm1, bad);
  // reflecting ghosts: BoundaryConditions.h:45-50 negates the normal velocity
  if (fx) q.u = -q.u;
  if (fy) q.v = -q.v;
  if (fz) q.w = -q.w;
  return q;
}

// slope at b from its neighbours a (minus side) and c (plus side).
// x and y: minmod(backward, forward); z: minmod(forward, backward)
// (Update.h:66-84).
template <typename real, bool ZORD>
__device__ __forceinline__ Prim<real> slope3(const Prim<real>& a, const Prim<real>& b,
                                             const Prim<real>& c) {
  Prim<real> s;
  if (!ZORD) {
    s.r = minmod(b.r - a.r, c.r - b.r);
    s.u = minmod(b.u - a.u, c.u - b.u);
    s.v = minmod(b.v - a.v, c.v - b.v);
    s.w = minmod(b.w - a.w, c.w - b.w);
    s.p = minmod(b.p - a.p, c.p - b.p);
  } else {
    s.r = minmod(c.r - b.r, b.r - a.r);
    s.u = minmod(c.u - b.u, b.u - a.u);
    s.v = minmod(c.v - b.v, b.v - a.v);
    s.w = minmod(c.w - b.w, b.w - a.w);
    s.p = minmod(c.p - b.p, b.p - a.p);
  }
  return s;
}

// PLM face states (Update.h:21): q + (slope*sign)*0.5 with sign = +1 on the
// plus face and -1 on the minus face; (s*-1)*0.5 == -(s*0.5) exactly.
template <typename real>
__device__ __forceinline__ Prim<real> plus_face(const Prim<real>& q, const Prim<real>& s) {
  return Prim<real>{q.r + s.r * (real)0.5, q.u + s.u * (real)0.5, q.v + s.v * (real)0.5,
                    q.w + s.w * (real)0.5, q.p + s.p * (real)0.5};
}
template <typename real>
__device__ __forceinline__ Prim<real> minus_face(const Prim<real>& q, const Prim<real>& s) {
  return Prim<real>{q.r - s.r * (real)0.5, q.u - s.u * (real)0.5, q.v - s.v * (real)0.5,
                    q.w - s.w * (real)0.5, q.p - s.p * (real)0.5};
}

// Flux through the face between cells b and c (a before b, d after c) along
// DIR, in the lab frame.
template <typename real, int DIR, int RECON, int RS, bool FAST = false>
__device__ __forceinline__ void face_flux4(const Prim<real>& a, const Prim<real>& b,
                                           const Prim<real>& c, const Prim<real>& d,
                                           const DevParams& P, real F[5], real& pout) {
  Prim<real> qL, qR;
  if (RECON == FV3D_PLM) {
    const Prim<real> sb = slope3<real, DIR == FV3D_IZ>(a, b, c);
    const Prim<real> sc = slope3<real, DIR == FV3D_IZ>(b, c, d);
    qL = plus_face(b, sb);
    qR = minus_face(c, sc);
  } else {
    qL = b;
    qR = c;
    if (DIR == FV3D_IZ && P.pcm_wb) {  // Update.h:27
      qL.p = b.p + b.r * (real)P.g * (real)P.dz * (real)0.5;
      qR.p = c.p - c.r * (real)P.g * (real)P.dz * (real)0.5;
    }
  }
  riemann<real, RS, FAST>(rotate<real, DIR>(qL), rotate<real, DIR>(qR), (real)P.gamma,
                          (real)P.gm1, (real)P.entho, F, pout);
  unrotate_flux<real, DIR>(F);
}

// face_flux4 along x or along y, chosen per lane (isy): on every lane the
// same operations on the same values as face_flux4<real, FV3D_IX or
// FV3D_IY, ...>; the rotation into the face frame and back is a select.
template <typename real, int RECON, int RS, bool FAST = false>
__device__ __forceinline__ void face_flux_xy(const Prim<real>& a, const Prim<real>& b,
                                             const Prim<real>& c, const Prim<real>& d, bool isy,
                                             const DevParams& P, real F[5], real& pout) {
  Prim<real> qL, qR;
  if (RECON == FV3D_PLM) {
    const Prim<real> sb = slope3<real, false>(a, b, c);
    const Prim<real> sc = slope3<real, false>(b, c, d);
    qL = plus_face(b, sb);
    qR = minus_face(c, sc);
  } else {
    qL = b;
    qR = c;
  }
  auto rot = [&](const Prim<real>& q) {
    return Prim<real>{q.r, isy ? q.v : q.u, isy ? q.u : q.v, q.w, q.p};
  };
  riemann<real, RS, FAST>(rot(qL), rot(qR), (real)P.gamma, (real)P.gm1, (real)P.entho, F, pout);
  const real f1 = F[1], f2 = F[2];
  F[1] = isy ? f2 : f1;
  F[2] = isy ? f1 : f2;
}

// One tile of one box: the z chunk `tile` (x fastest, then y, then z chunk)
// marched by the whole workgroup.  Called by the single-box kernel (one tile
// per workgroup) and by the persistent multi-box kernel (a work list of
// (box, tile) items).
template <typename real, int TY, int RECON, int RS, bool TC, bool VISC, int MODE,
          int VAR = STAGE_VAR_GENERAL, int WX = 64>
__device__ __forceinline__ void stage_tile(const real* __restrict__ Uin, const real* Urk,
                                           real* Uout, real* Cio,
                                           const unsigned long long* __restrict__ inv_cur,
                                           unsigned long long* inv_next,
                                           unsigned long long* inv_zero, int* health,
                                           const DevParams& P, const int tile) {
  constexpr int NR = TY + 4;   // window rows
  constexpr int NC = WX + 4;   // window columns
  constexpr int RPW = 64 / WX;  // window rows per wave
  constexpr int NW = NR / RPW;  // waves
  static_assert(WX == 64 || WX == 32, "tile width: one or two rows per wave");
  // interior rows start and end on a wave boundary; every halo column fits
  // the lanes given to it
  static_assert(TY % RPW == 0 && NR <= WX, "tile rows");
  constexpr bool C91Z = VAR == STAGE_VAR_C91;
  constexpr bool LEAN = VAR == STAGE_VAR_LEAN || VAR == STAGE_VAR_LEAN_ANYH;
  constexpr bool POW2 = VAR == STAGE_VAR_LEAN;
  static_assert(!LEAN || RECON == FV3D_PLM, "the lean variants are PLM only");
  // optional features, constant false (or true for the power-of-two cell
  // sizes) in the lean variant; gravity is tested at run time in every
  // variant (keeping it out of the lean one measured no faster)
  const bool wb_zflux = !LEAN && P.wb_flux_zbc, c91_zflux = !LEAN && P.c91_flux_zbc;
  const bool pcm_wb = !LEAN && P.pcm_wb, gravity = P.gravity;
  const bool p2x = POW2 || P.pow2x, p2y = POW2 || P.pow2y, p2z = POW2 || P.pow2z;
  const double* kztab = LEAN ? nullptr : P.kztab;
  // The primitive plane, the temperature plane and the k-1, k+1 velocity
  // planes are double-buffered by plane parity, so the next plane can be
  // published while slower waves still read this one, with no barrier
  // before the publish (the viscous forms, one workgroup per CU, take up to
  // 143 KB of LDS with both copies).
  constexpr int NBUF = 2;
  // viscosity: FP32 evaluates each x and y face once in the face phase
  // (VFACE); FP64 evaluates a cell's six sides in the update from one
  // 3x3x3 stencil, whose shared LDS reads cost less than re-reading each
  // face's stencil (measured 8% faster on the stratified box)
  constexpr bool VFACE = VISC && sizeof(real) == 4;
  // FINE: the face phase and the update are ordered per row instead of by a
  // barrier.  A row's update reads its own faces, the y faces above it and
  // the right edge's x faces (with viscosity and conduction, also the planes
  // published before the face phase); the wave that writes the y faces above
  // a row, and the wave that writes the edge, each raise a flag (the plane
  // index) in LDS after them, and the row's wave waits for those two flags.
  // Older waves then start their update while younger ones still solve
  // faces, and every wave solves its y faces first, so the flags rise early.
  // Taken by every kernel without conduction or viscosity and by the viscous
  // kernels of 64-column tiles (FP32: each row's viscous y face is solved
  // with its y face, before the flag); conduction alone keeps the barrier
  // (not measured).  Measured: 256^3 blast FP64 10.8 -> 11.3 Gcell/s, FP32
  // compensated 23.5 -> 25.3, FP32 plain 26.1 -> 27.6; stratified box FP64
  // 6.10 -> 6.39 and FP32 15.36 -> 15.93; C5 forward run 68.0 -> 63.7 ms;
  // 512-box RT ensemble (32-column tiles) 21.6 -> 22.7.  MOVE: on 32-column
  // tiles the merged edge job of the FP64 and second compensated stages goes
  // from the last wave to wave 0, the oldest, whose flags then rise first
  // (ensemble 22.47 -> 22.57; the first compensated stage ran slower with
  // it, 2151 -> 2337 us).  On 64-column tiles moving the edge jobs to the
  // oldest waves ran slower (FP64 blast -4%, stratified box -8%).
  constexpr bool FINE = (!TC && !VISC) || (VISC && RPW == 1);
  constexpr bool MOVE = FINE && RPW == 2 && (sizeof(real) == 8 || MODE == STAGE_INC2);
  // FP64 PLM-HLL faces (not C91, whose ghost planes are not range-checked)
  // take the FAST quotients and square roots in every plane where all the
  // cells they read are in the FAST range (cell_in_fast_range); other
  // planes, HLLC and the FP32 kernels keep their plain forms.  (HLLC in the
  // FAST form measured no faster on the 256^3 blast: its quotients share
  // fewer denominators than HLL's star flux, and the range test costs what
  // the rest saves.)
  constexpr bool FASTK = sizeof(real) == 8 && RECON == FV3D_PLM && RS == FV3D_HLL && !C91Z;
  __shared__ real sqb[NBUF][5][NR][NC];
  __shared__ real sTb[NBUF][TC ? NR : 1][TC ? NC : 1];
  __shared__ real svmb[VISC ? NBUF : 1][VISC ? 3 : 1][VISC ? NR : 1][VISC ? NC : 1];
  __shared__ real svpb[VISC ? NBUF : 1][VISC ? 3 : 1][VISC ? NR : 1][VISC ? NC : 1];
  __shared__ real sfx[5][TY][WX + 1];
  __shared__ real sfy[5][TY + 1][WX];
  // viscosity: mu * (stress terms) of each x face and y face, evaluated once
  // (a face's two cells read the same values in the same order)
  __shared__ real svx[VFACE ? 4 : 1][VFACE ? TY : 1][VFACE ? WX + 1 : 1];
  __shared__ real svy[VFACE ? 4 : 1][VFACE ? TY + 1 : 1][VFACE ? WX : 1];
  __shared__ double sred[NW];
  __shared__ int sfl[2][NW];  // FASTK: per wave, per plane parity
  // FINE: the last plane whose y face row j (0..TY) is in sfy, and whose
  // right edge x faces are in sfx
  __shared__ int syf[FINE ? TY + 1 : 1];
  __shared__ int sxf;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int xl = threadIdx.x % WX;  // interior column of the lane's row

  const int tx = tile % P.tiles_x;
  const int t2 = tile / P.tiles_x;
  const int ty = t2 % P.tiles_y;
  const int tz = t2 / P.tiles_y;
  const int i0 = tx * WX, j0 = ty * TY, k0 = tz * P.nzc;
  const int k1 = min(k0 + P.nzc, P.Nz);

  // every thread owns window cell (r, c), c = 2..WX+1.  The halo columns
  // 0, 1, WX+2, WX+3 of every row: for 64-column tiles the four halo-row
  // waves own one each in lanes 0..NR-1; for 32-column tiles the two
  // halo-row waves own two each, in lanes 0..NR-1 and 32..32+NR-1.
  const int r = threadIdx.x / WX, c = xl + 2;
  const int gi = i0 - 2 + c, gj = j0 - 2 + r;
  bool fx, fy;
  const int si = map_coord(gi, P.Nx, P.bcx, fx);
  const int sj = map_coord(gj, P.Ny, P.bcy, fy);
  const long long sxy = (long long)si + (long long)sj * P.sy;

  // wave-uniform: interior rows 2..TY+1 start and end on wave boundaries
  const bool interior = r >= 2 && r < TY + 2;
  // rows whose y face below them the row's own wave solves (the top row of
  // y faces too, unless the merged edge job holds it)
  const bool yworker = r >= 2 && r < TY + (MOVE ? 2 : 3);
  const bool store = interior && gi < P.Nx && gj < P.Ny;
  const bool hwave = RPW == 1 ? (wave < 2 || wave >= NW - 2) : (wave == 0 || wave == NW - 1);
  const int hidx = RPW == 1 ? (wave < 2 ? wave : wave - (NW - 4))
                            : (wave == 0 ? 0 : 2) + (lane >> 5);
  const int hl = RPW == 1 ? lane : (lane & 31);
  const bool hown = hwave && hl < NR;
  const int hr = hl < NR ? hl : 0;
  const int hc = hidx < 2 ? hidx : WX + hidx;
  // the right x face of the last interior column, row lane+2.  64-column
  // tiles: the last wave, which computes no other face.  32-column tiles
  // (MERGE): the last wave, or wave 0 (MOVE), lanes 32..32+TY-1, in the same
  // face evaluation as the top row of y faces in its lanes 0-31; with
  // viscosity in FP64 (where the merged evaluation spills 7 more VGPRs) wave
  // 0, lanes < TY
  constexpr bool MERGE = RPW == 2 && !(VISC && sizeof(real) == 8);
  static_assert(!MERGE || TY <= 32, "the edge faces take lanes 32..32+TY-1");
  static_assert(!MOVE || MERGE, "the moved edge job is the merged one");
  const bool xedge = (RPW == 1 ? wave == NW - 1 : !MERGE && wave == 0) && lane < TY;
  const bool mwave = MERGE && wave == (MOVE ? 0 : NW - 1);
  bool hfx, hfy;
  const int hsi = map_coord(i0 - 2 + hc, P.Nx, P.bcx, hfx);
  const int hsj = map_coord(j0 - 2 + hr, P.Ny, P.bcy, hfy);
  const long long hxy = (long long)hsi + (long long)hsj * P.sy;

  if (FINE && threadIdx.x <= TY) syf[threadIdx.x] = k0 - 1;
  if (FINE && threadIdx.x == 0) sxf = k0 - 1;

  const real gm1 = (real)P.gm1;
  const double dtd = dt_from_inv(u64_as_double(*inv_cur), P);
  const real dt = (real)dtd;
  if (inv_zero && tile == 0 && threadIdx.x == 0) *inv_zero = 0ull;

  // C91 z ghosts (C91Z: boundary_z == BC_C91; a template switch, since a
  // run-time test in these loads costs the plain kernels ~8%) come
  // precomputed (fillC91, BoundaryConditions.h:86-134): the column's value
  // with the x/y reflections applied
  auto c91_ghost = [&](int kk, long long xy, bool gx, bool gy) -> Prim<real> {
    const real* g = (const real*)P.c91g + (long long)(kk < 0 ? kk + 2 : kk - P.Nz + 2) * 5 * P.sz + xy;
    Prim<real> q{g[0], g[P.sz], g[2 * P.sz], g[3 * P.sz], g[4 * P.sz]};
    if (gx) q.u = -q.u;
    if (gy) q.v = -q.v;
    return q;
  };
  // bad: the cell is outside the FAST range (always false without FASTK)
  auto load_plane = [&](int kk, bool& bad) -> Prim<real> {
    if (C91Z && (kk < 0 || kk >= P.Nz)) { bad = true; return c91_ghost(kk, sxy, fx, fy); }
    bool fz;
    const int sk = map_coord(kk, P.Nz, P.bcz, fz);
    return load_prim<real, FASTK>(Uin, sxy + (long long)sk * P.sz, P.sf, gm1, fx, fy, fz, bad);
  };
  auto load_halo = [&](int kk, bool& bad) -> Prim<real> {
    if (C91Z && (kk < 0 || kk >= P.Nz)) { bad = true; return c91_ghost(kk, hxy, hfx, hfy); }
    bool fz;
    const int sk = map_coord(kk, P.Nz, P.bcz, fz);
    return load_prim<real, FASTK>(Uin, hxy + (long long)sk * P.sz, P.sf, gm1, hfx, hfy, fz, bad);
  };
  // raw conserved values of the halo cell at plane kk, converted later
  auto fetch_halo = [&](int kk, real h[5]) {
    bool fz;
    const int sk = map_coord(kk, P.Nz, P.bcz, fz);
    const long long base = hxy + (long long)sk * P.sz;
#pragma unroll
    for (int n = 0; n < 5; ++n) h[n] = Uin[base + n * P.sf];
  };

  // raw values of the halo cell at plane kk, or its C91 ghost primitives,
  // and their primitives with the reflections applied
  auto fetch_halo_any = [&](int kk, real h[5]) {
    if (C91Z && (kk < 0 || kk >= P.Nz)) {
      const Prim<real> g = c91_ghost(kk, hxy, hfx, hfy);
      h[0] = g.r; h[1] = g.u; h[2] = g.v; h[3] = g.w; h[4] = g.p;
      return;
    }
    fetch_halo(kk, h);
  };
  auto halo_prim = [&](int kk, const real h[5], bool& bad) -> Prim<real> {
    if (C91Z && (kk < 0 || kk >= P.Nz)) { bad = true; return Prim<real>{h[0], h[1], h[2], h[3], h[4]}; }
    bool hfz;
    map_coord(kk, P.Nz, P.bcz, hfz);
    Prim<real> q = cons_to_prim_checked<real, FASTK>(h[0], h[1], h[2], h[3], h[4], gm1, bad);
    if (hfx) q.u = -q.u;
    if (hfy) q.v = -q.v;
    if (hfz) q.w = -q.w;
    return q;
  };

  // Registers carried from plane to plane.  Interior waves: cz = z slope at
  // k, fz/pz = z face k-1/2 and its pressure.  Halo waves: cz = conserved
  // state of their halo cell at the next plane (prefetched one plane ahead);
  // with viscosity, which also needs the halo cell's velocities at k-1 and
  // k+1, cz = raw values of plane k+1 and fz = primitives of plane k, so no
  // load waits before the plane is published.
  real cz[5] = {0, 0, 0, 0, 0}, fz[5] = {0, 0, 0, 0, 0}, pz = 0;

  // ---- warm-up: planes k0-2 .. k0+1, z slope at k0 and the z face k0-1/2
  // (the z face always in the plain form).  bb: the lane's cell is outside
  // the FAST range in planes k-1, k, k+1 (bits 2..0; plane k+2 joins in the
  // face phase); hb1, with viscosity: so is its halo cell in the plane held
  // in fz (plane k at the top of each plane).
  bool bm1, b0, b1, hb1 = false;
  const Prim<real> Pm1 = load_plane(k0 - 1, bm1);
  Prim<real> P0 = load_plane(k0, b0);
  Prim<real> P1 = load_plane(k0 + 1, b1);
  unsigned bb = (bm1 ? 4u : 0u) | (b0 ? 2u : 0u) | (b1 ? 1u : 0u);
  if (interior) {
    bool bm2;
    const Prim<real> Pm2 = load_plane(k0 - 2, bm2);
    if (RECON == FV3D_PLM) {
      const Prim<real> s = slope3<real, true>(Pm1, P0, P1);
      cz[0] = s.r; cz[1] = s.u; cz[2] = s.v; cz[3] = s.w; cz[4] = s.p;
    }
    face_flux4<real, FV3D_IZ, RECON, RS>(Pm2, Pm1, P0, P1, P, fz, pz);
  } else if (hown) {
    if (VISC) {
      // halo column ring: fz = primitives of plane k, cz = raw values of
      // plane k+1, and the cell's slot in the other plane buffer holds the
      // velocities of plane k-1
      const Prim<real> A = load_halo(k0, hb1);
      fz[0] = A.r; fz[1] = A.u; fz[2] = A.v; fz[3] = A.w; fz[4] = A.p;
      bool bbm;
      const Prim<real> B = load_halo(k0 - 1, bbm);
      real (*sqm)[NR][NC] = sqb[(k0 - 1) & 1];
      sqm[1][hr][hc] = B.u; sqm[2][hr][hc] = B.v; sqm[3][hr][hc] = B.w;
      fetch_halo_any(k0 + 1, cz);
    } else {
      fetch_halo(k0, cz);
    }
  }
  real Tm = 0, T0 = 0, T1 = 0;
  if (TC) { Tm = qdiv(Pm1.p, Pm1.r); T0 = qdiv(P0.p, P0.r); T1 = qdiv(P1.p, P1.r); }
  real um = Pm1.u, vm = Pm1.v, wm = Pm1.w;

  const real kf = (real)P.kface;
  const real mu = (real)P.mu;
  const real oodx = (real)1.0 / (real)P.dx, oody = (real)1.0 / (real)P.dy,
             oodz = (real)1.0 / (real)P.dz;
  const real dtdx = dt / (real)P.dx, dtdy = dt / (real)P.dy, dtdz = dt / (real)P.dz;
  double inv_max = 0.0;
  int bad = 0;

  for (int k = k0; k < k1; ++k) {
    real (*sq)[NR][NC] = sqb[k & 1];
    real (*sT)[TC ? NC : 1] = sTb[k & 1];
    real (*svm)[VISC ? NR : 1][VISC ? NC : 1] = svmb[VISC ? (k & 1) : 0];
    real (*svp)[VISC ? NR : 1][VISC ? NC : 1] = svpb[VISC ? (k & 1) : 0];
    auto lds_prim = [&](int rr, int cc) -> Prim<real> {
      return Prim<real>{sq[0][rr][cc], sq[1][rr][cc], sq[2][rr][cc], sq[3][rr][cc], sq[4][rr][cc]};
    };
    // Viscosity.h's stencil[dk][dj][di] of u, v, w around window cell (rr, cc)
    auto vstencil = [&](int rr0, int cc0) {
      return [&, rr0, cc0](int dk, int dj, int di, int v) -> real {
        const int rr = rr0 + dj - 1, cc = cc0 + di - 1;
        if (dk == 1) return sq[v][rr][cc];
        if (dk == 0) return svm[v - 1][rr][cc];
        return svp[v - 1][rr][cc];
      };
    };
    // plane k+2: loaded here, converted in the face phase (its first use),
    // so its loads are in flight across the publish
    real raw2[5];
    bool flip2 = false;
    const bool ghost2 = C91Z && (k + 2 >= P.Nz);
    if (ghost2) {
      const Prim<real> g = c91_ghost(k + 2, sxy, fx, fy);
      raw2[0] = g.r; raw2[1] = g.u; raw2[2] = g.v; raw2[3] = g.w; raw2[4] = g.p;
    } else {
      const int sk = map_coord(k + 2, P.Nz, P.bcz, flip2);
      const long long base = sxy + (long long)sk * P.sz;
#pragma unroll
      for (int n = 0; n < 5; ++n) raw2[n] = Uin[base + n * P.sf];
    }
    Prim<real> P2{};
    real T2 = 0;
    auto convert_p2 = [&](bool& bad2) {
      if (ghost2) {
        bad2 = true;
        P2 = Prim<real>{raw2[0], raw2[1], raw2[2], raw2[3], raw2[4]};
      } else {
        P2 = cons_to_prim_checked<real, FASTK>(raw2[0], raw2[1], raw2[2], raw2[3], raw2[4], gm1,
                                               bad2);
        if (fx) P2.u = -P2.u;
        if (fy) P2.v = -P2.v;
        if (flip2) P2.w = -P2.w;
      }
      if (TC) T2 = qdiv(P2.p, P2.r);
    };
    Prim<real> H{}, Hm{}, Hp{};
    bool hbad = false;  // the halo cell of plane k
    if (VISC && hown) {
      H = Prim<real>{fz[0], fz[1], fz[2], fz[3], fz[4]};
      hbad = hb1;
      real (*sqm)[NR][NC] = sqb[(k - 1) & 1];  // plane k-1
      Hm.u = sqm[1][hr][hc]; Hm.v = sqm[2][hr][hc]; Hm.w = sqm[3][hr][hc];
      Hp = halo_prim(k + 1, cz, hb1);
    } else if (hown) {
      bool hfz;
      map_coord(k, P.Nz, P.bcz, hfz);
      H = cons_to_prim_checked<real, FASTK>(cz[0], cz[1], cz[2], cz[3], cz[4], gm1, hbad);
      if (hfx) H.u = -H.u;
      if (hfy) H.v = -H.v;
      if (hfz) H.w = -H.w;
    }

    sq[0][r][c] = P0.r; sq[1][r][c] = P0.u; sq[2][r][c] = P0.v; sq[3][r][c] = P0.w;
    sq[4][r][c] = P0.p;
    if (TC) sT[r][c] = T0;
    if (VISC) {
      svm[0][r][c] = um; svm[1][r][c] = vm; svm[2][r][c] = wm;
      svp[0][r][c] = P1.u; svp[1][r][c] = P1.v; svp[2][r][c] = P1.w;
    }
    if (hown) {
      sq[0][hr][hc] = H.r; sq[1][hr][hc] = H.u; sq[2][hr][hc] = H.v; sq[3][hr][hc] = H.w;
      sq[4][hr][hc] = H.p;
      if (TC) sT[hr][hc] = qdiv(H.p, H.r);
      if (VISC) {
        svm[0][hr][hc] = Hm.u; svm[1][hr][hc] = Hm.v; svm[2][hr][hc] = Hm.w;
        svp[0][hr][hc] = Hp.u; svp[1][hr][hc] = Hp.v; svp[2][hr][hc] = Hp.w;
      }
      if (VISC) {
        fz[0] = Hp.r; fz[1] = Hp.u; fz[2] = Hp.v; fz[3] = Hp.w; fz[4] = Hp.p;
        if (k + 1 < k1) fetch_halo_any(k + 2, cz);  // consumed next plane
      } else if (k + 1 < k1) {
        fetch_halo(k + 1, cz);  // consumed next plane
      }
    }
    // FASTK: one flag per wave, a cell of plane k (its x and y faces) is out
    // of range; bb holds planes k-1, k, k+1 in bits 2..0 here
    if (FASTK) {
      const bool wb = __ballot((bb & 2u) != 0u || hbad) != 0ull;  // plane k
      if (lane == 0) sfl[k & 1][wave] = wb ? 1 : 0;
    }
    __syncthreads();
    bool fast = false;
    if (FASTK) {
      int any = 0;
#pragma unroll
      for (int w = 0; w < NW; ++w) any |= sfl[k & 1][w];
      fast = any == 0;
    }

    // the update's base state (and U^n of the RK2 combine) are fetched here,
    // so their latency hides behind the face solves.  These loads are the
    // last use of their lines in this launch, and the stores below the only
    // one: all are non-temporal, so L2 keeps the lines of the planes still
    // to be read by this tile's window and its neighbours' halos (FP64 256^3
    // blast: 1.61x the algorithmic reads fetched with plain loads and
    // stores, 1.48x with non-temporal stores and U^n, 1.13x with all of
    // these).  Hints chosen per launch at run time would merge into plain
    // accesses in the compiled code; a 128^3 box, whose state the caches
    // could keep between launches, pays about 3% in FP32 for them.
    const long long cell = (long long)gi + (long long)gj * P.sy + (long long)k * P.sz;
    real ub[5], ur[5], cc[5];
    auto fetch_u = [&]() {
#pragma unroll
      for (int n = 0; n < 5; ++n) ub[n] = __builtin_nontemporal_load(&Uin[cell + n * P.sf]);
    };
    auto fetch_rk = [&]() {
      if (Urk) {
#pragma unroll
        for (int n = 0; n < 5; ++n) ur[n] = __builtin_nontemporal_load(&Urk[cell + n * P.sf]);
      }
      if (MODE == STAGE_INC2) {  // compensation of the U^n accumulation
#pragma unroll
        for (int n = 0; n < 5; ++n) cc[n] = __builtin_nontemporal_load(&Cio[cell + n * P.sf]);
      }
    };
    // With viscosity the base state is fetched after the face phase, so
    // that it does not hold its registers across the face solves: FP32
    // (whose face phase also evaluates the viscous faces) fetches all of it
    // late; FP64 fetches U^n of the RK2 combine early and the cell's own
    // state late (no spills; all late measured 2.5% slower in the second
    // stage, all early spills 4 VGPRs)
    constexpr bool LATE_U = VISC;
    constexpr bool LATE_RK = VISC && sizeof(real) == 4;
    if (store && !LATE_U) fetch_u();
    if (store && !LATE_RK) fetch_rk();

    real Fzh[5] = {0, 0, 0, 0, 0}, pouth = 0;
    Prim<real> sK1{};
    // The x and y face solves, in the FAST form (FASTK, every cell of plane k
    // in range: uniform over the workgroup) or the plain one.
    // FINE: raise flag f to this plane (stored by the lanes lead) once the
    // wave's LDS writes before it are complete
    auto raise = [&](int* f, bool lead) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lead) __hip_atomic_store(f, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    auto face_xy = [&](auto fast_tag) {
      constexpr bool FA = decltype(fast_tag)::value;
      // ---- y faces: face between window rows r-1 and r
      auto y_faces = [&]() {
        if (yworker && !mwave) {
          real F[5], po;
          face_flux4<real, FV3D_IY, RECON, RS, FA>(lds_prim(r - 2, c), lds_prim(r - 1, c), P0,
                                                   lds_prim(r + 1, c), P, F, po);
#pragma unroll
          for (int n = 0; n < 5; ++n) sfy[n][r - 2][xl] = F[n];
        }
      };
      if constexpr (FINE) {
        y_faces();
        if (VFACE && yworker) {  // the viscous y face below the row, before its flag
          real t[5];
          visc_y_side<real>(vstencil(r, c), 1, oodx, oody, oodz, t);
#pragma unroll
          for (int n = 1; n < 5; ++n) svy[n - 1][r - 2][xl] = mu * t[n];
        }
        // one flag per row: lane 0 of each row (lanes 0 and 32 of 32-column
        // tiles); the merged top row rises with the edge below
        if (yworker && !mwave) raise(&syf[r - 2], (lane & (WX - 1)) == 0);
      }
      // ---- x faces: face between window columns c-1 and c for every interior row
      if (interior) {
        real F[5], po;
        face_flux4<real, FV3D_IX, RECON, RS, FA>(lds_prim(r, c - 2), lds_prim(r, c - 1), P0,
                                                 lds_prim(r, c + 1), P, F, po);
#pragma unroll
        for (int n = 0; n < 5; ++n) sfx[n][r - 2][xl] = F[n];
      } else if (xedge) {
        // the right face of the last interior column for row lane+2
        real F[5], po;
        face_flux4<real, FV3D_IX, RECON, RS, FA>(lds_prim(lane + 2, WX),
                                                 lds_prim(lane + 2, WX + 1),
                                                 lds_prim(lane + 2, WX + 2),
                                                 lds_prim(lane + 2, WX + 3), P, F, po);
#pragma unroll
        for (int n = 0; n < 5; ++n) sfx[n][lane][WX] = F[n];
        if (FINE && VFACE) {  // its viscous face, before the edge flag
          real t[5];
          visc_x_side<real>(vstencil(lane + 2, WX + 2), 1, oodx, oody, oodz, t);
#pragma unroll
          for (int n = 1; n < 5; ++n) svx[n - 1][lane][WX] = mu * t[n];
        }
      }
      if (mwave) {
        // top y faces (lanes 0-31: row TY+2) and edge x faces (lanes 32..:
        // rows 2..TY+1) in one face evaluation
        const bool isy = lane < 32;
        const int er = lane - 32 + 2;
        if (isy || er < TY + 2) {
          const int r0 = isy ? TY : er, c0 = isy ? c : WX;
          const int dr = isy ? 1 : 0, dc = isy ? 0 : 1;
          real F[5], po;
          face_flux_xy<real, RECON, RS, FA>(lds_prim(r0, c0), lds_prim(r0 + dr, c0 + dc),
                                            lds_prim(r0 + 2 * dr, c0 + 2 * dc),
                                            lds_prim(r0 + 3 * dr, c0 + 3 * dc), isy, P, F, po);
          if (isy) {
#pragma unroll
            for (int n = 0; n < 5; ++n) sfy[n][TY][xl] = F[n];
          } else {
#pragma unroll
            for (int n = 0; n < 5; ++n) sfx[n][er - 2][WX] = F[n];
          }
        }
      }
      if constexpr (FINE) {
        // the edge x faces (and, merged, the top row of y faces)
        if (RPW == 1 ? wave == NW - 1 : mwave) {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          if (lane == 0) {
            __hip_atomic_store(&sxf, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (MERGE)
              __hip_atomic_store(&syf[TY], k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        }
      } else {
        y_faces();
      }
    };
    // The two extra face jobs of a 64-column tile, the top row of y faces
    // (window row TY+2) and the right edge's x faces (the last wave), sit on
    // the youngest waves, and a SIMD issues its oldest wave first: their
    // flags rose late, and the top interior row and every update waited
    // for them.  Those two waves solve their faces at priority 1, the others
    // stay at 0 (the same bits).  FP64 blast 11.3 -> 11.8-11.9 Gcell/s, FP32
    // plain 27.6 -> 28.3-28.7, stratified FP64 6.38 -> 6.57-6.63; the
    // compensated FP32 forms level.  The FP32 viscous form raises the top
    // y-face wave only (both: 1% slower; the top one alone: stratified 16.46
    // -> 16.67-16.76, 128^3 forward run 26.4-26.8 -> 25.4-25.5 ms).  The
    // 32-column tiles keep priority 0 (the ensemble 0.5% slower with it).
    // Priority for every wave's y faces as well was slower (11.6 in FP64),
    // and for the top interior row's update level.
    const bool edge_prio = FINE && RPW == 1 &&
                           ((!VISC || sizeof(real) == 8) ? (wave == NW - 1 || r == TY + 2)
                                                         : r == TY + 2);
    if (edge_prio) __builtin_amdgcn_s_setprio(1);
    if constexpr (FASTK) {
      if (fast) face_xy(std::true_type{});
      else face_xy(std::false_type{});
    } else {
      face_xy(std::false_type{});
    }
    if (edge_prio) __builtin_amdgcn_s_setprio(0);
    // plane k+2, then the z face k+1/2 in the FAST form when the wave's
    // cells of planes k-1 .. k+2 are all in range (uniform over the wave)
    bool bad2;
    convert_p2(bad2);
    const unsigned bb4 = ((bb << 1) | (bad2 ? 1u : 0u)) & 15u;
    auto face_z = [&](auto fast_tag) {
      constexpr bool FZ = decltype(fast_tag)::value;
      // ---- z face k+1/2 from registers
      if (interior) {
        Prim<real> qL, qR;
        if (RECON == FV3D_PLM) {
          const Prim<real> sK{cz[0], cz[1], cz[2], cz[3], cz[4]};
          sK1 = slope3<real, true>(P0, P1, P2);
          qL = plus_face(P0, sK);
          qR = minus_face(P1, sK1);
        } else {
          qL = P0;
          qR = P1;
          if (pcm_wb) {
            qL.p = P0.p + P0.r * (real)P.g * (real)P.dz * (real)0.5;
            qR.p = P1.p - P1.r * (real)P.g * (real)P.dz * (real)0.5;
          }
        }
        riemann<real, RS, FZ>(rotate<real, FV3D_IZ>(qL), rotate<real, FV3D_IZ>(qR),
                              (real)P.gamma, gm1, (real)P.entho, Fzh, pouth);
        unrotate_flux<real, FV3D_IZ>(Fzh);
      }
    };
    if constexpr (FASTK) {
      if (__ballot(bb4 != 0u) == 0ull) face_z(std::true_type{});
      else face_z(std::false_type{});
    } else {
      face_z(std::false_type{});
    }
    if (VFACE) {
      // viscous x face c-1/2 of the interior rows (side 1 of cell c) and the
      // last column's right face; y face r-1/2 of the interior columns
      if (interior) {
        real t[5];
        visc_x_side<real>(vstencil(r, c), 1, oodx, oody, oodz, t);
#pragma unroll
        for (int n = 1; n < 5; ++n) svx[n - 1][r - 2][xl] = mu * t[n];
      } else if (xedge && !FINE) {
        real t[5];
        visc_x_side<real>(vstencil(lane + 2, WX + 2), 1, oodx, oody, oodz, t);
#pragma unroll
        for (int n = 1; n < 5; ++n) svx[n - 1][lane][WX] = mu * t[n];
      } else if (mwave && lane >= 32 && lane - 32 < TY) {
        real t[5];
        visc_x_side<real>(vstencil(lane - 30, WX + 2), 1, oodx, oody, oodz, t);
#pragma unroll
        for (int n = 1; n < 5; ++n) svx[n - 1][lane - 32][WX] = mu * t[n];
      }
      if (yworker && !FINE) {
        real t[5];
        visc_y_side<real>(vstencil(r, c), 1, oodx, oody, oodz, t);
#pragma unroll
        for (int n = 1; n < 5; ++n) svy[n - 1][r - 2][xl] = mu * t[n];
      }
    }
    if (!FINE) __syncthreads();

    // ---- update (Update.h:145-160, then ThermalConduction.h, Viscosity.h)
    if (store) {
      if (LATE_U) fetch_u();
      if (LATE_RK) fetch_rk();
      real u[5];
#pragma unroll
      for (int n = 0; n < 5; ++n) u[n] = ub[n];
      // the face fluxes' differences.  The empty asm statements take all 20
      // flux values as operands, so their LDS reads are issued together: left
      // to itself, the FP64 kernels' register allocation sent them out one
      // pair at a time, each waiting for the last (ten LDS round trips per
      // interior wave and plane)
      if (FINE) {
        // the y faces above this row and the right edge's x faces.  Every
        // wave raises its flags once per plane, so the waits end; the bound
        // only turns a fault in that into an error flag instead of a hang.
        auto wait_for = [&](int* f) {
          for (int it = 0; __hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < k;
               ++it) {
            if (it >= (1 << 18)) { bad = 1; break; }
            __builtin_amdgcn_s_sleep(1);
          }
        };
        wait_for(&syf[r - 1]);
        wait_for(&sxf);
      }
      real dx[5], dy[5];
      {
        real a[5], b[5], e[5], f[5];
#pragma unroll
        for (int n = 0; n < 5; ++n) {
          a[n] = sfx[n][r - 2][xl];
          b[n] = sfx[n][r - 2][xl + 1];
          e[n] = sfy[n][r - 2][xl];
          f[n] = sfy[n][r - 1][xl];
        }
        asm volatile("" ::"v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(b[0]), "v"(b[1]),
                     "v"(b[2]), "v"(b[3]), "v"(b[4]));
        asm volatile("" ::"v"(e[0]), "v"(e[1]), "v"(e[2]), "v"(e[3]), "v"(e[4]), "v"(f[0]), "v"(f[1]),
                     "v"(f[2]), "v"(f[3]), "v"(f[4]));
#pragma unroll
        for (int n = 0; n < 5; ++n) {
          dx[n] = a[n] - b[n];
          dy[n] = e[n] - f[n];
        }
      }
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        const real t = dx[n] * dt;
        u[n] = u[n] + over_h(t, p2x, P.idx, P.dx);
      }
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        const real t = dy[n] * dt;
        u[n] = u[n] + over_h(t, p2y, P.idy, P.dy);
      }
      real zl[5], zh[5];
#pragma unroll
      for (int n = 0; n < 5; ++n) { zl[n] = fz[n]; zh[n] = Fzh[n]; }
      if (wb_zflux && (k == 0 || k == P.Nz - 1)) {  // Update.h:132-137
        if (k == 0) {
          zl[0] = 0; zl[1] = 0; zl[2] = 0; zl[4] = 0;
          zl[3] = pouth - P0.r * (real)P.g * (real)P.dz;
        } else {
          zh[0] = 0; zh[1] = 0; zh[2] = 0; zh[4] = 0;
          zh[3] = pz + P0.r * (real)P.g * (real)P.dz;
        }
      } else if (c91_zflux && (k == 0 || k == P.Nz - 1)) {  // Update.h:138-143
        if (k == 0) { zl[0] = 0; zl[1] = 0; zl[2] = 0; zl[3] = pz; zl[4] = 0; }
        else { zh[0] = 0; zh[1] = 0; zh[2] = 0; zh[3] = pouth; zh[4] = 0; }
      }
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        const real t = (zl[n] - zh[n]) * dt;
        u[n] = u[n] + over_h(t, p2z, P.idz, P.dz);
      }
      if (gravity) {  // Update.h:148-151
        u[3] = u[3] + dt * P0.r * (real)P.g;
        u[4] = u[4] + dt * (real)0.5 * (zl[0] + zh[0]) * (real)P.g;
      }
      u[0] = fmax((real)1.0e-6, u[0]);  // Update.h:160

      if (TC) {  // ThermalConduction.h:47-95
        // face conductivities: constant, or B02's z profile (kappaL..kappaB)
        real kxy = kf, kzF = kf, kzB = kf;
        if (kztab) {
          kxy = (real)kztab[k];
          kzF = (real)kztab[P.Nz + k];
          kzB = (real)kztab[2 * P.Nz + k];
        }
        u[4] = u[4] + conduction_increment<real>(T0, sT[r][c - 1], sT[r][c + 1], sT[r - 1][c],
                                                 sT[r + 1][c], Tm, T1, kxy, kzF, kzB, k == 0,
                                                 k == P.Nz - 1, P, p2x, p2y, p2z, dtdx, dtdy,
                                                 dtdz);
      }

      if (VISC && !VFACE) {  // Viscosity.h:28-153, six sides from one stencil
        real vx[5] = {0, 0, 0, 0, 0}, vy[5] = {0, 0, 0, 0, 0}, vz[5] = {0, 0, 0, 0, 0};
        auto S = [&](int dk, int dj, int di, int v) -> real {
          const int rr = r + dj - 1, cc = c + di - 1;
          if (dk == 1) return sq[v][rr][cc];
          if (dk == 0) return svm[v - 1][rr][cc];
          return svp[v - 1][rr][cc];
        };
        viscous_fluxes<real>(S, oodx, oody, oodz, mu, vx, vy, vz);
#pragma unroll
        for (int n = 1; n < 5; ++n) u[n] = u[n] + (vx[n] + vy[n] + vz[n]) * dt;
      }
      if (VFACE) {  // Viscosity.h:28-153
        // x and y: the two faces' terms from LDS; side 1 adds (-mu) t =
        // -(mu t) to the sum that starts at 0, side 2 adds mu t
        real vx[5] = {0, 0, 0, 0, 0}, vy[5] = {0, 0, 0, 0, 0}, vz[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int n = 1; n < 5; ++n) {
          vx[n] += -svx[n - 1][r - 2][xl];
          vx[n] += svx[n - 1][r - 2][xl + 1];
          vy[n] += -svy[n - 1][r - 2][xl];
          vy[n] += svy[n - 1][r - 1][xl];
        }
        // z: both sides per cell (the reference's z branch differs between
        // a face's two cells)
#pragma unroll
        for (int sd = 1; sd < 3; ++sd) {
          real t[5];
          visc_z_side<real>(vstencil(r, c), sd, oodx, oody, oodz, t);
          visc_accumulate(vz, sd, mu, t);
        }
#pragma unroll
        for (int n = 1; n < 5; ++n) u[n] = u[n] + (vx[n] + vy[n] + vz[n]) * dt;
      }

      if (MODE == STAGE_INC2) {
        // U^{n+1} = U^n + 0.5 (U** - U^n), Kahan-compensated
#pragma unroll
        for (int n = 0; n < 5; ++n) {
          const real inc = (real)0.5 * (u[n] - ur[n]);
          const real y = inc - cc[n];
          const real t = ur[n] + y;
          __builtin_nontemporal_store((t - ur[n]) - y, &Cio[cell + n * P.sf]);
          __builtin_nontemporal_store(t, &Uout[cell + n * P.sf]);
          u[n] = t;
        }
      } else {
        if (Urk) {  // Update.h:203
#pragma unroll
          for (int n = 0; n < 5; ++n) u[n] = (real)0.5 * (ur[n] + u[n]);
        }
#pragma unroll
        for (int n = 0; n < 5; ++n) __builtin_nontemporal_store(u[n], &Uout[cell + n * P.sf]);
      }

      if (inv_next) {  // next step's inv_dt_hyp from the stored state (ComputeDt.h:32-35)
        double ih;
        if (sizeof(real) == 8) {
          const Prim<double> q = cons_to_prim<double>(u[0], u[1], u[2], u[3], u[4], P.gm1);
          const double cs = sound_speed(q.r, q.p, P.gamma);
          const double a = cs + fabs(q.u), bb = cs + fabs(q.v), cc = cs + fabs(q.w);
          ih = (p2x ? a * P.idx : a / P.dx) + (p2y ? bb * P.idy : bb / P.dy) +
               (p2z ? cc * P.idz : cc / P.dz);
        } else {
          // FP32 state: the same FP64 expression, with its quotients and square
          // root to about an ulp of FP64 (Newton-refined rcp and rsq) instead of
          // the correctly rounded sequences; far below the state's FP32 rounding
          const double r0 = u[0], ir = recip_f64(r0);
          const double vx = u[1] * ir, vy = u[2] * ir, vz = u[3] * ir;
          const double pr = ((double)u[4] - 0.5 * r0 * (vx * vx + vy * vy + vz * vz)) * P.gm1;
          const double cs = sqrt_f64(pr * P.gamma * ir);
          ih = (cs + fabs(vx)) * P.idx + (cs + fabs(vy)) * P.idy + (cs + fabs(vz)) * P.idz;
        }
        if (ih > inv_max) inv_max = ih;
        if (!(ih <= 1.0e300)) bad = 1;  // NaN/inf: the reference's Max reducer drops NaN
      }
    }

    // ---- roll the z window
    if (VISC) { um = P0.u; vm = P0.v; wm = P0.w; }
    if (TC) { Tm = T0; T0 = T1; T1 = T2; }
    P0 = P1;
    P1 = P2;
    bb = bb4 & 7u;
    if (interior) {
      if (RECON == FV3D_PLM) { cz[0] = sK1.r; cz[1] = sK1.u; cz[2] = sK1.v; cz[3] = sK1.w; cz[4] = sK1.p; }
#pragma unroll
      for (int n = 0; n < 5; ++n) fz[n] = Fzh[n];
      pz = pouth;
    }
  }

  if (inv_next) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(inv_max, off);
      inv_max = (o > inv_max) ? o : inv_max;
    }
    if (lane == 0) sred[wave] = inv_max;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = 0.0;
      for (int w = 0; w < NW; ++w) m = (sred[w] > m) ? sred[w] : m;
      if (m > 0.0) atomicMax(inv_next, double_as_u64(m));
    }
    if (bad) atomicOr(health, 1);
  }
}

template <typename real, int TY, int RECON, int RS, bool TC, bool VISC, int MODE, int VAR, int WX>
__global__ void __launch_bounds__(WX * (TY + 4))
stage_kernel(const real* __restrict__ Uin, const real* Urk, real* Uout, real* Cio,
             const unsigned long long* __restrict__ inv_cur, unsigned long long* inv_next,
             unsigned long long* inv_zero, int* health, DevParams P) {
  // Blocks are dealt round-robin to the 8 XCDs (block b -> XCD b%8).  Give
  // each XCD a contiguous run of tiles so neighbouring tiles share halo rows
  // in that XCD's L2.
  const int nb = gridDim.x, b = blockIdx.x;
  const int q8 = nb >> 3, r8 = nb & 7, xcd = b & 7, loc = b >> 3;
  const int tile = xcd * q8 + (xcd < r8 ? xcd : r8) + loc;
  stage_tile<real, TY, RECON, RS, TC, VISC, MODE, VAR, WX>(Uin, Urk, Uout, Cio, inv_cur, inv_next,
                                                     inv_zero, health, P, tile);
}

// ---------------------------------------------------------------------------
// launcher
// ---------------------------------------------------------------------------
template <typename real, int TY, int RECON, int RS, bool TC, bool VISC, int MODE, int VAR, int WX>
static hipError_t launch_t(const StageLaunch<real>& L, hipStream_t s) {
  DevParams P = L.P;
  // the tile counts were sized for this shape
  if (P.twx != WX || P.tty != TY) return hipErrorInvalidValue;
  const int blocks = P.tiles_x * P.tiles_y * P.tiles_z;
  hipLaunchKernelGGL((stage_kernel<real, TY, RECON, RS, TC, VISC, MODE, VAR, WX>), dim3(blocks),
                     dim3(WX * (TY + 4)), 0, s, L.Uin, L.Urk, L.Uout, L.Cio, L.inv_cur,
                     L.inv_next, L.inv_zero, L.health, P);
  return hipGetLastError();
}

template <typename real, int RECON, int RS, int MODE, int VAR, int WX>
static hipError_t launch_phys(const StageLaunch<real>& L, hipStream_t s) {
  const bool tc = L.tc, vi = L.visc;
  constexpr int TV = stage_ty_of<real, true, WX>(), TH = stage_ty_of<real, false, WX>();
  if (tc && vi) return launch_t<real, TV, RECON, RS, true, true, MODE, VAR, WX>(L, s);
  if (tc) return launch_t<real, TH, RECON, RS, true, false, MODE, VAR, WX>(L, s);
  if (vi) return launch_t<real, TV, RECON, RS, false, true, MODE, VAR, WX>(L, s);
  return launch_t<real, TH, RECON, RS, false, false, MODE, VAR, WX>(L, s);
}

// fillC91 (BoundaryConditions.h:86-134) for the z-ghost planes of every
// domain column, from the stage's input state, as primitives.  The
// reference multiplies by the gravity flag where g belongs (its Q17); so
// does this.
template <typename real>
__global__ void __launch_bounds__(256)
k_c91_ghosts(const real* __restrict__ U, real* g, DevParams P) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= 4 * P.sz) return;
  const int plane = (int)(t / P.sz);
  const long long xy = t % P.sz;
  const int kk = plane < 2 ? plane - 2 : P.Nz + plane - 2;
  const int kref = kk < 0 ? 0 : P.Nz - 1;
  const int ksym = kk < 0 ? -kk - 1 : 2 * P.Nz - kk - 1;
  const real gm1 = (real)P.gm1;
  bool nb;
  const Prim<real> qr = load_prim<real>(U, xy + (long long)kref * P.sz, P.sf, gm1, false, false, false, nb);
  const Prim<real> qs = load_prim<real>(U, xy + (long long)ksym * P.sz, P.sf, gm1, false, false, false, nb);
  const real z = (real)P.zmin + ((real)kk + (real)0.5) * (real)P.dz;
  const real grav = P.gravity ? (real)1 : (real)0;
  real res[5];
  if (kk < 0) {
    const real dz = z - (real)0.5 * (real)P.dz;
    const real T = (real)1.0 + z * (real)P.theta1;
    res[4] = qr.p + dz * qr.r * grav;
    res[0] = res[4] / T;
    res[4] = fmax(res[4], (real)1.0e-6);
    res[0] = fmax(res[0], (real)1.0e-6);
    const real dratio = qs.r / res[0];
    res[1] = qs.u;
    res[2] = qs.v;
    res[3] = -qs.w * dratio;
  } else {
    const real Tref = qr.p / qr.r;
    const real dz = z - ((real)1.0 - (real)0.5 * (real)P.dz);
    const real T = Tref + dz * (real)P.theta1;
    const real Pp = qr.p + dz * qr.r * grav;
    const real rho = Pp / T;
    res[4] = fmax(Pp, (real)1.0e-6);
    res[0] = fmax(rho, (real)1.0e-6);
    const real dratio = qs.r / rho;
    res[1] = qs.u * dratio;
    res[2] = qs.v * dratio;
    res[3] = -qs.w * dratio;
  }
  real* d = g + (long long)plane * 5 * P.sz + xy;
#pragma unroll
  for (int n = 0; n < 5; ++n) d[n * P.sz] = res[n];
}

template <typename real, int MODE, int WX>
hipError_t launch_stage_wx(const StageLaunch<real>& L, hipStream_t s) {
  if (L.P.bcz == FV3D_BC_C91) {
    const long long n = 4 * L.P.sz;
    hipLaunchKernelGGL((k_c91_ghosts<real>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       L.Uin, (real*)L.P.c91g, L.P);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const bool plm = L.recon == FV3D_PLM;
  const bool hll = L.riemann == FV3D_HLL;
  constexpr int G = STAGE_VAR_GENERAL, C = STAGE_VAR_C91, N = STAGE_VAR_LEAN,
                H = STAGE_VAR_LEAN_ANYH;
  const int var = stage_variant(L.P, L.recon);
  if (var == C) {
    if (plm && hll) return launch_phys<real, FV3D_PLM, FV3D_HLL, MODE, C, WX>(L, s);
    if (plm) return launch_phys<real, FV3D_PLM, FV3D_HLLC, MODE, C, WX>(L, s);
    if (hll) return launch_phys<real, FV3D_PCM, FV3D_HLL, MODE, C, WX>(L, s);
    return launch_phys<real, FV3D_PCM, FV3D_HLLC, MODE, C, WX>(L, s);
  }
  if (var == N) {
    if (hll) return launch_phys<real, FV3D_PLM, FV3D_HLL, MODE, N, WX>(L, s);
    return launch_phys<real, FV3D_PLM, FV3D_HLLC, MODE, N, WX>(L, s);
  }
  if (var == H) {
    if (hll) return launch_phys<real, FV3D_PLM, FV3D_HLL, MODE, H, WX>(L, s);
    return launch_phys<real, FV3D_PLM, FV3D_HLLC, MODE, H, WX>(L, s);
  }
  if (plm && hll) return launch_phys<real, FV3D_PLM, FV3D_HLL, MODE, G, WX>(L, s);
  if (plm) return launch_phys<real, FV3D_PLM, FV3D_HLLC, MODE, G, WX>(L, s);
  if (hll) return launch_phys<real, FV3D_PCM, FV3D_HLL, MODE, G, WX>(L, s);
  return launch_phys<real, FV3D_PCM, FV3D_HLLC, MODE, G, WX>(L, s);
}

}  // namespace fv3d_hip
