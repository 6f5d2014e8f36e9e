// g5_step2d_body.h -- fused 2-D five-equation step, one 16x16 block per workgroup (FP64 or mixed:
// the primitive/face path in the arithmetic T, the state accumulation in FP64).
//
// One launch does, for every owned block, what the reference does in seven launches
// (scheme/GodunovImplemV0.cpp:83-151 plus the border fill and the next dt reduction):
//   1. gather the 20x20 tile (ghost width 2) from the neighbour blocks, resolving the outside
//      border conditions on the fly, and convert it to primitives into LDS
//      (ConvertToPrimitivesVariablesFunctor + FillOutsideCellFunctor);
//   2. limited slopes, sound speed and Hancock half-step terms per cell, then the reconstructed
//      face states (ComputeLimitedSlopesFunctor + reconstruct_state_2d, MUSCL or THINC);
//   3. HLLC on the 17x16 x-faces and 16x17 y-faces (riemann_hllc);
//   4. flux exchange through LDS and the x then y conservative update with clamps,
//      volume-fraction normalisation (ReadFluxesAndConservativeUpdateFunctor,
//      FixVolumeFractionsNormalization);
//   5. the CFL contribution of the new state, reduced over the wavefront with shuffles and over
//      the workgroup in LDS, then one atomicMax per block (ComputeDtFiveEqFunctor).
// Thread t owns interior cell (t%16, t/16); threads 0..63 (wave 0) also own the 64 ring cells
// whose single face state the block needs, threads 0..143 convert the 144 halo cells.
#pragma once

#include "g5_device.h"
#include "g5_kernels.h"

namespace g5
{

namespace
{

constexpr int B2 = 16;
constexpr int T2 = B2 + 4;   // tile edge
constexpr int NT2 = T2 * T2; // 400 tile cells
constexpr int NF2 = 17 * 16; // faces per direction

G5D int
tidx2(int gi, int gj)
{
  return (gi + 2) + T2 * (gj + 2);
}
// x-face fi in [0,16] of row j ; y-face fj in [0,16] of column i
G5D int
fx2(int fi, int j)
{
  return fi + 17 * j;
}
G5D int
fy2(int i, int fj)
{
  return i + 16 * fj;
}

// halo cell h in [0,144): 64 predictor-ring cells then 80 outer cells
G5D void
halo_cell2(int h, int & gi, int & gj)
{
  if (h < 64)
  {
    const int r = h & 15, s = h >> 4;
    gi = s == 0 ? -1 : (s == 1 ? 16 : r);
    gj = s == 0 ? r : (s == 1 ? r : (s == 2 ? -1 : 16));
    return;
  }
  int c = h - 64;
  if (c < 4)
  {
    gi = (c & 1) ? 16 : -1;
    gj = (c & 2) ? 16 : -1;
    return;
  }
  c -= 4;
  if (c < 40)
  {
    gi = (c % 20) - 2;
    gj = c < 20 ? -2 : 17;
    return;
  }
  c -= 40; // 36 cells: columns -2 and 17, rows -1..16
  gi = c < 18 ? -2 : 17;
  gj = (c % 18) - 1;
}

template <typename T>
G5D PrimT<T>
sq_load2(T const * sQ, int idx)
{
  PrimT<T> p;
  p.r0 = sQ[0 * NT2 + idx];
  p.r1 = sQ[1 * NT2 + idx];
  p.a = sQ[2 * NT2 + idx];
  p.p = sQ[3 * NT2 + idx];
  p.u = sQ[4 * NT2 + idx];
  p.v = sQ[5 * NT2 + idx];
  p.w = T(0.0);
  return p;
}

template <typename T>
G5D void
sq_store2(T * sQ, int idx, PrimT<T> const & p)
{
  sQ[0 * NT2 + idx] = p.r0;
  sQ[1 * NT2 + idx] = p.r1;
  sQ[2 * NT2 + idx] = p.a;
  sQ[3 * NT2 + idx] = p.p;
  sQ[4 * NT2 + idx] = p.u;
  sQ[5 * NT2 + idx] = p.v;
}

// face states: 6 values (r0, r1, a, p, u, v) at slot k of a [6][NF2] array
template <typename T>
G5D void
st_store2(T * A, int k, FaceT<T> const & f)
{
  A[0 * NF2 + k] = f.r0;
  A[1 * NF2 + k] = f.r1;
  A[2 * NF2 + k] = f.a;
  A[3 * NF2 + k] = f.p;
  A[4 * NF2 + k] = f.u;
  A[5 * NF2 + k] = f.v;
}
template <typename T>
G5D FaceT<T>
st_load2(T const * A, int k)
{
  FaceT<T> f;
  f.r0 = A[0 * NF2 + k];
  f.r1 = A[1 * NF2 + k];
  f.a = A[2 * NF2 + k];
  f.p = A[3 * NF2 + k];
  f.u = A[4 * NF2 + k];
  f.v = A[5 * NF2 + k];
  f.w = T(0.0);
  return f;
}
// face records at slot k of a [7][NF2] array plus E[k]: the FP64 flux (ar0, ar1, a0, a1, mu, mv,
// us; energy in E) or the mixed path's star state (ro0, ro1, ro, uo, pt, t1, phi0; etot in E)
G5D void
fl_store2(double * A, double * E, int k, Flux const & F)
{
  A[0 * NF2 + k] = F.ar0;
  A[1 * NF2 + k] = F.ar1;
  A[2 * NF2 + k] = F.a0;
  A[3 * NF2 + k] = F.a1;
  A[4 * NF2 + k] = F.mu;
  A[5 * NF2 + k] = F.mv;
  A[6 * NF2 + k] = F.us;
  E[k] = F.e;
}
G5D Flux
fl_load2(double const * A, double const * E, int k)
{
  Flux F;
  F.ar0 = A[0 * NF2 + k];
  F.ar1 = A[1 * NF2 + k];
  F.a0 = A[2 * NF2 + k];
  F.a1 = A[3 * NF2 + k];
  F.mu = A[4 * NF2 + k];
  F.mv = A[5 * NF2 + k];
  F.mw = 0.0;
  F.us = A[6 * NF2 + k];
  F.e = E[k];
  return F;
}
G5D void
fl_store2(float * A, double * E, int k, FluxT<float> const & F)
{
  A[0 * NF2 + k] = F.ro0;
  A[1 * NF2 + k] = F.ro1;
  A[2 * NF2 + k] = F.ro;
  A[3 * NF2 + k] = F.uo;
  A[4 * NF2 + k] = F.pt;
  A[5 * NF2 + k] = F.t1;
  A[6 * NF2 + k] = F.phi0;
  E[k] = F.etot;
}
G5D FluxT<float>
fl_load2(float const * A, double const * E, int k)
{
  FluxT<float> F;
  F.ro0 = A[0 * NF2 + k];
  F.ro1 = A[1 * NF2 + k];
  F.ro = A[2 * NF2 + k];
  F.uo = A[3 * NF2 + k];
  F.pt = A[4 * NF2 + k];
  F.t1 = A[5 * NF2 + k];
  F.t2 = 0.0f;
  F.phi0 = A[6 * NF2 + k];
  F.etot = E[k];
  return F;
}

// the reconstructed face state of tile cell (gi,gj) on axis ax with side ib (+1 right, -1 left)
template <bool THINC, bool HANCOCK, typename T, bool FAST>
G5D void
predict_faces2(SolveDesc const & S, T const * sQ, int gi, int gj, T dtdx, bool all, int ax1, int ib1,
               FaceT<T> & fxm, FaceT<T> & fxp, FaceT<T> & fym, FaceT<T> & fyp, FaceT<T> & f1)
{
  const PrimT<T>  qc = sq_load2(sQ, tidx2(gi, gj));
  const PrimT<T>  qxm = sq_load2(sQ, tidx2(gi - 1, gj));
  const PrimT<T>  qxp = sq_load2(sQ, tidx2(gi + 1, gj));
  const PrimT<T>  qym = sq_load2(sQ, tidx2(gi, gj - 1));
  const PrimT<T>  qyp = sq_load2(sQ, tidx2(gi, gj + 1));
  const T         st = static_cast<T>(S.slope_type);
  // the limiter tested per axis here: tested once for both axes, the mixed 2-D step needs 83
  // registers instead of 79, and one wavefront per SIMD less
  const PrimT<T>  qpx[1] = { qxp }, qmx[1] = { qxm }, qpy[1] = { qyp }, qmy[1] = { qym };
  SlopeT<T>       slx[1], sly[1];
  slopes_axes<T, FAST, 1>(qc, qpx, qmx, st, slx);
  slopes_axes<T, FAST, 1>(qc, qpy, qmy, st, sly);
  const SlopeT<T> sx = slx[0];
  const SlopeT<T> sy = sly[0];
  const SlopeT<T> sz = { T(0.0), T(0.0), T(0.0), T(0.0), T(0.0), T(0.0) };
  HancockT<T>     h = { T(0.0), T(0.0), T(0.0), T(0.0), T(0.0), T(0.0) };
  if (HANCOCK)
    h = hancock_terms<2, T, FAST>(S, qc, sx, sy, sz, dtdx);
  if (!THINC)
  {
    if (all)
    {
      fxm = face_muscl<2, 0, HANCOCK, T, FAST>(S, qc, sx, sy, sz, h, T(-0.5));
      fxp = face_muscl<2, 0, HANCOCK, T, FAST>(S, qc, sx, sy, sz, h, T(0.5));
      fym = face_muscl<2, 1, HANCOCK, T, FAST>(S, qc, sx, sy, sz, h, T(-0.5));
      fyp = face_muscl<2, 1, HANCOCK, T, FAST>(S, qc, sx, sy, sz, h, T(0.5));
    }
    else
    {
      // one reconstruction on the lane's axis ax1
      f1 = face_muscl<2, -1, HANCOCK, T, FAST>(S, qc, sx, sy, sz, h, T(0.5) * static_cast<T>(ib1), ax1);
    }
    return;
  }
  // THINC: Youngs normal from the 3x3 alpha stencil, needed only in mixed cells
  T nx = T(0.0), ny = T(0.0);
  if (qc.a > Par<T>::eps(S) and qc.a < T(1.0) - Par<T>::eps(S))
  {
    T a9[9];
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int di = -1; di <= 1; ++di)
        a9[(di + 1) + 3 * (dj + 1)] = sQ[2 * NT2 + tidx2(gi + di, gj + dj)];
    youngs2<T, FAST>(a9, nx, ny);
  }
  if (all)
  {
    const ThincAxisT<T> tx = thinc_axis<T, FAST>(S, qc.a, qxm.a, qc.a, qxp.a, fabs(nx));
    const ThincAxisT<T> ty = thinc_axis<T, FAST>(S, qc.a, qym.a, qc.a, qyp.a, fabs(ny));
    fxm = face_thinc<2, 0, HANCOCK, T, FAST>(qc, sx, sy, sz, h, tx, -1);
    fxp = face_thinc<2, 0, HANCOCK, T, FAST>(qc, sx, sy, sz, h, tx, 1);
    fym = face_thinc<2, 1, HANCOCK, T, FAST>(qc, sx, sy, sz, h, ty, -1);
    fyp = face_thinc<2, 1, HANCOCK, T, FAST>(qc, sx, sy, sz, h, ty, 1);
  }
  else
  {
    const bool          x = ax1 == 0;
    const ThincAxisT<T> t1 =
      thinc_axis<T, FAST>(S, qc.a, x ? qxm.a : qym.a, qc.a, x ? qxp.a : qyp.a, fabs(x ? nx : ny));
    f1 = face_thinc<2, -1, HANCOCK, T, FAST>(qc, sx, sy, sz, h, t1, ib1, ax1);
  }
}

// ---- two-level static mesh ----------------------------------------------------------------
// Level 0 is the brick of nb[0] x nb[1] blocks, some of them refined into 2 x 2 blocks of half
// the cell length (level 1), at least one block away from the brick's border.  The solve's slot
// tables: level 0 at S.base (-1 where refined), level 1 at S.base1 over the 2nb[0] x 2nb[1] fine
// grid (-1 outside the refined blocks).  A cell is named by its global index at its level.
G5D int
slot_l0(SolveDesc const & S, int32_t const * __restrict__ slots, int x, int y)
{
  return slots[S.base + x + S.nb[0] * y];
}
G5D int
slot_l1(SolveDesc const & S, int32_t const * __restrict__ slots, int x, int y)
{
  return slots[S.base1 + x + 2 * S.nb[0] * y];
}
G5D Cons
cell_of(double const * __restrict__ U, int slot, int ci, int cj)
{
  return load_cons<2, B2>(U, static_cast<int64_t>(slot) * 8 * 256 + ci + B2 * cj, -1);
}

// coarse cell (gx, gy) inside a refined block: the average of its four fine cells, summed x
// fastest then divided by 4, for E, rho u, rho v, alpha_rho 0/1, alpha 0; alpha 1 = 1 - alpha 0
// and rho = alpha_rho 0 + alpha_rho 1 (get_conservative_vars_restriction,
// scheme/ConvertToPrimitivesVariablesFunctor.cpp:214-250, with the siblings' average of the core)
G5D Cons
restrict_cell(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U, int gx, int gy)
{
  Cons sum{};
#pragma unroll
  for (int k = 0; k < 4; ++k)
  {
    const int  fx = 2 * gx + (k & 1), fy = 2 * gy + (k >> 1);
    const Cons f = cell_of(U, slot_l1(S, slots, fx >> 4, fy >> 4), fx & 15, fy & 15);
    sum.e = k == 0 ? f.e : sum.e + f.e;
    sum.mu = k == 0 ? f.mu : sum.mu + f.mu;
    sum.mv = k == 0 ? f.mv : sum.mv + f.mv;
    sum.ar0 = k == 0 ? f.ar0 : sum.ar0 + f.ar0;
    sum.ar1 = k == 0 ? f.ar1 : sum.ar1 + f.ar1;
    sum.a0 = k == 0 ? f.a0 : sum.a0 + f.a0;
  }
  Cons u;
  u.e = sum.e / 4.0;
  u.mu = sum.mu / 4.0;
  u.mv = sum.mv / 4.0;
  u.mw = 0.0;
  u.ar0 = sum.ar0 / 4.0;
  u.ar1 = sum.ar1 / 4.0;
  u.a0 = sum.a0 / 4.0;
  u.a1 = 1.0 - u.a0;
  u.d = u.ar0 + u.ar1;
  return u;
}

// coarse cell (gx, gy) of the parent's 3x3-block neighbourhood (all inside the brick): a coarse
// leaf's cell, or the restriction where the block is refined
G5D Cons
coarse_cell(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U, int gx, int gy)
{
  const int sl = slot_l0(S, slots, gx >> 4, gy >> 4);
  return sl >= 0 ? cell_of(U, sl, gx & 15, gy & 15) : restrict_cell(S, slots, U, gx, gy);
}

// variable at offset var of the 8-variable layout of level-0 cell (gx, gy): a coarse leaf's value,
// or the average of the four fine cells where the block is refined (summed x fastest, over 4)
G5D double
coarse_var(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U, int gx, int gy,
           int var)
{
  const int sl = slot_l0(S, slots, gx >> 4, gy >> 4);
  if (sl >= 0)
    return U[static_cast<int64_t>(sl) * 8 * 256 + var * 256 + (gx & 15) + B2 * (gy & 15)];
  double sum = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
  {
    const int    fx = 2 * gx + (k & 1), fy = 2 * gy + (k >> 1);
    const int    fs = slot_l1(S, slots, fx >> 4, fy >> 4);
    const double v = U[static_cast<int64_t>(fs) * 8 * 256 + var * 256 + (fx & 15) + B2 * (fy & 15)];
    sum = k == 0 ? v : sum + v;
  }
  return sum / 4.0;
}

// CONSERVATIVE_INTERPOLATION_ORDER_2 / _4 (S.prolong 2 / 3; conservative_interpolation_order{2,4}
// and interpol_order{2,4}, scheme/ConvertToPrimitivesVariablesFunctor.cpp:339-490,751-827): the
// fine cell's half of its parent (cx, cy) along x and y, by 1-D conservative interpolations over
// the parent's 3x3 (5x5) coarse neighbourhood, x first, then y.  Weights: the averages over the
// left half of the polynomial of degree 2 (4) with the given cell averages, (1/8, 1, -1/8) and
// (-3/128, 22/128, 1, -22/128, 3/128), mirrored for the right half; sums in stencil order.
G5D double
conservative_var(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U, int cx,
                 int cy, int px, int py, int var)
{
  const double w2[3] = { 0.125, 1.0, -0.125 };
  const double w4[5] = { -3.0 / 128.0, 22.0 / 128.0, 1.0, -22.0 / 128.0, 3.0 / 128.0 };
  const bool   o2 = S.prolong == 2;
  const int    m = o2 ? 1 : 2, n = 2 * m + 1;
  double       valy = 0.0;
  for (int ky = 0; ky < n; ++ky)
  {
    double valx = 0.0;
    for (int kx = 0; kx < n; ++kx)
    {
      const int    ix = px == 0 ? kx : n - 1 - kx;
      const double x = (o2 ? w2[ix] : w4[ix]) * coarse_var(S, slots, U, cx + kx - m, cy + ky - m, var);
      valx = kx == 0 ? x : valx + x;
    }
    const int    iy = py == 0 ? ky : n - 1 - ky;
    const double wy = o2 ? w2[iy] : w4[iy];
    valy = ky == 0 ? wy * valx : valy + wy * valx;
  }
  return valy;
}

// fine cell (gx, gy) outside the fine blocks: its coarse parent copied (S.prolong 0, SIMPLE_COPY),
// extrapolated to the fine cell's side with minmod slopes of the conservative variables
// (S.prolong 1, EXTRAPOLATE_LINEAR_MINMOD: linear_extrapolate_using_limited_slopes,
// scheme/ConvertToPrimitivesVariablesFunctor.cpp:553-770, fill_ghosts :889-910), or
// conservatively interpolated (S.prolong 2, 3)
G5D Cons
prolong_cell(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U, int gx, int gy)
{
  const int cx = gx >> 1, cy = gy >> 1; // fine cells of the ghost ring are at positive indices
  if (S.prolong >= 2)
  {
    const int px = gx & 1, py = gy & 1;
    Cons      u;
    u.ar0 = conservative_var(S, slots, U, cx, cy, px, py, 5);
    u.ar1 = conservative_var(S, slots, U, cx, cy, px, py, 7);
    u.a0 = conservative_var(S, slots, U, cx, cy, px, py, 4);
    u.e = conservative_var(S, slots, U, cx, cy, px, py, 1);
    u.mu = conservative_var(S, slots, U, cx, cy, px, py, 2);
    u.mv = conservative_var(S, slots, U, cx, cy, px, py, 3);
    u.mw = 0.0;
    u.d = u.ar0 + u.ar1;
    u.a1 = 1.0 - u.a0;
    return u;
  }
  const Cons c = coarse_cell(S, slots, U, cx, cy);
  if (S.prolong == 0)
    return c;
  const double hx = (gx & 1) ? 0.25 : -0.25, hy = (gy & 1) ? 0.25 : -0.25;
  const Cons   xl = coarse_cell(S, slots, U, cx - 1, cy), xr = coarse_cell(S, slots, U, cx + 1, cy);
  const Cons   yl = coarse_cell(S, slots, U, cx, cy - 1), yr = coarse_cell(S, slots, U, cx, cy + 1);
  Cons         u;
  u.ar0 = c.ar0 + hx * slope1_limited<double>(c.ar0, xr.ar0, xl.ar0, 1.0) + hy * slope1_limited<double>(c.ar0, yr.ar0, yl.ar0, 1.0);
  u.ar1 = c.ar1 + hx * slope1_limited<double>(c.ar1, xr.ar1, xl.ar1, 1.0) + hy * slope1_limited<double>(c.ar1, yr.ar1, yl.ar1, 1.0);
  u.a0 = c.a0 + hx * slope1_limited<double>(c.a0, xr.a0, xl.a0, 1.0) + hy * slope1_limited<double>(c.a0, yr.a0, yl.a0, 1.0);
  u.e = c.e + hx * slope1_limited<double>(c.e, xr.e, xl.e, 1.0) + hy * slope1_limited<double>(c.e, yr.e, yl.e, 1.0);
  u.mu = c.mu + hx * slope1_limited<double>(c.mu, xr.mu, xl.mu, 1.0) + hy * slope1_limited<double>(c.mu, yr.mu, yl.mu, 1.0);
  u.mv = c.mv + hx * slope1_limited<double>(c.mv, xr.mv, xl.mv, 1.0) + hy * slope1_limited<double>(c.mv, yr.mv, yl.mv, 1.0);
  u.mw = 0.0;
  u.d = u.ar0 + u.ar1;
  u.a1 = 1.0 - u.a0;
  return u;
}

// conservative state of ghost cell (gi, gj) of a block of a two-level solve (fill_ghosts,
// scheme/ConvertToPrimitivesVariablesFunctor.cpp:884-918): a same-level neighbour's cell (border
// conditions resolved as on one level), a restriction of a finer neighbour, or a prolongation of a
// coarser one
G5D Cons
ghost_two_level(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U,
                BlockDesc const & bd, int gi, int gj)
{
  const int dx = gi < 0 ? -1 : (gi >= B2 ? 1 : 0), dy = gj < 0 ? -1 : (gj >= B2 ? 1 : 0);
  const int nx = bd.bx + dx, ny = bd.by + dy;
  const int gx = bd.bx * B2 + gi, gy = bd.by * B2 + gj;
  if (bd.level == 0)
  {
    if (nx >= 0 and nx < S.nb[0] and ny >= 0 and ny < S.nb[1] and slot_l0(S, slots, nx, ny) < 0)
      return restrict_cell(S, slots, U, gx, gy);
    int           neg;
    const int64_t off = resolve_cell<2, B2>(S, slots, bd.bx, bd.by, 0, gi, gj, 0, neg);
    Cons          u = load_cons<2, B2>(U, off, -1);
    apply_border<2>(S, u, neg);
    return u;
  }
  const int sl = slot_l1(S, slots, nx, ny); // fine blocks sit inside the fine grid's interior
  return sl >= 0 ? cell_of(U, sl, gi - dx * B2, gj - dy * B2) : prolong_cell(S, slots, U, gx, gy);
}

// Face records across a level interface, per coarse slot: [side xmin, xmax, ymin, ymax][cell
// 0..15][ar0, ar1, a0, a1, mu, mv, us, e].  Entry (side, j) of coarse block C holds, for C's
// boundary cell j on that side, the sum of the records of the two fine faces that cover the
// coarse face divided by 4 (the volume ratio): get_flux_from_fine_neighbor /
// get_ustar_flux_from_fine_neighbor (scheme/ReadFluxesAndConservativeUpdateFunctor.h:256-307) and
// the NEIGHBOR_IS_FINER branches of read_fluxes_and_update_2d (.cpp:137-199).  The records hold
// flux * dt / dx_fine; u* enters the update as st * (u*_sum / 4) * alpha, the same bits as the
// reference's st * u*_sum * alpha / 4 (scalings by 4 are exact).
G5D int64_t
iflux_at(int slot, int side, int cell)
{
  static_assert(kIfluxDoubles == 4 * 16 * 8, "interface record layout");
  return ((static_cast<int64_t>(slot) * 4 + side) * 16 + cell) * 8;
}
G5D Flux
iflux_load(double const * __restrict__ R, int64_t o)
{
  Flux F;
  F.ar0 = R[o + 0];
  F.ar1 = R[o + 1];
  F.a0 = R[o + 2];
  F.a1 = R[o + 3];
  F.mu = R[o + 4];
  F.mv = R[o + 5];
  F.mw = 0.0;
  F.us = R[o + 6];
  F.e = R[o + 7];
  return F;
}
G5D void
iflux_store_sum(double * __restrict__ R, int64_t o, Flux const & a, Flux const & b)
{
  R[o + 0] = (a.ar0 + b.ar0) / 4.0;
  R[o + 1] = (a.ar1 + b.ar1) / 4.0;
  R[o + 2] = (a.a0 + b.a0) / 4.0;
  R[o + 3] = (a.a1 + b.a1) / 4.0;
  R[o + 4] = (a.mu + b.mu) / 4.0;
  R[o + 5] = (a.mv + b.mv) / 4.0;
  R[o + 6] = (a.us + b.us) / 4.0;
  R[o + 7] = (a.e + b.e) / 4.0;
}

} // namespace

// One LDS arena, in doubles.  Phases 1-3 use [primitives 6 x 400 | face states 12 x 272] (T);
// phase 4 reuses it from the start for the face records (14 x 272 T + 2 x 272 double), once the
// barrier after the Riemann solves has retired every read of the primitives and face states.
// FP64: 45.3 KB, so three workgroups fit in a CU's 160 KB.
template <typename T>
constexpr int ARENA2_BYTES_A = (6 * NT2 + 12 * NF2) * static_cast<int>(sizeof(T));
template <typename T>
constexpr int ARENA2_BYTES_B = 14 * NF2 * static_cast<int>(sizeof(T)) + 2 * NF2 * 8;
template <typename T>
constexpr int ARENA2_DOUBLES =
  ((ARENA2_BYTES_A<T> > ARENA2_BYTES_B<T> ? ARENA2_BYTES_A<T> : ARENA2_BYTES_B<T>)+7) / 8;

// RW: a fifth wavefront (threads 256..319) does the 64 ring-cell face states and the 32 ring
// Riemann problems in parallel with the owned cells instead of after them on wave 0 -- the form
// for launch lists too short to fill the GPU, where a block's latency is the step time.
// AMR: the launch list may hold blocks of two-level solves (FP64 records only): their ghosts come
// from ghost_two_level, a coarse block takes the faces it shares with fine blocks from A.iflux,
// and a fine block bordering coarse blocks writes those records there (the fine blocks that do
// run in a launch of their own before the rest of the list).
template <bool THINC, bool HANCOCK, typename T, bool RW = false, bool MP = false, bool FAST = false, bool AMR = false>
G5D void
step2d_block(StepArgs const & A, int blk)
{
  static_assert(!AMR or sizeof(T) == sizeof(double), "two-level solves exchange FP64 face records");
  __shared__ double sArena[ARENA2_DOUBLES<T>];
  __shared__ double sRed[5];
  T * const         sQ = reinterpret_cast<T *>(sArena);

  const int       t = threadIdx.x;
  const bool      owner = !RW or t < 256;       // lane with an owned cell
  const int       tr = RW ? t - 256 : t;        // ring lane index when in [0, 64)
  const bool      ring = RW ? !owner : t < 64; // lane with a ring cell
  const BlockDesc bd = A.blocks[blk];
  const SolveDesc S = AMR ? at_level(A.solves[bd.solve], bd.level) : A.solves[bd.solve];
  SolveCtl *      C = A.ctl + bd.solve;
  const int64_t   own = static_cast<int64_t>(bd.slot) * 8 * 256 + t;
  const bool      host_dt = A.dt_fixed > 0.0;
  if (!host_dt and C->active == 0)
    return; // finished solve: its state stays in buffer `parity`
  // two-level solve: sides (bit s: xmin, xmax, ymin, ymax) whose neighbour is finer (a coarse
  // block) or coarser (a fine block)
  const bool two_level = AMR and S.base1 >= 0;
  int        finer = 0, coarser = 0;
  if (two_level)
  {
#pragma unroll
    for (int sd = 0; sd < 4; ++sd)
    {
      const int nx = bd.bx + (sd == 0 ? -1 : (sd == 1 ? 1 : 0)), ny = bd.by + (sd == 2 ? -1 : (sd == 3 ? 1 : 0));
      if (bd.level == 0)
        finer |= (nx >= 0 and nx < S.nb[0] and ny >= 0 and ny < S.nb[1] and slot_l0(S, A.slots, nx, ny) < 0) ? 1 << sd
                                                                                                             : 0;
      else
        coarser |= slot_l1(S, A.slots, nx, ny) < 0 ? 1 << sd : 0;
    }
  }
  const int      par = C->parity;
  double const * U = A.U_in ? A.U_in : (par ? A.buf1 : A.buf0);
  double *       U2 = A.U_out ? A.U_out : (par ? A.buf0 : A.buf1);
  const double   dt = host_dt ? A.dt_fixed : C->dt;
  const double   dtdx = dt / S.dx;
  const T        dtdxT = static_cast<T>(dtdx);
  const int      i = t & 15, j = t >> 4;

  // ---- 1. gather + primitives ----
  Cons uown{};
  if (owner)
    uown = load_cons<2, B2>(U, own, -1);
  if (owner)
  {
    PrimT<T> q;
    to_prim<2, FAST>(S, uown, q);
    sq_store2(sQ, tidx2(i, j), q);
  }
  if (t < 144)
  {
    int  gi, gj, neg;
    Cons u;
    halo_cell2(t, gi, gj);
    if (two_level)
      u = ghost_two_level(S, A.slots, U, bd, gi, gj);
    else
    {
      const int64_t off = resolve_cell<2, B2>(S, A.slots, bd.bx, bd.by, 0, gi, gj, 0, neg);
      u = load_cons<2, B2>(U, off, -1);
      apply_border<2>(S, u, neg);
    }
    PrimT<T>      q;
    to_prim<2, FAST>(S, u, q);
    sq_store2(sQ, tidx2(gi, gj), q);
  }
  __syncthreads();

  // ---- 2. slopes, Hancock terms, face states ----
  T *      XL = sQ + 6 * NT2; // [6][NF2] left states of the x faces
  T *      YL = XL + 6 * NF2;  // [6][NF2] left states of the y faces
  FaceT<T> fxm, fxp, fym, fyp, fdummy, fring;
  if (owner)
  {
    predict_faces2<THINC, HANCOCK, T, FAST>(S, sQ, i, j, dtdxT, true, 0, 0, fxm, fxp, fym, fyp, fdummy);
    st_store2(XL, fx2(i + 1, j), fxp);
    st_store2(YL, fy2(i, j + 1), fyp);
  }
  // ring cells: s=0 (-1,r) +x ; s=1 (16,r) -x ; s=2 (r,-1) +y ; s=3 (r,16) -y
  const int rs = (tr >> 4) & 3, rr = tr & 15;
  if (ring)
  {
    int gi, gj;
    halo_cell2(tr, gi, gj);
    const int ax = rs < 2 ? 0 : 1;
    const int ib = (rs & 1) ? -1 : 1;
    predict_faces2<THINC, HANCOCK, T, FAST>(S, sQ, gi, gj, dtdxT, false, ax, ib, fdummy, fdummy, fdummy, fdummy,
                                            fring);
    if (rs == 0)
      st_store2(XL, fx2(0, rr), fring);
    else if (rs == 2)
      st_store2(YL, fy2(rr, 0), fring);
  }
  __syncthreads();

  // ---- 3. HLLC ----
  FluxT<T> Fx, Fy;
  if (owner)
  {
    Fx = riemann<2, FAST>(S, st_load2(XL, fx2(i, j)), fxm, dtdx, 0);
    Fy = riemann<2, FAST>(S, st_load2(YL, fy2(i, j)), fym, dtdx, 1);
  }
  FluxT<T>   Fr;
  const bool ring_face = ring and (rs & 1);
  if (ring_face)
  {
    const int      ax = rs == 1 ? 0 : 1;
    const FaceT<T> L = ax == 0 ? st_load2(XL, fx2(16, rr)) : st_load2(YL, fy2(rr, 16));
    Fr = riemann<2, FAST>(S, L, fring, dtdx, ax);
  }
  if constexpr (AMR)
  {
    // a coarse block's faces on a finer neighbour take the fine faces' records
    if (finer != 0)
    {
      if (owner and i == 0 and (finer & 1))
        Fx = iflux_load(A.iflux, iflux_at(bd.slot, 0, j));
      if (owner and j == 0 and (finer & 4))
        Fy = iflux_load(A.iflux, iflux_at(bd.slot, 2, i));
      const int sd = rs == 1 ? 1 : 3;
      if (ring_face and ((finer >> sd) & 1))
        Fr = iflux_load(A.iflux, iflux_at(bd.slot, sd, rr));
    }
  }
  __syncthreads();

  // ---- 4. flux exchange and update ----
  T *      FX = reinterpret_cast<T *>(sArena); // [7][NF2] (primitives and face states are dead)
  T *      FY = FX + 7 * NF2;              // [7][NF2]
  double * FXE = reinterpret_cast<double *>(FY + 7 * NF2);
  double * FYE = FXE + NF2;
  if (owner)
  {
    fl_store2(FX, FXE, fx2(i, j), Fx);
    fl_store2(FY, FYE, fy2(i, j), Fy);
  }
  if (ring_face)
  {
    if (rs == 1)
      fl_store2(FX, FXE, fx2(16, rr), Fr);
    else
      fl_store2(FY, FYE, fy2(rr, 16), Fr);
  }
  __syncthreads();
  if constexpr (AMR)
  {
    // a fine block on a coarse neighbour: per coarse boundary cell, the two fine faces' records
    // summed and divided by 4 into the coarse block's entry (lane: side t/8, coarse cell t%8 of
    // the 8 this block covers)
    const int sd = t >> 3, c = t & 7;
    if (coarser != 0 and t < 32 and ((coarser >> sd) & 1))
    {
      Flux r0, r1;
      if (sd < 2)
      {
        const int fi = sd == 0 ? 0 : 16;
        r0 = fl_load2(FX, FXE, fx2(fi, 2 * c));
        r1 = fl_load2(FX, FXE, fx2(fi, 2 * c + 1));
      }
      else
      {
        const int fj = sd == 2 ? 0 : 16;
        r0 = fl_load2(FY, FYE, fy2(2 * c, fj));
        r1 = fl_load2(FY, FYE, fy2(2 * c + 1, fj));
      }
      const int nx = bd.bx + (sd == 0 ? -1 : (sd == 1 ? 1 : 0)), ny = bd.by + (sd == 2 ? -1 : (sd == 3 ? 1 : 0));
      const int cslot = slot_l0(S, A.slots, nx >> 1, ny >> 1);
      const int cell = 8 * ((sd < 2 ? bd.by : bd.bx) & 1) + c;
      iflux_store_sum(A.iflux, iflux_at(cslot, sd ^ 1, cell), r0, r1);
    }
  }
  double   rate = 0.0;
  uint32_t fail = 0u;
  if (owner)
  {
    Cons   u2 = uown;
    double st0 = 1.0, st1 = 1.0; // volume-fraction source factors (MILLER_PUCKETT: from the time-n cell)
    if constexpr (MP) // the list holds a MILLER_PUCKETT solve; this block's solve decides
      if (S.source_term == G5_SOURCE_MILLER_PUCKETT)
        source_factors<2>(S, uown, st0, st1);
    apply_face_pair<2, FAST>(u2, Fx, fl_load2(FX, FXE, fx2(i + 1, j)), uown.a0, uown.a1, dtdx, 0, st0, st1);
    apply_face_pair<2, FAST>(u2, Fy, fl_load2(FY, FYE, fy2(i, j + 1)), uown.a0, uown.a1, dtdx, 1, st0, st1);
    fix_volume_fractions<FAST>(u2);
    store_cons<2, B2>(U2, own, u2);
    double pc;
    rate = cell_rate<2, T, FAST>(S, u2, pc);
    fail = cell_failure(S, u2, rate, pc);
  }
  report_failure(C, fail);

  // ---- 5. CFL of the new state (the ring wavefront has no cell) ----
  double inv = wave_max(rate);
  if ((t & 63) == 0)
    sRed[t >> 6] = inv;
  __syncthreads();
  if (t == 0)
  {
    inv = fmax(fmax(sRed[0], sRed[1]), fmax(sRed[2], sRed[3]));
    atomicMax(&C->invdt_bits, static_cast<unsigned long long>(__double_as_longlong(inv)));
  }
}

// one workgroup per block
template <bool THINC, bool HANCOCK, typename T, bool MP, bool FAST>
__global__ void __launch_bounds__(256)
k_step2d(StepArgs A)
{
  int b, end, stride;
  block_range(A, b, end, stride);
  if (b < end)
    step2d_block<THINC, HANCOCK, T, false, MP, FAST>(A, b);
}

// one workgroup per block, blocks of two-level solves allowed
template <bool THINC, bool HANCOCK, bool MP, bool FAST>
__global__ void __launch_bounds__(256)
k_step2d_two_level(StepArgs A)
{
  int b, end, stride;
  block_range(A, b, end, stride);
  if (b < end)
    step2d_block<THINC, HANCOCK, double, false, MP, FAST, true>(A, b);
}

// one workgroup of five wavefronts per block (ring work on the fifth)
template <bool THINC, bool HANCOCK, typename T, bool MP, bool FAST>
__global__ void __launch_bounds__(320)
k_step2d_ringwave(StepArgs A)
{
  int b, end, stride;
  block_range(A, b, end, stride);
  if (b < end)
    step2d_block<THINC, HANCOCK, T, true, MP, FAST>(A, b);
}

// persistent workgroups looping over their range of the block list
template <bool THINC, bool HANCOCK, typename T, bool MP, bool FAST>
__global__ void __launch_bounds__(256)
k_step2d_persistent(StepArgs A)
{
  int b, end, stride;
  block_range(A, b, end, stride);
  for (; b < end; b += stride)
  {
    step2d_block<THINC, HANCOCK, T, false, MP, FAST>(A, b);
    __syncthreads(); // the next block reuses the LDS
  }
}

// MP: the launch list holds a solve with the MILLER_PUCKETT source term
template <bool THINC, bool HANCOCK, typename T, bool MP, bool FAST>
static void
launch2k(StepArgs const & A, bool ring_wave, hipStream_t stream)
{
  const dim3 block(256);
  if (A.wg_per_xcd > 0)
    hipLaunchKernelGGL((k_step2d_persistent<THINC, HANCOCK, T, MP, FAST>), dim3(8 * A.wg_per_xcd), block, 0, stream,
                       A);
  else if (ring_wave)
    hipLaunchKernelGGL((k_step2d_ringwave<THINC, HANCOCK, T, MP, FAST>), dim3(8 * A.per_xcd), dim3(320), 0, stream,
                       A);
  else
    hipLaunchKernelGGL((k_step2d<THINC, HANCOCK, T, MP, FAST>), dim3(8 * A.per_xcd), block, 0, stream, A);
}

// a launch list with blocks of two-level solves: the fine blocks that border coarse blocks
// (the first n_first of the list) write their interface records, then the rest of the list reads them
template <bool MP, bool FAST>
static void
launch2_two_level(StepArgs const & A, int n_first, bool thinc, bool hancock, hipStream_t stream)
{
  StepArgs P[2] = { A, A };
  P[0].nblocks = n_first;
  P[1].blocks = A.blocks + n_first;
  P[1].nblocks = A.nblocks - n_first;
  for (StepArgs & a : P)
  {
    if (a.nblocks <= 0)
      continue;
    a.per_xcd = (a.nblocks + 7) / 8;
    a.wg_per_xcd = 0;
    const dim3 grid(8 * a.per_xcd), block(256);
    if (thinc and hancock)
      hipLaunchKernelGGL((k_step2d_two_level<true, true, MP, FAST>), grid, block, 0, stream, a);
    else if (thinc)
      hipLaunchKernelGGL((k_step2d_two_level<true, false, MP, FAST>), grid, block, 0, stream, a);
    else if (hancock)
      hipLaunchKernelGGL((k_step2d_two_level<false, true, MP, FAST>), grid, block, 0, stream, a);
    else
      hipLaunchKernelGGL((k_step2d_two_level<false, false, MP, FAST>), grid, block, 0, stream, a);
  }
}

template <typename T, bool MP, bool FAST>
static void
launch2(StepArgs const & A, bool thinc, bool hancock, bool ring_wave, hipStream_t stream)
{
  if (thinc and hancock)
    launch2k<true, true, T, MP, FAST>(A, ring_wave, stream);
  else if (thinc)
    launch2k<true, false, T, MP, FAST>(A, ring_wave, stream);
  else if (hancock)
    launch2k<false, true, T, MP, FAST>(A, ring_wave, stream);
  else
    launch2k<false, false, T, MP, FAST>(A, ring_wave, stream);
}

} // namespace g5
