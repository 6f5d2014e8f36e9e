// g5_step3d_body.h -- fused 3-D five-equation step, one 16^3 block per workgroup, marching in z.
//
// A 20^3 primitive tile with ghost width 2 does not fit in LDS, so the block is swept plane by
// plane (2.5-D): LDS holds a ring of three primitive planes (P-1, P, P+1; 20x20 each) for the x,
// y and z stencils and the THINC normals.  Per plane P = -1..16 the workgroup
//   a. converts plane P+1 (gathered one phase earlier, ghost/border resolution on the fly with
//      the neighbour block slots from a per-block table in LDS),
//   b. computes slopes, sound speed and Hancock terms of plane P and its face states,
//   c. solves HLLC on the x and y faces of plane P and on the z face between P-1 and P,
//   d. exchanges the x/y fluxes through LDS, applies the x then y update to plane P and the z
//      update (with normalisation, store and CFL contribution) to plane P-1.
// Same numerics as scheme/GodunovImplemV0.cpp:83-151 (see g5_step2d.hip for the per-function
// citations); the z face flux of plane P-1 is kept in registers until plane P is done.
//
// Layout: one workgroup of 4 wavefronts per block; thread t owns the interior column
// (t%16, t/16).  The register demand of the FP64 z-march allows one wavefront per SIMD, so the
// FP64 form uses up to 139 KB of LDS; the mixed form (T = float primitives, face states and
// fluxes, FP64 energy flux and state) halves it.  Extra work is spread so that every wavefront
// has one extra item per plane: wave 0 converts the 64 ring cells and solves the 32 ring Riemann
// problems, wave 1 computes the ring cells' face states, wave 2 converts 16 outer cells, wave 3
// converts 64 outer cells (lanes on different axes share one run-time-axis code path).  The raw
// state of the owned column is carried from its conversion to its update, and the global loads
// of plane P+2 are issued one phase ahead of their use.  The z Riemann solve runs before the x/y
// exchange (with ZSPLIT on waves 0, 2, 3 only: wave 1, which carries the ring predictor, solves
// after it); the completion of plane P-1 follows the last barrier of the plane on every wave.
#pragma once

#include "g5_device.h"
#include "g5_kernels.h"

namespace g5
{

namespace
{

constexpr int B3 = 16;
constexpr int T3 = 20;
constexpr int NT3 = T3 * T3; // 400
constexpr int NF3 = 17 * 16; // 272

G5D int
tidx3(int gi, int gj)
{
  return (gi + 2) + T3 * (gj + 2);
}
G5D int
fx3(int fi, int j)
{
  return fi + 17 * j;
}
G5D int
fy3(int i, int fj)
{
  return i + 16 * fj;
}

// outer (non-interior, non-ring) plane cell c in [0,80)
G5D void
outer_cell3(int c, int & gi, int & gj)
{
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
  c -= 40;
  gi = c < 18 ? -2 : 17;
  gj = (c % 18) - 1;
}

// ring cell r in [0,64): s=0 (-1,r) s=1 (16,r) s=2 (r,-1) s=3 (r,16)
G5D void
ring_cell3(int r, int & gi, int & gj)
{
  const int rr = r & 15, s = r >> 4;
  gi = s == 0 ? -1 : (s == 1 ? 16 : rr);
  gj = s == 0 ? rr : (s == 1 ? rr : (s == 2 ? -1 : 16));
}

template <typename T>
G5D PrimT<T>
sq_load3(T const * sQ, int idx)
{
  PrimT<T> p;
  p.r0 = sQ[0 * NT3 + idx];
  p.r1 = sQ[1 * NT3 + idx];
  p.a = sQ[2 * NT3 + idx];
  p.p = sQ[3 * NT3 + idx];
  p.u = sQ[4 * NT3 + idx];
  p.v = sQ[5 * NT3 + idx];
  p.w = sQ[6 * NT3 + idx];
  return p;
}
template <typename T>
G5D void
sq_store3(T * sQ, int idx, PrimT<T> const & p)
{
  sQ[0 * NT3 + idx] = p.r0;
  sQ[1 * NT3 + idx] = p.r1;
  sQ[2 * NT3 + idx] = p.a;
  sQ[3 * NT3 + idx] = p.p;
  sQ[4 * NT3 + idx] = p.u;
  sQ[5 * NT3 + idx] = p.v;
  sQ[6 * NT3 + idx] = p.w;
}
// face states: 7 values at slot k of a [7][N] array (N = 272 faces, or 16 ring faces)
template <int N = NF3, typename T>
G5D void
st_store3(T * A, int k, FaceT<T> const & f)
{
  A[0 * N + k] = f.r0;
  A[1 * N + k] = f.r1;
  A[2 * N + k] = f.a;
  A[3 * N + k] = f.p;
  A[4 * N + k] = f.u;
  A[5 * N + k] = f.v;
  A[6 * N + k] = f.w;
}
template <typename T>
G5D void
st_store3n(T * A, int n, FaceT<T> const & f) // row stride n
{
  A[0 * n] = f.r0;
  A[1 * n] = f.r1;
  A[2 * n] = f.a;
  A[3 * n] = f.p;
  A[4 * n] = f.u;
  A[5 * n] = f.v;
  A[6 * n] = f.w;
}
template <int N = NF3, typename T>
G5D FaceT<T>
st_load3(T const * A, int k)
{
  FaceT<T> f;
  f.r0 = A[0 * N + k];
  f.r1 = A[1 * N + k];
  f.a = A[2 * N + k];
  f.p = A[3 * N + k];
  f.u = A[4 * N + k];
  f.v = A[5 * N + k];
  f.w = A[6 * N + k];
  return f;
}
// face records at slot k of a [8][NF3] array plus E[k]: the FP64 flux (ar0, ar1, a0, a1, mu, mv,
// mw, us; energy in E) or the mixed path's star state (ro0, ro1, ro, uo, pt, t1, t2, phi0; etot)
G5D void
fl_store3(double * A, double * E, int k, Flux const & F)
{
  A[0 * NF3 + k] = F.ar0;
  A[1 * NF3 + k] = F.ar1;
  A[2 * NF3 + k] = F.a0;
  A[3 * NF3 + k] = F.a1;
  A[4 * NF3 + k] = F.mu;
  A[5 * NF3 + k] = F.mv;
  A[6 * NF3 + k] = F.mw;
  A[7 * NF3 + k] = F.us;
  E[k] = F.e;
}
G5D Flux
fl_load3(double const * A, double const * E, int k)
{
  Flux F;
  F.ar0 = A[0 * NF3 + k];
  F.ar1 = A[1 * NF3 + k];
  F.a0 = A[2 * NF3 + k];
  F.a1 = A[3 * NF3 + k];
  F.mu = A[4 * NF3 + k];
  F.mv = A[5 * NF3 + k];
  F.mw = A[6 * NF3 + k];
  F.us = A[7 * NF3 + k];
  F.e = E[k];
  return F;
}
G5D void
fl_store3(float * A, double * E, int k, FluxT<float> const & F)
{
  A[0 * NF3 + k] = F.ro0;
  A[1 * NF3 + k] = F.ro1;
  A[2 * NF3 + k] = F.ro;
  A[3 * NF3 + k] = F.uo;
  A[4 * NF3 + k] = F.pt;
  A[5 * NF3 + k] = F.t1;
  A[6 * NF3 + k] = F.t2;
  A[7 * NF3 + k] = F.phi0;
  E[k] = F.etot;
}
G5D FluxT<float>
fl_load3(float const * A, double const * E, int k)
{
  FluxT<float> F;
  F.ro0 = A[0 * NF3 + k];
  F.ro1 = A[1 * NF3 + k];
  F.ro = A[2 * NF3 + k];
  F.uo = A[3 * NF3 + k];
  F.pt = A[4 * NF3 + k];
  F.t1 = A[5 * NF3 + k];
  F.t2 = A[6 * NF3 + k];
  F.phi0 = A[7 * NF3 + k];
  F.etot = E[k];
  return F;
}

G5D int
aslot(int plane)
{
  return (plane + 3) % 3;
}

} // namespace

namespace
{
constexpr int QPL = 7 * NT3; // values per primitive plane

template <typename T>
G5D PrimT<T>
q3_load(T const * sQ, int slot, int idx)
{
  return sq_load3(sQ + slot * QPL, idx);
}

template <typename T, bool FAST>
G5D void
youngs_ring(T const * sQ, int P, int gi, int gj, T & nx, T & ny, T & nz)
{
  T a27[27];
#pragma unroll
  for (int dk = -1; dk <= 1; ++dk)
  {
    T const * plane = sQ + aslot(P + dk) * QPL + 2 * NT3; // alpha of plane P+dk
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int di = -1; di <= 1; ++di)
        a27[(di + 1) + 3 * (dj + 1) + 9 * (dk + 1)] = plane[tidx3(gi + di, gj + dj)];
  }
  youngs3<T, FAST>(a27, nx, ny, nz);
}

// slopes, Hancock terms and the requested face states of tile cell (gi,gj) in plane P
template <bool THINC, bool HANCOCK, typename T, bool FAST>
struct Pred3
{
  PrimT<T>    q;
  SlopeT<T>   sx, sy, sz;
  HancockT<T> h;
  T           nx, ny, nz;
  T           axm, axp, aym, ayp, azm, azp;

  G5D void
  compute(SolveDesc const & S, T const * sQ, int P, int gi, int gj, T dtdx)
  {
    const int      sp = aslot(P);
    const int      c = tidx3(gi, gj);
    q = q3_load(sQ, sp, c);
    const PrimT<T> qxm = q3_load(sQ, sp, tidx3(gi - 1, gj));
    const PrimT<T> qxp = q3_load(sQ, sp, tidx3(gi + 1, gj));
    const PrimT<T> qym = q3_load(sQ, sp, tidx3(gi, gj - 1));
    const PrimT<T> qyp = q3_load(sQ, sp, tidx3(gi, gj + 1));
    const PrimT<T> qzm = q3_load(sQ, aslot(P - 1), c);
    const PrimT<T> qzp = q3_load(sQ, aslot(P + 1), c);
    const T        st = static_cast<T>(S.slope_type);
    {
      const PrimT<T> qp[3] = { qxp, qyp, qzp }, qm[3] = { qxm, qym, qzm };
      SlopeT<T>      sl[3];
      slopes_axes<T, FAST, 3>(q, qp, qm, st, sl);
      sx = sl[0];
      sy = sl[1];
      sz = sl[2];
    }
    axm = qxm.a;
    axp = qxp.a;
    aym = qym.a;
    ayp = qyp.a;
    azm = qzm.a;
    azp = qzp.a;
    h = HancockT<T>{ T(0.0), T(0.0), T(0.0), T(0.0), T(0.0), T(0.0) };
    if (HANCOCK)
      h = hancock_terms<3, T, FAST>(S, q, sx, sy, sz, dtdx);
    nx = ny = nz = T(0.0);
    if (THINC and q.a > Par<T>::eps(S) and q.a < T(1.0) - Par<T>::eps(S))
      youngs_ring<T, FAST>(sQ, P, gi, gj, nx, ny, nz);
  }

  // THINC quantities of axis AX (AX = -1: of the run-time axis ax), shared by its two faces
  template <int AX>
  G5D ThincAxisT<T>
  axis(SolveDesc const & S, int ax = AX) const
  {
    if (!THINC)
      return ThincAxisT<T>{ false, T(0.0), T(0.0), T(0.0), T(0.0) };
    const int a = AX >= 0 ? AX : ax;
    const T   am = a == 0 ? axm : (a == 1 ? aym : azm);
    const T   ap = a == 0 ? axp : (a == 1 ? ayp : azp);
    const T   nd = a == 0 ? nx : (a == 1 ? ny : nz);
    return thinc_axis<T, FAST>(S, q.a, am, q.a, ap, fabs(nd));
  }

  // face state on axis AX (AX = -1: on the run-time axis ax), ib = -1 left face, +1 right face;
  // ta: axis<AX>(S, ax)
  template <int AX>
  G5D FaceT<T>
  face(SolveDesc const & S, int ib, ThincAxisT<T> const & ta, int ax = AX) const
  {
    if (!THINC)
      return face_muscl<3, AX, HANCOCK, T, FAST>(S, q, sx, sy, sz, h, T(0.5) * static_cast<T>(ib), ax);
    return face_thinc<3, AX, HANCOCK, T, FAST>(q, sx, sy, sz, h, ta, ib, ax);
  }
};
} // namespace

// ---- two-level static mesh (3-D) -------------------------------------------------------------
// Level 0 is the brick of nb[0] x nb[1] x nb[2] blocks, some of them refined into 2 x 2 x 2 blocks
// of half the cell length (level 1), at least one block away from the brick's border.  Slot
// tables: level 0 at S.base (-1 where refined), level 1 at S.base1 over the 2nb^3 fine grid (-1
// outside the refined blocks).  A cell is named by its global index at its level.
namespace
{
G5D int
slot3_l0(SolveDesc const & S, int32_t const * __restrict__ slots, int x, int y, int z)
{
  return slots[S.base + x + S.nb[0] * (y + S.nb[1] * z)];
}
G5D int
slot3_l1(SolveDesc const & S, int32_t const * __restrict__ slots, int x, int y, int z)
{
  return slots[S.base1 + x + 2 * S.nb[0] * (y + 2 * S.nb[1] * z)];
}
G5D Cons
cell3_of(double const * __restrict__ U, int slot, int ci, int cj, int ck)
{
  return load_cons<3, B3>(U, static_cast<int64_t>(slot) * 9 * 4096 + ci + B3 * (cj + B3 * ck), -1);
}
G5D double
var3_of(double const * __restrict__ U, int slot, int var, int ci, int cj, int ck)
{
  return U[static_cast<int64_t>(slot) * 9 * 4096 + var * 4096 + ci + B3 * (cj + B3 * ck)];
}

// coarse cell g inside a refined block: the average of its eight fine cells (summed x fastest, then
// y, then z; over 8) for E, rho u, rho v, rho w, alpha_rho 0/1, alpha 0; alpha 1 = 1 - alpha 0,
// rho = alpha_rho 0 + alpha_rho 1 (get_conservative_vars_restriction,
// scheme/ConvertToPrimitivesVariablesFunctor.cpp:214-250)
G5D Cons
restrict3(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U, int gx, int gy, int gz)
{
  Cons sum{};
  for (int k = 0; k < 8; ++k)
  {
    const int  fx = 2 * gx + (k & 1), fy = 2 * gy + ((k >> 1) & 1), fz = 2 * gz + (k >> 2);
    const Cons f = cell3_of(U, slot3_l1(S, slots, fx >> 4, fy >> 4, fz >> 4), fx & 15, fy & 15, fz & 15);
    sum.e = k == 0 ? f.e : sum.e + f.e;
    sum.mu = k == 0 ? f.mu : sum.mu + f.mu;
    sum.mv = k == 0 ? f.mv : sum.mv + f.mv;
    sum.mw = k == 0 ? f.mw : sum.mw + f.mw;
    sum.ar0 = k == 0 ? f.ar0 : sum.ar0 + f.ar0;
    sum.ar1 = k == 0 ? f.ar1 : sum.ar1 + f.ar1;
    sum.a0 = k == 0 ? f.a0 : sum.a0 + f.a0;
  }
  Cons u;
  u.e = sum.e / 8.0;
  u.mu = sum.mu / 8.0;
  u.mv = sum.mv / 8.0;
  u.mw = sum.mw / 8.0;
  u.ar0 = sum.ar0 / 8.0;
  u.ar1 = sum.ar1 / 8.0;
  u.a0 = sum.a0 / 8.0;
  u.a1 = 1.0 - u.a0;
  u.d = u.ar0 + u.ar1;
  return u;
}

// level-0 cell g of the parent's 3x3x3-block neighbourhood: a coarse leaf's cell or a restriction
G5D Cons
coarse3(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U, int gx, int gy, int gz)
{
  const int sl = slot3_l0(S, slots, gx >> 4, gy >> 4, gz >> 4);
  return sl >= 0 ? cell3_of(U, sl, gx & 15, gy & 15, gz & 15) : restrict3(S, slots, U, gx, gy, gz);
}
// one variable (offset var of the 9-variable layout) of that cell
G5D double
coarse3_var(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U, int gx, int gy,
            int gz, int var)
{
  const int sl = slot3_l0(S, slots, gx >> 4, gy >> 4, gz >> 4);
  if (sl >= 0)
    return var3_of(U, sl, var, gx & 15, gy & 15, gz & 15);
  double sum = 0.0;
  for (int k = 0; k < 8; ++k)
  {
    const int    fx = 2 * gx + (k & 1), fy = 2 * gy + ((k >> 1) & 1), fz = 2 * gz + (k >> 2);
    const double v = var3_of(U, slot3_l1(S, slots, fx >> 4, fy >> 4, fz >> 4), var, fx & 15, fy & 15, fz & 15);
    sum = k == 0 ? v : sum + v;
  }
  return sum / 8.0;
}

// the conservative interpolations of order 2 / 4 (S.prolong 2 / 3): tensor product of the 1-D
// half-cell weights over the parent's 3^3 (5^3) neighbourhood, x first, then y, then z
G5D double
conservative3_var(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U, int cx,
                  int cy, int cz, int px, int py, int pz, int var)
{
  const double w2[3] = { 0.125, 1.0, -0.125 };
  const double w4[5] = { -3.0 / 128.0, 22.0 / 128.0, 1.0, -22.0 / 128.0, 3.0 / 128.0 };
  const bool   o2 = S.prolong == 2;
  const int    m = o2 ? 1 : 2, n = 2 * m + 1;
  double       valz = 0.0;
  for (int kz = 0; kz < n; ++kz)
  {
    double valy = 0.0;
    for (int ky = 0; ky < n; ++ky)
    {
      double valx = 0.0;
      for (int kx = 0; kx < n; ++kx)
      {
        const int    ix = px == 0 ? kx : n - 1 - kx;
        const double x =
          (o2 ? w2[ix] : w4[ix]) * coarse3_var(S, slots, U, cx + kx - m, cy + ky - m, cz + kz - m, var);
        valx = kx == 0 ? x : valx + x;
      }
      const int    iy = py == 0 ? ky : n - 1 - ky;
      const double wy = o2 ? w2[iy] : w4[iy];
      valy = ky == 0 ? wy * valx : valy + wy * valx;
    }
    const int    iz = pz == 0 ? kz : n - 1 - kz;
    const double wz = o2 ? w2[iz] : w4[iz];
    valz = kz == 0 ? wz * valy : valz + wz * valy;
  }
  return valz;
}

// fine cell g outside the fine blocks: its coarse parent copied (SIMPLE_COPY), extrapolated with
// minmod slopes (EXTRAPOLATE_LINEAR_MINMOD, linear_extrapolate_using_limited_slopes 3-D,
// ConvertToPrimitivesVariablesFunctor.cpp:553-770), or conservatively interpolated
G5D Cons
prolong3(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U, int gx, int gy, int gz)
{
  const int cx = gx >> 1, cy = gy >> 1, cz = gz >> 1; // the fine ghost cells are at positive indices
  Cons      u;
  if (S.prolong >= 2)
  {
    const int px = gx & 1, py = gy & 1, pz = gz & 1;
    u.ar0 = conservative3_var(S, slots, U, cx, cy, cz, px, py, pz, 6);
    u.ar1 = conservative3_var(S, slots, U, cx, cy, cz, px, py, pz, 8);
    u.a0 = conservative3_var(S, slots, U, cx, cy, cz, px, py, pz, 5);
    u.e = conservative3_var(S, slots, U, cx, cy, cz, px, py, pz, 1);
    u.mu = conservative3_var(S, slots, U, cx, cy, cz, px, py, pz, 2);
    u.mv = conservative3_var(S, slots, U, cx, cy, cz, px, py, pz, 3);
    u.mw = conservative3_var(S, slots, U, cx, cy, cz, px, py, pz, 4);
    u.d = u.ar0 + u.ar1;
    u.a1 = 1.0 - u.a0;
    return u;
  }
  const Cons c = coarse3(S, slots, U, cx, cy, cz);
  if (S.prolong == 0)
    return c;
  const double hx = (gx & 1) ? 0.25 : -0.25, hy = (gy & 1) ? 0.25 : -0.25, hz = (gz & 1) ? 0.25 : -0.25;
  const Cons   xl = coarse3(S, slots, U, cx - 1, cy, cz), xr = coarse3(S, slots, U, cx + 1, cy, cz);
  const Cons   yl = coarse3(S, slots, U, cx, cy - 1, cz), yr = coarse3(S, slots, U, cx, cy + 1, cz);
  const Cons   zl = coarse3(S, slots, U, cx, cy, cz - 1), zr = coarse3(S, slots, U, cx, cy, cz + 1);
  auto         ex = [&](double C, double XL, double XR, double YL, double YR, double ZL, double ZR) {
    return C + hx * slope1_limited<double>(C, XR, XL, 1.0) + hy * slope1_limited<double>(C, YR, YL, 1.0) +
           hz * slope1_limited<double>(C, ZR, ZL, 1.0);
  };
  u.ar0 = ex(c.ar0, xl.ar0, xr.ar0, yl.ar0, yr.ar0, zl.ar0, zr.ar0);
  u.ar1 = ex(c.ar1, xl.ar1, xr.ar1, yl.ar1, yr.ar1, zl.ar1, zr.ar1);
  u.a0 = ex(c.a0, xl.a0, xr.a0, yl.a0, yr.a0, zl.a0, zr.a0);
  u.e = ex(c.e, xl.e, xr.e, yl.e, yr.e, zl.e, zr.e);
  u.mu = ex(c.mu, xl.mu, xr.mu, yl.mu, yr.mu, zl.mu, zr.mu);
  u.mv = ex(c.mv, xl.mv, xr.mv, yl.mv, yr.mv, zl.mv, zr.mv);
  u.mw = ex(c.mw, xl.mw, xr.mw, yl.mw, yr.mw, zl.mw, zr.mw);
  u.d = u.ar0 + u.ar1;
  u.a1 = 1.0 - u.a0;
  return u;
}

// conservative state of tile cell (gi, gj, gk) of a block of a two-level solve (fill_ghosts,
// ConvertToPrimitivesVariablesFunctor.cpp:884-918): the block's own cell, a same-level
// neighbour's (border conditions resolved as on one level), a restriction or a prolongation
template <bool INLET>
G5D Cons
ghost3_two_level(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U,
                 BlockDesc const & bd, int gi, int gj, int gk)
{
  const int dx = gi < 0 ? -1 : (gi >= B3 ? 1 : 0), dy = gj < 0 ? -1 : (gj >= B3 ? 1 : 0),
            dz = gk < 0 ? -1 : (gk >= B3 ? 1 : 0);
  const int nx = bd.bx + dx, ny = bd.by + dy, nz = bd.bz + dz;
  const int gx = bd.bx * B3 + gi, gy = bd.by * B3 + gj, gz = bd.bz * B3 + gk;
  if (bd.level == 0)
  {
    if (nx >= 0 and nx < S.nb[0] and ny >= 0 and ny < S.nb[1] and nz >= 0 and nz < S.nb[2] and
        slot3_l0(S, slots, nx, ny, nz) < 0)
      return restrict3(S, slots, U, gx, gy, gz);
    int           neg;
    const int64_t off = resolve_cell<3, B3>(S, slots, bd.bx, bd.by, bd.bz, gi, gj, gk, neg);
    Cons          u = load_cons<3, B3>(U, off, -1);
    apply_border<3, INLET>(S, u, neg);
    return u;
  }
  const int sl = slot3_l1(S, slots, nx, ny, nz); // fine blocks sit inside the fine grid's interior
  return sl >= 0 ? cell3_of(U, sl, gi - dx * B3, gj - dy * B3, gk - dz * B3) : prolong3(S, slots, U, gx, gy, gz);
}

// interface records of a coarse slot: [side][face cell c][fine face sub][9 values] (see
// kIflux3Doubles); the four fine faces over a coarse face cell are summed in sub order 0..3 and
// divided by 8 (get_flux_from_fine_neighbor, ReadFluxesAndConservativeUpdateFunctor.h:256-307)
G5D int64_t
iflux3_at(int slot, int side, int c, int sub)
{
  static_assert(kIflux3Doubles == 6 * 256 * 4 * 9, "interface record layout");
  return (((static_cast<int64_t>(slot) * 6 + side) * 256 + c) * 4 + sub) * 9;
}
G5D void
iflux3_store(double * __restrict__ R, int64_t o, Flux const & F)
{
  R[o + 0] = F.ar0;
  R[o + 1] = F.ar1;
  R[o + 2] = F.a0;
  R[o + 3] = F.a1;
  R[o + 4] = F.e;
  R[o + 5] = F.mu;
  R[o + 6] = F.mv;
  R[o + 7] = F.mw;
  R[o + 8] = F.us;
}
G5D Flux
iflux3_sum(double const * __restrict__ R, int slot, int side, int c)
{
  const int64_t o = iflux3_at(slot, side, c, 0);
  double        v[9];
#pragma unroll
  for (int q = 0; q < 9; ++q)
    v[q] = ((R[o + q] + R[o + 9 + q]) + R[o + 18 + q]) + R[o + 27 + q];
  Flux F;
  F.ar0 = v[0] / 8.0;
  F.ar1 = v[1] / 8.0;
  F.a0 = v[2] / 8.0;
  F.a1 = v[3] / 8.0;
  F.e = v[4] / 8.0;
  F.mu = v[5] / 8.0;
  F.mv = v[6] / 8.0;
  F.mw = v[7] / 8.0;
  F.us = v[8] / 8.0;
  return F;
}
// where a fine block's boundary face record goes: the coarse neighbour's slot, side, face cell and
// sub index, for the fine face at global fine cell (gx, gy, gz) on the fine block's side sd (0..5:
// xmin, xmax, ymin, ymax, zmin, zmax); transverse axes t1, t2 = (axis+1)%3, (axis+2)%3 as in the
// reference's face-sibling sum, the face cell index t1 + 16 t2 in the coarse block
G5D int64_t
iflux3_dest(SolveDesc const & S, int32_t const * __restrict__ slots, BlockDesc const & bd, int sd, int gx, int gy,
            int gz)
{
  const int ax = sd >> 1;
  const int n1[3] = { bd.bx + (ax == 0 ? ((sd & 1) ? 1 : -1) : 0), bd.by + (ax == 1 ? ((sd & 1) ? 1 : -1) : 0),
                      bd.bz + (ax == 2 ? ((sd & 1) ? 1 : -1) : 0) };
  const int cb[3] = { n1[0] >> 1, n1[1] >> 1, n1[2] >> 1 };
  const int g[3] = { gx, gy, gz };
  const int t1 = (ax + 1) % 3, t2 = (ax + 2) % 3;
  const int c = ((g[t1] >> 1) - 16 * cb[t1]) + 16 * ((g[t2] >> 1) - 16 * cb[t2]);
  const int sub = (g[t1] & 1) + 2 * (g[t2] & 1);
  return iflux3_at(slot3_l0(S, slots, cb[0], cb[1], cb[2]), sd ^ 1, c, sub);
}
} // namespace

// AMR: the launch list may hold blocks of two-level solves (FP64 records only): their ghosts come
// from ghost3_two_level, a coarse block takes the faces it shares with fine blocks from A.iflux,
// and a fine block bordering coarse blocks writes its boundary face records there (the fine blocks
// that do run in a launch of their own before the rest of the list)
template <bool THINC, bool HANCOCK, bool ZSPLIT, bool INLET, bool MP, typename T, bool FAST, bool AMR = false>
G5D void
step3d_body(StepArgs const & A, int item)
{
  static_assert(!AMR or sizeof(T) == sizeof(double), "two-level solves exchange FP64 face records");
  __shared__ T      sQ[3 * QPL];
  __shared__ T      sXL[7 * NF3];
  __shared__ T      sYL[7 * NF3];
  __shared__ T      sXR[7 * 16];
  __shared__ T      sYR[7 * 16];
  __shared__ T      sFX[8 * NF3];
  __shared__ T      sFY[8 * NF3];
  __shared__ double sFXE[NF3];
  __shared__ double sFYE[NF3];
  __shared__ double sRed[4];
  __shared__ int    sSlot[kMaxColumn * 27];

  // the work item: L <= col_len consecutive z blocks (x, y, z0 + k) of one solve, marched through as
  // one column of planes 0 .. 16 L - 1 (a block's own planes are 16 k .. 16 k + 15); two-level
  // solves run one block per item
  const int       t = threadIdx.x;
  const int       per = A.col_stride / A.col_len;
  const int       first = (item / per) * A.col_stride + (item % per) * A.col_len;
  int             L = 1;
  while (L < A.col_len and A.blocks[first + L].solve >= 0)
    ++L;
  const BlockDesc bd = A.blocks[first];
  if (bd.solve < 0)
    return; // padding of a short column
  const SolveDesc S = AMR ? at_level(A.solves[bd.solve], bd.level) : A.solves[bd.solve];
  SolveCtl *      C = A.ctl + bd.solve;
  constexpr int   NC = 4096;
  // offset of block k of the column in the state buffers
  auto            col_off = [&](int k) { return static_cast<int64_t>(A.blocks[first + k].slot) * 9 * NC; };
  const int       zend = 16 * L; // first plane above the column
  const bool      host_dt = A.dt_fixed > 0.0;
  if (!host_dt and C->active == 0)
    return; // finished solve: its state stays in buffer `parity`
  const int      par = C->parity;
  double const * U = A.U_in ? A.U_in : (par ? A.buf1 : A.buf0);
  double *       U2 = A.U_out ? A.U_out : (par ? A.buf0 : A.buf1);
  const double   dt = host_dt ? A.dt_fixed : C->dt;
  const double   dtdx = dt / S.dx;
  const T        dtdxT = static_cast<T>(dtdx);

  const int  wave = t >> 6, lane = t & 63;
  const int  i = t & 15, j = t >> 4;
  // extra conversion item
  const bool has_x = wave == 0 or wave == 3 or (wave == 2 and lane < 16);
  int        xgi = 0, xgj = 0;
  if (wave == 0)
    ring_cell3(lane, xgi, xgj);
  else if (wave == 3)
    outer_cell3(lane, xgi, xgj);
  else if (wave == 2 and lane < 16)
    outer_cell3(64 + lane, xgi, xgj);
  // ring predictor (wave 1) and ring Riemann problems (wave 2)
  const int rs = (lane >> 4) & 3, rr = lane & 15;
  int       rgi = 0, rgj = 0;
  ring_cell3(lane, rgi, rgj);
  const bool z_early = !ZSPLIT or wave != 1;

  // two-level solve: sides (bit s: xmin, xmax, ymin, ymax, zmin, zmax) whose neighbour is finer (a
  // coarse block) or coarser (a fine block)
  const bool two_level = AMR and S.base1 >= 0;
  int        finer = 0, coarser = 0;
  if (two_level)
  {
#pragma unroll
    for (int sd = 0; sd < 6; ++sd)
    {
      const int a = sd >> 1, o = (sd & 1) ? 1 : -1;
      const int nx = bd.bx + (a == 0 ? o : 0), ny = bd.by + (a == 1 ? o : 0), nz = bd.bz + (a == 2 ? o : 0);
      if (bd.level == 0)
        finer |= (nx >= 0 and nx < S.nb[0] and ny >= 0 and ny < S.nb[1] and nz >= 0 and nz < S.nb[2] and
                  slot3_l0(S, A.slots, nx, ny, nz) < 0)
                   ? 1 << sd
                   : 0;
      else
        coarser |= slot3_l1(S, A.slots, nx, ny, nz) < 0 ? 1 << sd : 0;
    }
  }

  // neighbour slots once per block (LDS), so that a fetch issues its loads without a dependent
  // global read of the slot table (a two-level solve's ghosts are resolved by level instead)
  // 27 entries per block of the column: more than one per thread beyond 9 blocks
  if (!two_level)
    for (int e = t; e < 27 * L; e += 256)
      fill_slot_table<3>(S, A.slots, bd.bx, bd.by, bd.bz + e / 27, e % 27, sSlot + 27 * (e / 27));
  __syncthreads();
  // raw loads; the WALL momentum sign is applied at the conversion, so that the prefetch of a
  // plane is not waited for in the phase that issues it
  // the block of the column that plane g (in [-2, 16 L + 1]) is resolved from, and g relative to it
  auto blk_of = [&](int g) { return g < 0 ? 0 : (g >= zend ? L - 1 : g >> 4); };
  auto fetch_raw = [&](int gi, int gj, int gk, int & neg) {
    if constexpr (AMR)
      if (two_level)
      {
        neg = -1; // border conditions resolved inside
        return ghost3_two_level<INLET>(S, A.slots, U, bd, gi, gj, gk);
      }
    const int     kb = blk_of(gk);
    const int64_t off = resolve_cell_tab<3, B3, INLET>(S, sSlot + 27 * kb, bd.bx, bd.by, bd.bz + kb, gi, gj,
                                                       gk - 16 * kb, neg);
    return load_cons<3, B3>(U, off, -1);
  };
  auto fetch = [&](int gi, int gj, int gk) {
    int  neg;
    Cons u = fetch_raw(gi, gj, gk, neg);
    apply_border<3, INLET>(S, u, neg);
    return u;
  };
  auto put = [&](int P, int gi, int gj, Cons const & u) {
    PrimT<T> q;
    to_prim<3, FAST>(S, u, q);
    sq_store3(sQ + aslot(P) * QPL, tidx3(gi, gj), q);
  };

  // carried state
  Cons     pf_i, pf_x;   // prefetched raw state of the next plane to convert
  int      neg_i = -1, neg_x = -1; // ... and its WALL momentum component
  Cons     uraw_cur;     // raw state of the owned column in plane P
  Cons     uraw_next;    // ... in plane P+1
  Cons     u2prev;       // plane P-1 after its x and y updates
  double   uaprev0 = 0.0, uaprev1 = 0.0;
  double   stprev0 = 1.0, stprev1 = 1.0; // source factors of plane P-1 (1 unless MILLER_PUCKETT)
  FluxT<T> zFprev;       // z flux on face P-1
  FaceT<T> zL;           // +z face state of plane P-1
  FaceT<T> zLnext;       // +z face state of plane P (wave 1 with ZSPLIT, until its z solve)
  double   inv_max = 0.0;
  uint32_t fail = 0u; // G5_FAIL_* bits of the lane's new cells

  // column offsets at plane 0 of the current block of the own column and of the extra conversion
  // item: within a block's z range (dz = 0) no border transform acts on z, so its plane k is at
  // offset + 256 k (the extra item's offset is resolved again when the march enters a block)
  const int     own_c = i + 16 * j;
  const bool    inlet_block = INLET and S.inlet != G5_INLET_NONE and bd.bx == 0; // halo cells may be inflow cells
  int           x_neg0 = -1;
  int64_t       x_off0 = 0;
  if (has_x and !two_level)
    x_off0 = resolve_cell_tab<3, B3, INLET>(S, sSlot, bd.bx, bd.by, bd.bz, xgi, xgj, 0, x_neg0);

  // prologue: planes -2 and -1 into LDS, plane 0 prefetched
  {
    const Cons a = fetch(i, j, -2);
    put(-2, i, j, a);
    if (has_x)
      put(-2, xgi, xgj, fetch(xgi, xgj, -2));
    const Cons b = fetch(i, j, -1);
    put(-1, i, j, b);
    if (has_x)
      put(-1, xgi, xgj, fetch(xgi, xgj, -1));
    pf_i = fetch_raw(i, j, 0, neg_i);
    if (has_x)
      pf_x = fetch_raw(xgi, xgj, 0, neg_x);
  }

  // z face P (between planes P-1 and P): the Riemann solve, then the completion of plane P-1 --
  // in FP64 after the plane's last barrier (the solve's flux stays in registers meanwhile), in the
  // mixed form, whose registers are capped at 256, right after the solve
  constexpr bool LATE = sizeof(T) == sizeof(double);
  FluxT<T>       zF; // flux of z face P, from its solve to the completion
  auto           z_complete = [&](int P) {
    if (P >= 1)
    {
      apply_face_pair<3, FAST>(u2prev, zFprev, zF, uaprev0, uaprev1, dtdx, 2, stprev0, stprev1);
      fix_volume_fractions<FAST>(u2prev);
      store_cons<3, B3>(U2, col_off((P - 1) >> 4) + i + 16 * j + 256 * ((P - 1) & 15), u2prev);
      double       pc;
      const double rate = cell_rate<3, T, FAST>(S, u2prev, pc);
      inv_max = fmax(inv_max, rate);
      fail |= cell_failure(S, u2prev, rate, pc);
    }
    zFprev = zF;
  };
  auto z_solve = [&](int P, FaceT<T> const & fzm) {
    zF = riemann<3, FAST>(S, zL, fzm, dtdx, 2);
    if constexpr (AMR)
    {
      // the block's bottom (face 0) and top (face 16) z faces across a level interface
      const int sd = P == 0 ? 4 : 5;
      if (P == 0 or P == zend)
      {
        if ((finer >> sd) & 1)
          zF = iflux3_sum(A.iflux, bd.slot, sd, i + 16 * j);
        else if ((coarser >> sd) & 1)
          iflux3_store(A.iflux, iflux3_dest(S, A.slots, bd, sd, bd.bx * B3 + i, bd.by * B3 + j,
                                            bd.bz * B3 + (P == 0 ? 0 : zend - 1)), zF);
      }
    }
    if (!LATE)
      z_complete(P);
  };

  for (int P = -1; P <= zend; ++P)
  {
    const bool inner_plane = P >= 0 and P < zend;
    // (a) plane P+1 from the prefetched raw state
    negate_momentum(pf_i, neg_i); // the own column is never left of x-min: no inflow cell
    put(P + 1, i, j, pf_i);
    uraw_next = pf_i;
    if (has_x)
    {
      apply_border<3, INLET>(S, pf_x, neg_x);
      put(P + 1, xgi, xgj, pf_x);
    }
    __syncthreads(); // (b)

    // (c) prefetch plane P+2, predictor of plane P
    if (P + 2 < zend and not inlet_block)
    {
      // planes of the column's own z range: within a block each lane's column offset is linear in
      // the plane
      const int kb = (P + 2) >> 4, pl = (P + 2) & 15;
      pf_i = load_cons<3, B3>(U, col_off(kb) + own_c + 256 * pl, -1);
      neg_i = -1;
      if (has_x)
      {
        if (two_level)
          pf_x = fetch_raw(xgi, xgj, P + 2, neg_x);
        else
        {
          if (pl == 0 and kb > 0) // the march enters block kb
            x_off0 = resolve_cell_tab<3, B3, INLET>(S, sSlot + 27 * kb, bd.bx, bd.by, bd.bz + kb, xgi, xgj, 0, x_neg0);
          pf_x = load_cons<3, B3>(U, x_off0 + 256 * pl, -1);
          neg_x = x_neg0;
        }
      }
    }
    else if (P + 2 <= zend + 1)
    {
      pf_i = fetch_raw(i, j, P + 2, neg_i);
      if (has_x)
        pf_x = fetch_raw(xgi, xgj, P + 2, neg_x);
    }
    FaceT<T> fxm, fym, fzm;
    {
      Pred3<THINC, HANCOCK, T, FAST> pr;
      pr.compute(S, sQ, P, i, j, dtdxT);
      const ThincAxisT<T> tz = pr.template axis<2>(S);
      fzm = pr.template face<2>(S, -1, tz);
      const FaceT<T> fzp = pr.template face<2>(S, 1, tz);
      if (inner_plane)
      {
        const ThincAxisT<T> tx = pr.template axis<0>(S);
        const ThincAxisT<T> ty = pr.template axis<1>(S);
        fxm = pr.template face<0>(S, -1, tx);
        fym = pr.template face<1>(S, -1, ty);
        st_store3(sXL, fx3(i + 1, j), pr.template face<0>(S, 1, tx));
        st_store3(sYL, fy3(i, j + 1), pr.template face<1>(S, 1, ty));
      }
      if (P >= 0 and z_early)
        z_solve(P, fzm);
      if (z_early or P < 0)
        zL = fzp;
      else
        zLnext = fzp; // wave 1 keeps both z faces until after the exchange
    }
    if (wave == 1 and inner_plane)
    {
      // ring cell of side rs: its face on the block boundary, one reconstruction for all lanes
      Pred3<THINC, HANCOCK, T, FAST> pr;
      pr.compute(S, sQ, P, rgi, rgj, dtdxT);
      const FaceT<T> f = pr.template face<-1>(S, (rs & 1) ? -1 : 1, pr.template axis<-1>(S, rs >> 1), rs >> 1);
      T *            dst = rs == 0 ? sXL + fx3(0, rr) : (rs == 1 ? sXR + rr : (rs == 2 ? sYL + fy3(rr, 0) : sYR + rr));
      st_store3n(dst, (rs & 1) ? 16 : NF3, f);
    }
    __syncthreads(); // (d)

    // (e) Riemann problems on the x and y faces of plane P
    if (inner_plane)
    {
      FluxT<T> fx = riemann<3, FAST>(S, st_load3(sXL, fx3(i, j)), fxm, dtdx, 0);
      FluxT<T> fy = riemann<3, FAST>(S, st_load3(sYL, fy3(i, j)), fym, dtdx, 1);
      if constexpr (AMR)
      {
        // a coarse block's faces on a finer neighbour take the fine faces' records (x sides: face
        // cell y + 16 z; y sides: z + 16 x)
        if (i == 0 and (finer & 1))
          fx = iflux3_sum(A.iflux, bd.slot, 0, j + 16 * P);
        if (j == 0 and (finer & 4))
          fy = iflux3_sum(A.iflux, bd.slot, 2, P + 16 * i);
      }
      fl_store3(sFX, sFXE, fx3(i, j), fx);
      fl_store3(sFY, sFYE, fy3(i, j), fy);
      if (wave == 0 and (rs & 1))
      {
        // the last x face (rs 1) or y face (rs 3) of row/column rr: one solve on the run-time axis
        const int ax = rs >> 1;
        const int k = ax == 0 ? fx3(16, rr) : fy3(rr, 16);
        const FaceT<T> L = st_load3(ax == 0 ? sXL : sYL, k);
        const FaceT<T> R = st_load3<16>(ax == 0 ? sXR : sYR, rr);
        FluxT<T>       fr = riemann<3, FAST>(S, L, R, dtdx, ax);
        if constexpr (AMR)
          if ((finer >> rs) & 1)
            fr = iflux3_sum(A.iflux, bd.slot, rs, ax == 0 ? rr + 16 * P : P + 16 * rr);
        fl_store3(ax == 0 ? sFX : sFY, ax == 0 ? sFXE : sFYE, k, fr);
      }
    }
    if (!z_early)
    {
      if (P >= 0)
      {
        z_solve(P, fzm);
        zL = zLnext;
      }
    }
    __syncthreads(); // (f)

    // (g) completion of plane P-1 (z update, normalisation, store, CFL), then (i) the x and y
    // update of plane P
    if constexpr (AMR)
    {
      // a fine block on a coarse neighbour: its boundary x/y face records of plane P, one per lane
      // of wave 0 (side t/16, row or column t%16)
      const int sd = t >> 4, r = t & 15;
      if (inner_plane and coarser != 0 and t < 64 and ((coarser >> sd) & 1))
      {
        const Flux F = sd < 2 ? fl_load3(sFX, sFXE, fx3(sd == 0 ? 0 : 16, r)) : fl_load3(sFY, sFYE, fy3(r, sd == 2 ? 0 : 16));
        const int  gi = sd == 0 ? 0 : (sd == 1 ? 15 : r), gj = sd < 2 ? r : (sd == 2 ? 0 : 15);
        iflux3_store(A.iflux, iflux3_dest(S, A.slots, bd, sd, bd.bx * B3 + gi, bd.by * B3 + gj, bd.bz * B3 + P), F);
      }
    }
    if (LATE and P >= 0)
      z_complete(P);
    if (inner_plane)
    {
      u2prev = uraw_cur;
      double stc0 = 1.0, stc1 = 1.0;
      if constexpr (MP) // the list holds a MILLER_PUCKETT solve; this block's solve decides
        if (S.source_term == G5_SOURCE_MILLER_PUCKETT)
          source_factors<3>(S, uraw_cur, stc0, stc1);
      apply_face_pair<3, FAST>(u2prev, fl_load3(sFX, sFXE, fx3(i, j)), fl_load3(sFX, sFXE, fx3(i + 1, j)), uraw_cur.a0,
                         uraw_cur.a1, dtdx, 0, stc0, stc1);
      apply_face_pair<3, FAST>(u2prev, fl_load3(sFY, sFYE, fy3(i, j)), fl_load3(sFY, sFYE, fy3(i, j + 1)), uraw_cur.a0,
                         uraw_cur.a1, dtdx, 1, stc0, stc1);
      uaprev0 = uraw_cur.a0;
      uaprev1 = uraw_cur.a1;
      if constexpr (MP)
      {
        stprev0 = stc0;
        stprev1 = stc1;
      }
    }
    uraw_cur = uraw_next;
  }

  report_failure(C, fail);
  const double inv = wave_max(inv_max);
  if (lane == 0)
    sRed[wave] = inv;
  __syncthreads();
  if (t == 0)
  {
    const double m = fmax(fmax(sRed[0], sRed[1]), fmax(sRed[2], sRed[3]));
    atomicMax(&C->invdt_bits, static_cast<unsigned long long>(__double_as_longlong(m)));
  }
}

// FP64: one wavefront per SIMD (about 330 registers); mixed: two wavefronts per SIMD (at most
// 256 registers, and 72 KB of LDS per workgroup)
// One workgroup per block.  (A persistent form, workgroups looping over several blocks, made
// an illegal memory access on the GPU with no cause found in the code; the runtime never asks
// for it in 3-D: the launch must have wg_per_xcd == 0.)
template <bool THINC, bool HANCOCK, bool ZSPLIT, bool INLET, bool MP, typename T, bool FAST = false, bool AMR = false>
G5D void
step3d_blocks(StepArgs const & A)
{
  int b, end, stride;
  block_range(A, b, end, stride);
  if (b < end)
    step3d_body<THINC, HANCOCK, ZSPLIT, INLET, MP, T, FAST, AMR>(A, b);
}

// a launch list with blocks of two-level solves: the first n_first blocks (the fine blocks that
// border coarse blocks) write their interface records, then the rest of the list reads them.
// The two-level instances compile the inflow-border and Miller-Puckett code (each block applies
// its own solve's choices) and the split z solve.
template <bool THINC, bool HANCOCK, bool FAST>
__global__ void __launch_bounds__(256)
k_step3d_two_level(StepArgs A)
{
  step3d_blocks<THINC, HANCOCK, true, true, true, double, FAST, true>(A);
}

template <bool FAST>
static void
launch3_two_level(StepArgs const & A, int n_first, bool thinc, bool hancock, hipStream_t stream)
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
      hipLaunchKernelGGL((k_step3d_two_level<true, true, FAST>), grid, block, 0, stream, a);
    else if (thinc)
      hipLaunchKernelGGL((k_step3d_two_level<true, false, FAST>), grid, block, 0, stream, a);
    else if (hancock)
      hipLaunchKernelGGL((k_step3d_two_level<false, true, FAST>), grid, block, 0, stream, a);
    else
      hipLaunchKernelGGL((k_step3d_two_level<false, false, FAST>), grid, block, 0, stream, a);
  }
}

} // namespace g5
