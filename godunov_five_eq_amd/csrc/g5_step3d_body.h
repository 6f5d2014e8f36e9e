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
    sx = slopes_of<T, FAST>(q, qxp, qxm, st);
    sy = slopes_of<T, FAST>(q, qyp, qym, st);
    sz = slopes_of<T, FAST>(q, qzp, qzm, st);
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

template <bool THINC, bool HANCOCK, bool ZSPLIT, bool INLET, bool MP, typename T, bool FAST>
G5D void
step3d_body(StepArgs const & A, int blk)
{
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
  __shared__ int    sSlot[27];

  const int       t = threadIdx.x;
  const BlockDesc bd = A.blocks[blk];
  const SolveDesc S = A.solves[bd.solve];
  SolveCtl *      C = A.ctl + bd.solve;
  constexpr int   NC = 4096;
  const int64_t   blk_off = static_cast<int64_t>(bd.slot) * 9 * NC;
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

  // neighbour slots once per block (LDS), so that a fetch issues its loads without a dependent
  // global read of the slot table
  fill_slot_table<3>(S, A.slots, bd.bx, bd.by, bd.bz, t, sSlot);
  __syncthreads();
  // raw loads; the WALL momentum sign is applied at the conversion, so that the prefetch of a
  // plane is not waited for in the phase that issues it
  auto fetch_raw = [&](int gi, int gj, int gk, int & neg) {
    const int64_t off = resolve_cell_tab<3, B3, INLET>(S, sSlot, bd.bx, bd.by, bd.bz, gi, gj, gk, neg);
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

  // column offsets at plane 0 of the own column and of the extra conversion item: within the
  // block's z range (dz = 0) no border transform acts on z, so plane k is at offset + 256 k
  const int64_t own_off0 = blk_off + i + 16 * j;
  const bool    inlet_block = INLET and S.inlet != G5_INLET_NONE and bd.bx == 0; // halo cells may be inflow cells
  int           x_neg0 = -1;
  int64_t       x_off0 = 0;
  if (has_x)
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
      apply_face_pair<3>(u2prev, zFprev, zF, uaprev0, uaprev1, dtdx, 2, stprev0, stprev1);
      fix_volume_fractions<FAST>(u2prev);
      store_cons<3, B3>(U2, blk_off + i + 16 * j + 256 * (P - 1), u2prev);
      double       pc;
      const double rate = cell_rate<3, T, FAST>(S, u2prev, pc);
      inv_max = fmax(inv_max, rate);
      fail |= cell_failure(S, u2prev, rate, pc);
    }
    zFprev = zF;
  };
  auto z_solve = [&](int P, FaceT<T> const & fzm) {
    zF = riemann<3, FAST>(S, zL, fzm, dtdx, 2);
    if (!LATE)
      z_complete(P);
  };

  for (int P = -1; P <= 16; ++P)
  {
    const bool inner_plane = P >= 0 and P <= 15;
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
    if (P + 2 <= 15 and not inlet_block)
    {
      // planes of the block's own z range: each lane's column offset is linear in the plane
      pf_i = load_cons<3, B3>(U, own_off0 + 256 * (P + 2), -1);
      neg_i = -1;
      if (has_x)
      {
        pf_x = load_cons<3, B3>(U, x_off0 + 256 * (P + 2), -1);
        neg_x = x_neg0;
      }
    }
    else if (P + 2 <= 17)
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
      fl_store3(sFX, sFXE, fx3(i, j), riemann<3, FAST>(S, st_load3(sXL, fx3(i, j)), fxm, dtdx, 0));
      fl_store3(sFY, sFYE, fy3(i, j), riemann<3, FAST>(S, st_load3(sYL, fy3(i, j)), fym, dtdx, 1));
      if (wave == 0 and (rs & 1))
      {
        // the last x face (rs 1) or y face (rs 3) of row/column rr: one solve on the run-time axis
        const int ax = rs >> 1;
        const int k = ax == 0 ? fx3(16, rr) : fy3(rr, 16);
        const FaceT<T> L = st_load3(ax == 0 ? sXL : sYL, k);
        const FaceT<T> R = st_load3<16>(ax == 0 ? sXR : sYR, rr);
        fl_store3(ax == 0 ? sFX : sFY, ax == 0 ? sFXE : sFYE, k, riemann<3, FAST>(S, L, R, dtdx, ax));
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
    if (LATE and P >= 0)
      z_complete(P);
    if (inner_plane)
    {
      u2prev = uraw_cur;
      double stc0 = 1.0, stc1 = 1.0;
      if constexpr (MP) // the list holds a MILLER_PUCKETT solve; this block's solve decides
        if (S.source_term == G5_SOURCE_MILLER_PUCKETT)
          source_factors<3>(S, uraw_cur, stc0, stc1);
      apply_face_pair<3>(u2prev, fl_load3(sFX, sFXE, fx3(i, j)), fl_load3(sFX, sFXE, fx3(i + 1, j)), uraw_cur.a0,
                         uraw_cur.a1, dtdx, 0, stc0, stc1);
      apply_face_pair<3>(u2prev, fl_load3(sFY, sFYE, fy3(i, j)), fl_load3(sFY, sFYE, fy3(i, j + 1)), uraw_cur.a0,
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
template <bool THINC, bool HANCOCK, bool ZSPLIT, bool INLET, bool MP, typename T, bool FAST = false>
G5D void
step3d_blocks(StepArgs const & A)
{
  int b, end, stride;
  block_range(A, b, end, stride);
  if (b < end)
    step3d_body<THINC, HANCOCK, ZSPLIT, INLET, MP, T, FAST>(A, b);
}

} // namespace g5
