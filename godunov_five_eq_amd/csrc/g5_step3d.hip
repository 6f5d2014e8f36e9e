// g5_step3d.hip -- fused 3-D five-equation step, one 16^3 block per workgroup, marching in z.
//
// A 20^3 FP64 primitive tile with ghost width 2 (576 KB for 9 variables) does not fit in LDS, so
// the block is swept plane by plane (2.5-D): LDS holds the primitives of the current z plane
// (20x20) for the x/y stencils, the z neighbours live in registers of the thread that owns the
// column.  Per plane P = -1..16 the workgroup
//   a. gathers and converts plane P+1 (ghost/border resolution on the fly),
//   b. computes slopes, sound speed and Hancock terms of plane P and its face states,
//   c. solves HLLC on the x and y faces of plane P and on the z face between P-1 and P,
//   d. exchanges the x/y fluxes through LDS, applies the x then y update to plane P and the z
//      update (with normalisation, store and CFL contribution) to plane P-1.
// Same numerics as scheme/GodunovImplemV0.cpp:83-151 (see g5_step2d.hip for the per-function
// citations); the z face flux of plane P-1 is kept in registers until plane P is done.
// Threads 0..255 (waves 0-3) own the interior columns (t%16, t/16); threads 256..319 (wave 4)
// own the 64 ring columns whose single face state the block needs and convert the 80 outer
// columns of the 20x20 plane.
#include <cstdlib>

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

G5D Prim
sq_load3(double const * sQ, int idx)
{
  Prim p;
  p.r0 = sQ[0 * NT3 + idx];
  p.r1 = sQ[1 * NT3 + idx];
  p.a = sQ[2 * NT3 + idx];
  p.p = sQ[3 * NT3 + idx];
  p.u = sQ[4 * NT3 + idx];
  p.v = sQ[5 * NT3 + idx];
  p.w = sQ[6 * NT3 + idx];
  return p;
}
G5D void
sq_store3(double * sQ, int idx, Prim const & p)
{
  sQ[0 * NT3 + idx] = p.r0;
  sQ[1 * NT3 + idx] = p.r1;
  sQ[2 * NT3 + idx] = p.a;
  sQ[3 * NT3 + idx] = p.p;
  sQ[4 * NT3 + idx] = p.u;
  sQ[5 * NT3 + idx] = p.v;
  sQ[6 * NT3 + idx] = p.w;
}
G5D void
st_store3(double * A, int k, Face const & f)
{
  A[0 * NF3 + k] = f.r0;
  A[1 * NF3 + k] = f.r1;
  A[2 * NF3 + k] = f.a;
  A[3 * NF3 + k] = f.p;
  A[4 * NF3 + k] = f.u;
  A[5 * NF3 + k] = f.v;
  A[6 * NF3 + k] = f.w;
}
G5D Face
st_load3(double const * A, int k)
{
  Face f;
  f.r0 = A[0 * NF3 + k];
  f.r1 = A[1 * NF3 + k];
  f.a = A[2 * NF3 + k];
  f.p = A[3 * NF3 + k];
  f.u = A[4 * NF3 + k];
  f.v = A[5 * NF3 + k];
  f.w = A[6 * NF3 + k];
  return f;
}
G5D void
fl_store3(double * A, int k, Flux const & F)
{
  A[0 * NF3 + k] = F.ar0;
  A[1 * NF3 + k] = F.ar1;
  A[2 * NF3 + k] = F.a0;
  A[3 * NF3 + k] = F.a1;
  A[4 * NF3 + k] = F.e;
  A[5 * NF3 + k] = F.mu;
  A[6 * NF3 + k] = F.mv;
  A[7 * NF3 + k] = F.mw;
  A[8 * NF3 + k] = F.us;
}
G5D Flux
fl_load3(double const * A, int k)
{
  Flux F;
  F.ar0 = A[0 * NF3 + k];
  F.ar1 = A[1 * NF3 + k];
  F.a0 = A[2 * NF3 + k];
  F.a1 = A[3 * NF3 + k];
  F.e = A[4 * NF3 + k];
  F.mu = A[5 * NF3 + k];
  F.mv = A[6 * NF3 + k];
  F.mw = A[7 * NF3 + k];
  F.us = A[8 * NF3 + k];
  return F;
}

G5D int
aslot(int plane)
{
  return (plane + 3) % 3;
}

// Youngs normal of tile cell (gi,gj) in plane P from the alpha ring (slots of P-1, P, P+1)
G5D void
youngs_at(double const * sA, int P, int gi, int gj, double & nx, double & ny, double & nz)
{
  double a27[27];
#pragma unroll
  for (int dk = -1; dk <= 1; ++dk)
  {
    double const * plane = sA + aslot(P + dk) * NT3;
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int di = -1; di <= 1; ++di)
        a27[(di + 1) + 3 * (dj + 1) + 9 * (dk + 1)] = plane[tidx3(gi + di, gj + dj)];
  }
  youngs3(a27, nx, ny, nz);
}

G5D Prim
convert_at(SolveDesc const & S, double const * __restrict__ U, BlockDesc const & bd, int gi, int gj, int gk)
{
  int           neg;
  const int64_t off = resolve_cell<3, B3>(S, bd.bx, bd.by, bd.bz, gi, gj, gk, neg);
  const Cons    u = load_cons<3, B3>(U, off, neg);
  Prim          q;
  cons_to_prim<3>(S, u, q);
  return q;
}

} // namespace

G5D int
logical_block3(StepArgs const & A)
{
  const int w = blockIdx.x;
  return (w & 7) * A.per_xcd + (w >> 3);
}

// NTH = 320: waves 0-3 own the interior columns, wave 4 owns the ring columns and converts the
//            outer cells (capped at 256 registers per lane by the 5-wave workgroup);
// NTH = 256: every thread owns an interior column; threads 0..63 also own a ring column and
//            threads 64..143 convert one outer cell (one wave per SIMD, up to 512 registers).
template <bool THINC, bool HANCOCK, int NTH>
__global__ void __launch_bounds__(NTH)
k_step3d(StepArgs A)
{
  __shared__ double sQ[7 * NT3];
  __shared__ double sA[THINC ? 3 * NT3 : 1];
  __shared__ double sE[18 * NF3];
  __shared__ double sRed[NTH / 64];

  const int t = threadIdx.x;
  const int blk = logical_block3(A);
  if (blk >= A.nblocks)
    return;
  const BlockDesc bd = A.blocks[blk];
  const SolveDesc S = A.solves[bd.solve];
  SolveCtl *      C = A.ctl + bd.solve;
  constexpr int   NC = 4096;
  const int64_t   blk_off = static_cast<int64_t>(blk) * 9 * NC;
  const bool      host_dt = A.dt_fixed > 0.0;
  if (!host_dt and C->active == 0)
  {
    for (int c = t; c < NC; c += NTH)
      for (int v = 0; v < 9; ++v)
        A.U2[blk_off + v * NC + c] = A.U[blk_off + v * NC + c];
    return;
  }
  const double dt = host_dt ? A.dt_fixed : C->dt;
  const double dtdx = dt / S.dx;
  const double st = S.slope_type;

  // work items of this thread
  const bool has_int = t < 256;
  const bool has_ring = NTH == 320 ? t >= 256 : t < 64;
  const int  r = NTH == 320 ? t - 256 : t; // ring index when has_ring
  const int  rs = (r >> 4) & 3, rr = r & 15;
  const int  n_out = NTH == 320 ? (t >= 256 ? (r < 16 ? 2 : 1) : 0) : (t >= 64 and t < 144 ? 1 : 0);
  const int  i = t & 15, j = (t >> 4) & 15; // interior column (valid when has_int)
  int        rgi = 0, rgj = 0, ogi0 = 0, ogj0 = 0, ogi1 = 0, ogj1 = 0;
  if (has_ring)
    ring_cell3(r, rgi, rgj);
  if (n_out >= 1)
    outer_cell3(NTH == 320 ? r : t - 64, ogi0, ogj0);
  if (n_out >= 2)
    outer_cell3(r + 64, ogi1, ogj1);

  double * XL = sE;
  double * YL = sE + 7 * NF3;
  double * FX = sE;
  double * FY = sE + 9 * NF3;

  // plane history of the owned interior column (q*) and ring column (r*)
  Prim qm, q0, qp, rm, r0, rp;
  Prim qo0, qo1; // outer cells of the plane being converted
  // interior: carried state of plane P-1
  Cons   u2prev;
  double uaprev0 = 0.0, uaprev1 = 0.0;
  Flux   zFprev;
  Face   zL; // +z face state of plane P-1
  double inv_max = 0.0;

  auto convert_plane = [&](int P, Prim & qi, Prim & qr) {
    if (has_int)
      qi = convert_at(S, A.U, bd, i, j, P);
    if (has_ring)
      qr = convert_at(S, A.U, bd, rgi, rgj, P);
    if (n_out >= 1)
      qo0 = convert_at(S, A.U, bd, ogi0, ogj0, P);
    if (n_out >= 2)
      qo1 = convert_at(S, A.U, bd, ogi1, ogj1, P);
    if (THINC)
    {
      double * plane = sA + aslot(P) * NT3;
      if (has_int)
        plane[tidx3(i, j)] = qi.a;
      if (has_ring)
        plane[tidx3(rgi, rgj)] = qr.a;
      if (n_out >= 1)
        plane[tidx3(ogi0, ogj0)] = qo0.a;
      if (n_out >= 2)
        plane[tidx3(ogi1, ogj1)] = qo1.a;
    }
  };
  auto store_plane = [&](Prim const & qi, Prim const & qr) {
    if (has_int)
      sq_store3(sQ, tidx3(i, j), qi);
    if (has_ring)
      sq_store3(sQ, tidx3(rgi, rgj), qr);
    if (n_out >= 1)
      sq_store3(sQ, tidx3(ogi0, ogj0), qo0);
    if (n_out >= 2)
      sq_store3(sQ, tidx3(ogi1, ogj1), qo1);
  };

  // ---- prologue: planes -2 and -1 ----
  convert_plane(-2, qm, rm);
  convert_plane(-1, q0, r0);
  store_plane(q0, r0);

  for (int P = -1; P <= 16; ++P)
  {
    const bool inner_plane = P >= 0 and P <= 15;
    // (a) convert plane P+1
    convert_plane(P + 1, qp, rp);
    __syncthreads(); // (b)

    // (c) predictor of plane P; z face P and the completion of plane P-1 (all column-local)
    Face fxm, fym, fring;
    if (has_int)
    {
      const Prim  qxm = sq_load3(sQ, tidx3(i - 1, j));
      const Prim  qxp = sq_load3(sQ, tidx3(i + 1, j));
      const Prim  qym = sq_load3(sQ, tidx3(i, j - 1));
      const Prim  qyp = sq_load3(sQ, tidx3(i, j + 1));
      const Slope sx = slopes_of(q0, qxp, qxm, st);
      const Slope sy = slopes_of(q0, qyp, qym, st);
      const Slope sz = slopes_of(q0, qp, qm, st);
      Hancock     h = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
      if (HANCOCK)
        h = hancock_terms<3>(S, q0, sx, sy, sz, dtdx);
      Face fzm, fzp;
      if (!THINC)
      {
        fzm = face_muscl<3, 2, HANCOCK>(S, q0, sx, sy, sz, h, -0.5);
        fzp = face_muscl<3, 2, HANCOCK>(S, q0, sx, sy, sz, h, 0.5);
        if (inner_plane)
        {
          fxm = face_muscl<3, 0, HANCOCK>(S, q0, sx, sy, sz, h, -0.5);
          fym = face_muscl<3, 1, HANCOCK>(S, q0, sx, sy, sz, h, -0.5);
          st_store3(XL, fx3(i + 1, j), face_muscl<3, 0, HANCOCK>(S, q0, sx, sy, sz, h, 0.5));
          st_store3(YL, fy3(i, j + 1), face_muscl<3, 1, HANCOCK>(S, q0, sx, sy, sz, h, 0.5));
        }
      }
      else
      {
        double nx = 0.0, ny = 0.0, nz = 0.0;
        if (q0.a > S.thinc_eps and q0.a < 1.0 - S.thinc_eps)
          youngs_at(sA, P, i, j, nx, ny, nz);
        const ThincAxis tz = thinc_axis(S, q0.a, qm.a, q0.a, qp.a, fabs(nz));
        fzm = face_thinc<3, 2, HANCOCK>(q0, sx, sy, sz, h, tz, -1);
        fzp = face_thinc<3, 2, HANCOCK>(q0, sx, sy, sz, h, tz, 1);
        if (inner_plane)
        {
          const ThincAxis tx = thinc_axis(S, q0.a, qxm.a, q0.a, qxp.a, fabs(nx));
          const ThincAxis ty = thinc_axis(S, q0.a, qym.a, q0.a, qyp.a, fabs(ny));
          fxm = face_thinc<3, 0, HANCOCK>(q0, sx, sy, sz, h, tx, -1);
          fym = face_thinc<3, 1, HANCOCK>(q0, sx, sy, sz, h, ty, -1);
          st_store3(XL, fx3(i + 1, j), face_thinc<3, 0, HANCOCK>(q0, sx, sy, sz, h, tx, 1));
          st_store3(YL, fy3(i, j + 1), face_thinc<3, 1, HANCOCK>(q0, sx, sy, sz, h, ty, 1));
        }
      }
      if (P >= 0)
      {
        // z face P between planes P-1 and P
        const Flux Fz = hllc<3>(S, zL, fzm, dtdx, 2);
        if (P >= 1)
        {
          // plane P-1 gets its z update, normalisation, store and CFL
          apply_flux_pair<3>(u2prev, zFprev, Fz, uaprev0, uaprev1);
          fix_volume_fractions(u2prev);
          store_cons<3, B3>(A.U2, blk_off + i + 16 * j + 256 * (P - 1), u2prev);
          inv_max = fmax(inv_max, cell_inv_dt<3>(S, u2prev));
        }
        zFprev = Fz;
      }
      zL = fzp;
    }
    if (has_ring and inner_plane)
    {
      const Prim  qxm = sq_load3(sQ, tidx3(rgi - 1, rgj));
      const Prim  qxp = sq_load3(sQ, tidx3(rgi + 1, rgj));
      const Prim  qym = sq_load3(sQ, tidx3(rgi, rgj - 1));
      const Prim  qyp = sq_load3(sQ, tidx3(rgi, rgj + 1));
      const Slope sx = slopes_of(r0, qxp, qxm, st);
      const Slope sy = slopes_of(r0, qyp, qym, st);
      const Slope sz = slopes_of(r0, rp, rm, st);
      Hancock     h = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
      if (HANCOCK)
        h = hancock_terms<3>(S, r0, sx, sy, sz, dtdx);
      if (!THINC)
      {
        const double o = (rs & 1) ? -0.5 : 0.5;
        fring = rs < 2 ? face_muscl<3, 0, HANCOCK>(S, r0, sx, sy, sz, h, o)
                       : face_muscl<3, 1, HANCOCK>(S, r0, sx, sy, sz, h, o);
      }
      else
      {
        double nx = 0.0, ny = 0.0, nz = 0.0;
        if (r0.a > S.thinc_eps and r0.a < 1.0 - S.thinc_eps)
          youngs_at(sA, P, rgi, rgj, nx, ny, nz);
        const int ib = (rs & 1) ? -1 : 1;
        if (rs < 2)
        {
          const ThincAxis tx = thinc_axis(S, r0.a, qxm.a, r0.a, qxp.a, fabs(nx));
          fring = face_thinc<3, 0, HANCOCK>(r0, sx, sy, sz, h, tx, ib);
        }
        else
        {
          const ThincAxis ty = thinc_axis(S, r0.a, qym.a, r0.a, qyp.a, fabs(ny));
          fring = face_thinc<3, 1, HANCOCK>(r0, sx, sy, sz, h, ty, ib);
        }
      }
      if (rs == 0)
        st_store3(XL, fx3(0, rr), fring);
      else if (rs == 2)
        st_store3(YL, fy3(rr, 0), fring);
    }
    __syncthreads(); // (d)

    // (e) Riemann problems on the x and y faces of plane P
    Flux Fx, Fy, Fr;
    if (inner_plane)
    {
      if (has_int)
      {
        Fx = hllc<3>(S, st_load3(XL, fx3(i, j)), fxm, dtdx, 0);
        Fy = hllc<3>(S, st_load3(YL, fy3(i, j)), fym, dtdx, 1);
      }
      if (has_ring and (rs & 1))
      {
        const int  ax = rs == 1 ? 0 : 1;
        const Face L = ax == 0 ? st_load3(XL, fx3(16, rr)) : st_load3(YL, fy3(rr, 16));
        Fr = hllc<3>(S, L, fring, dtdx, ax);
      }
    }
    __syncthreads(); // (f)

    // (g) flux exchange, next plane into LDS
    if (inner_plane)
    {
      if (has_int)
      {
        fl_store3(FX, fx3(i, j), Fx);
        fl_store3(FY, fy3(i, j), Fy);
      }
      if (has_ring and (rs & 1))
      {
        if (rs == 1)
          fl_store3(FX, fx3(16, rr), Fr);
        else
          fl_store3(FY, fy3(rr, 16), Fr);
      }
    }
    store_plane(qp, rp);
    __syncthreads(); // (h)

    // (i) x then y update of plane P
    if (has_int and inner_plane)
    {
      const Cons uc = load_cons<3, B3>(A.U, blk_off + i + 16 * j + 256 * P, -1);
      u2prev = uc;
      apply_flux_pair<3>(u2prev, Fx, fl_load3(FX, fx3(i + 1, j)), uc.a0, uc.a1);
      apply_flux_pair<3>(u2prev, Fy, fl_load3(FY, fy3(i, j + 1)), uc.a0, uc.a1);
      uaprev0 = uc.a0;
      uaprev1 = uc.a1;
    }
    qm = q0;
    q0 = qp;
    rm = r0;
    r0 = rp;
  }

  // CFL of the new state
  const double inv = wave_max(inv_max);
  if ((t & 63) == 0)
    sRed[t >> 6] = inv;
  __syncthreads();
  if (t == 0)
  {
    double m = sRed[0];
    for (int w = 1; w < NTH / 64; ++w)
      m = fmax(m, sRed[w]);
    atomicMax(&C->invdt_bits, static_cast<unsigned long long>(__double_as_longlong(m)));
  }
}

// block shape of the 3-D step: 256 (default) or 320 threads, G5_STEP3D_THREADS overrides
static int
step3d_threads()
{
  static int n = [] {
    const char * e = getenv("G5_STEP3D_THREADS");
    return (e and atoi(e) == 320) ? 320 : 256;
  }();
  return n;
}

template <int NTH>
static void
launch3(StepArgs const & A, bool thinc, bool hancock, hipStream_t stream)
{
  const dim3 grid(8 * A.per_xcd), block(NTH);
  if (thinc and hancock)
    hipLaunchKernelGGL((k_step3d<true, true, NTH>), grid, block, 0, stream, A);
  else if (thinc)
    hipLaunchKernelGGL((k_step3d<true, false, NTH>), grid, block, 0, stream, A);
  else if (hancock)
    hipLaunchKernelGGL((k_step3d<false, true, NTH>), grid, block, 0, stream, A);
  else
    hipLaunchKernelGGL((k_step3d<false, false, NTH>), grid, block, 0, stream, A);
}

void
launch_step3d(StepArgs const & A, bool thinc, bool hancock, hipStream_t stream)
{
  if (step3d_threads() == 320)
    launch3<320>(A, thinc, hancock, stream);
  else
    launch3<256>(A, thinc, hancock, stream);
}

} // namespace g5
