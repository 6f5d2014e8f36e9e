// g5_refine.hip -- refine flags of the blocks of a solve (ComputeRefineFlags,
// scheme/ComputeRefineFlags.cpp:18-372): the Lohner indicator (split or unsplit) or the simple
// normalised gradient of the partial densities IAD0 and IAD1 per cell, the flag vote of each
// indicator value (update_flag, :236-249) and the maximum over the block.
//
// One workgroup of 256 threads per block of the launch list (2-D: a cell per thread; 3-D: a
// column of 16 cells per thread).  The kernel reads only the two partial densities of the block
// and of its face neighbours' edge cells: 2 x 8 B per cell, plus the halo, from HBM.  The votes
// and the largest indicator value are reduced over each wavefront with shuffles and over the
// workgroup in LDS; one atomicMax per block writes each.
#include "g5_device.h"
#include "g5_kernels.h"

namespace g5
{

namespace
{

constexpr int RB = 16;

template <int DIM>
G5D double
var_at(double const * __restrict__ U, int slot, int var, int ci, int cj, int ck)
{
  constexpr int NC = DIM == 2 ? RB * RB : RB * RB * RB;
  constexpr int NV = DIM == 2 ? 8 : 9;
  return U[(static_cast<int64_t>(slot) * NV + var) * NC + ci + RB * (cj + RB * ck)];
}

// variable var of level-0 cell g inside the brick of a two-level solve: a coarse leaf's value, or
// the average of the 2^DIM fine cells of a refined block (summed x fastest, then y, then z, over
// 2^DIM)
template <int DIM>
G5D double
coarse_value(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U, int gx, int gy,
             int gz, int var)
{
  const int sl = slots[S.base + (gx >> 4) + S.nb[0] * ((gy >> 4) + S.nb[1] * (gz >> 4))];
  if (sl >= 0)
    return var_at<DIM>(U, sl, var, gx & 15, gy & 15, gz & 15);
  double sum = 0.0;
#pragma unroll
  for (int k = 0; k < (1 << DIM); ++k)
  {
    const int    fx = 2 * gx + (k & 1), fy = 2 * gy + ((k >> 1) & 1), fz = DIM == 3 ? 2 * gz + (k >> 2) : 0;
    const int    fs = slots[S.base1 + (fx >> 4) + 2 * S.nb[0] * ((fy >> 4) + 2 * S.nb[1] * (fz >> 4))];
    const double v = var_at<DIM>(U, fs, var, fx & 15, fy & 15, fz & 15);
    sum = k == 0 ? v : sum + v;
  }
  return sum / (DIM == 3 ? 8.0 : 4.0);
}

// getNeighborDataSameLevel (scheme/ComputeRefineFlags.cpp:67-96): variable var of the cell one
// step from cell c of block bd along dir (side -1 or +1).  Outside a non-periodic border: the
// current cell's own value; a same-level or coarser neighbour: its cell value (a fine cell outside
// the fine blocks reads its coarse parent); a finer neighbour: the average of its fine cells.
template <int DIM>
G5D double
neighbour_value(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U,
                BlockDesc const & bd, int const (&c)[3], int dir, int side, int var, double self)
{
  int g[3] = { bd.bx * RB + c[0], bd.by * RB + c[1], DIM == 3 ? bd.bz * RB + c[2] : 0 };
  g[dir] += side;
  if (S.base1 >= 0) // two-level solve; fine blocks lie at least one block inside the brick's border
  {
    if (bd.level == 1)
    {
      const int fs = slots[S.base1 + (g[0] >> 4) + 2 * S.nb[0] * ((g[1] >> 4) + 2 * S.nb[1] * (g[2] >> 4))];
      return fs >= 0 ? var_at<DIM>(U, fs, var, g[0] & 15, g[1] & 15, g[2] & 15)
                     : coarse_value<DIM>(S, slots, U, g[0] >> 1, g[1] >> 1, g[2] >> 1, var);
    }
    const int n = S.nb[dir] * RB;
    if (g[dir] < 0 or g[dir] >= n)
    {
      if (!S.periodic[dir])
        return self;
      g[dir] = g[dir] < 0 ? g[dir] + n : g[dir] - n;
    }
    return coarse_value<DIM>(S, slots, U, g[0], g[1], g[2], var);
  }
  const int n = S.nb[dir] * RB;
  if (g[dir] < 0 or g[dir] >= n)
  {
    if (!S.periodic[dir])
      return self;
    g[dir] = g[dir] < 0 ? g[dir] + n : g[dir] - n;
  }
  const int sl = slots[S.base + (g[0] >> 4) + S.nb[0] * ((g[1] >> 4) + S.nb[1] * (g[2] >> 4))];
  return var_at<DIM>(U, sl, var, g[0] & 15, g[1] & 15, g[2] & 15);
}

// compute_lohner_split (:117-135)
G5D double
lohner_split(double L, double C, double R, double eps)
{
  return fabs(L - 2 * C + R) / (fabs(R - C) + fabs(C - L) + eps * (fabs(L) + 2 * fabs(C) + fabs(R)));
}

// normalized_gradient (:100-113)
G5D double
normalized_gradient(double v1, double v2)
{
  double vmax = fmax(fabs(v1), fabs(v2));
  if (vmax < 0.001)
    return 0.0;
  vmax = fabs(v1 - v2) / vmax;
  return fmax(fmin(vmax, 1.0), 0.0);
}

struct Vote
{
  int    flag = G5_AMR_FLAG_INIT;
  double imax = 0.0; // fmax drops a NaN indicator value (0/0 where a partial density vanishes)
};

// update_flag (:236-249)
G5D void
vote(Vote & v, RefineArgs const & R, int level, double ind)
{
  if (level < R.level_max and ind > R.refine_th)
    v.flag = max(v.flag, static_cast<int>(G5_AMR_REFINE));
  else if (level > R.level_min and ind < R.coarsen_th)
    v.flag = max(v.flag, static_cast<int>(G5_AMR_COARSEN));
  else
    v.flag = max(v.flag, static_cast<int>(G5_AMR_NOTHING));
  v.imax = fmax(v.imax, ind);
}

// the indicator values of one cell and variable (operator() bodies, :253-372)
template <int DIM>
G5D void
cell_votes(SolveDesc const & S, int32_t const * __restrict__ slots, double const * __restrict__ U,
           BlockDesc const & bd, RefineArgs const & R, int level, int const (&c)[3], int var, Vote & v)
{
  const double C = var_at<DIM>(U, bd.slot, var, c[0], c[1], c[2]);
  double       L[DIM], Rv[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d)
  {
    L[d] = neighbour_value<DIM>(S, slots, U, bd, c, d, -1, var, C);
    Rv[d] = neighbour_value<DIM>(S, slots, U, bd, c, d, 1, var, C);
  }
  if (R.indicator == G5_INDICATOR_LOHNER_SPLIT)
  {
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      vote(v, R, level, lohner_split(L[d], C, Rv[d], R.eps));
  }
  else if (R.indicator == G5_INDICATOR_LOHNER_UNSPLIT)
  {
    // compute_lohner_unsplit (:140-191)
    double num = 0.0, den = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d)
    {
      const double s2 = L[d] - 2 * C + Rv[d];
      const double t = fabs(Rv[d] - C) + fabs(C - L[d]) + R.eps * (fabs(L[d]) + 2 * fabs(C) + fabs(Rv[d]));
      num = d == 0 ? s2 * s2 : num + s2 * s2;
      den = d == 0 ? t * t : den + t * t;
    }
    vote(v, R, level, sqrt(num / den));
  }
  else
  {
    // compute_simple_gradient (:196-231), from the identity of a max reduction
    double res = -1.7976931348623157e308;
#pragma unroll
    for (int d = 0; d < DIM; ++d)
    {
      res = fmax(res, normalized_gradient(C, L[d]));
      res = fmax(res, normalized_gradient(C, Rv[d]));
    }
    vote(v, R, level, res);
  }
}

template <int DIM>
__global__ void __launch_bounds__(256)
k_refine_flags(double const * buf0, double const * buf1, BlockDesc const * blocks, SolveDesc const * solves,
               SolveCtl const * ctl, int32_t const * slots, RefineArgs R, int nblocks)
{
  __shared__ int    sFlag[4];
  __shared__ double sMax[4];
  const int blk = blockIdx.x;
  if (blk >= nblocks)
    return;
  const BlockDesc bd = blocks[blk];
  const SolveDesc S = solves[bd.solve];
  double const *  U = ctl[bd.solve].parity ? buf1 : buf0;
  const int       level = R.base_level + bd.level;
  constexpr int   IAD0 = DIM == 2 ? 5 : 6, IAD1 = DIM == 2 ? 7 : 8;
  const int       t = threadIdx.x;
  Vote            v;
  for (int k = 0; k < (DIM == 2 ? 1 : RB); ++k)
  {
    const int c[3] = { t & 15, t >> 4, k };
    cell_votes<DIM>(S, slots, U, bd, R, level, c, IAD0, v);
    cell_votes<DIM>(S, slots, U, bd, R, level, c, IAD1, v);
  }
  int    f = v.flag;
  double m = v.imax;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
  {
    f = max(f, __shfl_xor(f, off, 64));
    m = fmax(m, __shfl_xor(m, off, 64));
  }
  if ((t & 63) == 0)
  {
    sFlag[t >> 6] = f;
    sMax[t >> 6] = m;
  }
  __syncthreads();
  if (t == 0)
  {
    f = max(max(sFlag[0], sFlag[1]), max(sFlag[2], sFlag[3]));
    m = fmax(fmax(sMax[0], sMax[1]), fmax(sMax[2], sMax[3]));
    atomicMax(R.flags + bd.slot, f);
    // non-negative doubles order as their bit patterns
    atomicMax(reinterpret_cast<unsigned long long *>(R.imax) + bd.slot,
              static_cast<unsigned long long>(__double_as_longlong(m)));
  }
}

} // namespace

void
launch_refine_flags(int dim, double const * buf0, double const * buf1, BlockDesc const * blocks,
                    SolveDesc const * solves, SolveCtl const * ctl, int32_t const * slots, RefineArgs const & R,
                    int nblocks, hipStream_t stream)
{
  if (nblocks <= 0)
    return;
  if (dim == 2)
    hipLaunchKernelGGL(k_refine_flags<2>, dim3(nblocks), dim3(256), 0, stream, buf0, buf1, blocks, solves, ctl, slots,
                       R, nblocks);
  else
    hipLaunchKernelGGL(k_refine_flags<3>, dim3(nblocks), dim3(256), 0, stream, buf0, buf1, blocks, solves, ctl, slots,
                       R, nblocks);
}

} // namespace g5
