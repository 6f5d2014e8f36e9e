// g5_aux.hip -- dt reduction, time-step control, outside-ring gather, linear combination.
#include "g5_device.h"
#include "g5_kernels.h"

namespace g5
{

// ComputeDtFiveEqFunctor (scheme/ComputeDtFiveEqFunctor.cpp:45-134): one workgroup per block,
// per-cell CFL rate, wavefront butterfly max (__shfl_xor), LDS max over the 4 wavefronts,
// one atomicMax per block on the bit pattern of the (non-negative) double.
template <int DIM>
__global__ void __launch_bounds__(256)
k_inv_dt(double const * __restrict__ U, BlockDesc const * blocks, SolveDesc const * solves,
         SolveCtl * ctl, int nblocks)
{
  __shared__ double sRed[4];
  const int blk = blockIdx.x;
  if (blk >= nblocks)
    return;
  const BlockDesc bd = blocks[blk];
  const SolveDesc S = solves[bd.solve];
  constexpr int   NC = DIM == 2 ? 256 : 4096;
  constexpr int   NV = DIM == 2 ? 8 : 9;
  double          inv = 0.0;
  for (int c = threadIdx.x; c < NC; c += 256)
  {
    const Cons u = load_cons<DIM, 16>(U, static_cast<int64_t>(blk) * NV * NC + c, -1);
    inv = fmax(inv, cell_inv_dt<DIM>(S, u));
  }
  inv = wave_max(inv);
  if ((threadIdx.x & 63) == 0)
    sRed[threadIdx.x >> 6] = inv;
  __syncthreads();
  if (threadIdx.x == 0)
  {
    inv = fmax(fmax(sRed[0], sRed[1]), fmax(sRed[2], sRed[3]));
    atomicMax(&ctl[bd.solve].invdt_bits, static_cast<unsigned long long>(__double_as_longlong(inv)));
  }
}

void
launch_inv_dt(int dim, double const * U, BlockDesc const * blocks, SolveDesc const * solves,
              SolveCtl * ctl, int nblocks, hipStream_t stream)
{
  if (nblocks <= 0)
    return;
  if (dim == 2)
    hipLaunchKernelGGL(k_inv_dt<2>, dim3(nblocks), dim3(256), 0, stream, U, blocks, solves, ctl, nblocks);
  else
    hipLaunchKernelGGL(k_inv_dt<3>, dim3(nblocks), dim3(256), 0, stream, U, blocks, solves, ctl, nblocks);
}

// SolverGodunovFiveEq::compute_dt_local (SolverGodunovFiveEq.cpp:386): dt = cfl / invDt, then
// clamped so that t + dt does not pass tend.
G5D double
next_dt(SolveCtl & c, SolveDesc const & S)
{
  const double inv = __longlong_as_double(static_cast<long long>(c.invdt_bits));
  double       dt = S.cfl / inv;
  if (c.t + dt > c.tend)
    dt = c.tend - c.t;
  return dt;
}

__global__ void
k_start(SolveCtl * ctl, SolveDesc const * solves, int nsolves)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsolves)
    return;
  SolveCtl & c = ctl[s];
  c.dt = next_dt(c, solves[s]);
  c.invdt_bits = 0ull;
  c.active = (c.t < c.tend and c.steps < c.max_steps) ? 1 : 0;
}

__global__ void
k_finalize(SolveCtl * ctl, SolveDesc const * solves, int nsolves)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsolves)
    return;
  SolveCtl & c = ctl[s];
  if (c.active == 0)
  {
    c.invdt_bits = 0ull;
    return;
  }
  c.t = c.t + c.dt;
  c.steps += 1;
  c.dt = next_dt(c, solves[s]);
  c.invdt_bits = 0ull;
  if (not(c.t < c.tend and c.steps < c.max_steps))
    c.active = 0;
}

__global__ void
k_reset_invdt(SolveCtl * ctl, int nsolves)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsolves)
    ctl[s].invdt_bits = 0ull;
}

void
launch_start(SolveCtl * ctl, SolveDesc const * solves, int nsolves, hipStream_t stream)
{
  hipLaunchKernelGGL(k_start, dim3((nsolves + 255) / 256), dim3(256), 0, stream, ctl, solves, nsolves);
}

void
launch_finalize(SolveCtl * ctl, SolveDesc const * solves, int nsolves, hipStream_t stream)
{
  hipLaunchKernelGGL(k_finalize, dim3((nsolves + 255) / 256), dim3(256), 0, stream, ctl, solves, nsolves);
}

void
launch_reset_invdt(SolveCtl * ctl, int nsolves, hipStream_t stream)
{
  hipLaunchKernelGGL(k_reset_invdt, dim3((nsolves + 255) / 256), dim3(256), 0, stream, ctl, nsolves);
}

// FillOutsideCellFunctor (border_conditions/FillOutside.cpp:69-170) as a gather: every cell of
// every outside block reads its source cell directly (no hash lookup).
template <int DIM>
__global__ void __launch_bounds__(256)
k_fill_outside(double const * __restrict__ U, SolveDesc const * solve, int32_t const * coords, int n_out,
               double * __restrict__ ring)
{
  constexpr int NC = DIM == 2 ? 256 : 4096;
  constexpr int NV = DIM == 2 ? 8 : 9;
  const int     o = blockIdx.x;
  if (o >= n_out)
    return;
  const SolveDesc S = *solve;
  const int       N[3] = { coords[3 * o], coords[3 * o + 1], coords[3 * o + 2] };
  for (int cell = threadIdx.x; cell < NC; cell += 256)
  {
    const int     c[3] = { cell & 15, (cell >> 4) & 15, DIM == 3 ? cell >> 8 : 0 };
    int           neg;
    const int64_t off = resolve_block_cell<DIM, 16>(S, N, c, neg);
    const Cons    u = load_cons<DIM, 16>(U, off, neg);
    store_cons<DIM, 16>(ring, static_cast<int64_t>(o) * NV * NC + cell, u);
  }
}

void
launch_fill_outside(int dim, double const * U, SolveDesc const * solve, int32_t const * coords, int n_out,
                    double * ring, hipStream_t stream)
{
  if (n_out <= 0)
    return;
  if (dim == 2)
    hipLaunchKernelGGL(k_fill_outside<2>, dim3(n_out), dim3(256), 0, stream, U, solve, coords, n_out, ring);
  else
    hipLaunchKernelGGL(k_fill_outside<3>, dim3(n_out), dim3(256), 0, stream, U, solve, coords, n_out, ring);
}

__global__ void
k_lincomb(double a, double const * __restrict__ U, double b, double * __restrict__ U2, size_t n)
{
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<size_t>(gridDim.x) * blockDim.x)
    U2[k] = a * U[k] + b * U2[k];
}

void
launch_lincomb(double a, double const * U, double b, double * U2, size_t n, hipStream_t stream)
{
  size_t nb = (n + 255) / 256;
  if (nb > 65536)
    nb = 65536;
  hipLaunchKernelGGL(k_lincomb, dim3(static_cast<unsigned>(nb)), dim3(256), 0, stream, a, U, b, U2, n);
}

} // namespace g5
