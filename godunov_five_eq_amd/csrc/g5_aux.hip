// g5_aux.hip -- dt reduction, time-step control, state gather/scatter, outside-ring gather,
// RK2 linear combination, device initial conditions.
#include "g5_device.h"
#include "g5_kernels.h"

namespace g5
{

G5D double const *
cur_buf(double * b0, double * b1, int parity)
{
  return parity ? b1 : b0;
}

// ComputeDtFiveEqFunctor (scheme/ComputeDtFiveEqFunctor.cpp:45-134): one workgroup per block,
// per-cell CFL rate, wavefront butterfly max (__shfl_xor), LDS max over the 4 wavefronts,
// one atomicMax per block on the bit pattern of the (non-negative) double.  T selects the FP64
// rate or the mixed-path rate (FP32 sound speed from the FP64-recovered pressure); FAST the rate of
// the reassociated FP64 path.
template <int DIM, typename T, bool FAST = false>
__global__ void __launch_bounds__(256)
k_inv_dt(double * buf0, double * buf1, BlockDesc const * blocks, SolveDesc const * solves, SolveCtl * ctl,
         int nblocks)
{
  __shared__ double sRed[4];
  const int blk = blockIdx.x;
  if (blk >= nblocks)
    return;
  const BlockDesc bd = blocks[blk];
  const SolveDesc S = at_level(solves[bd.solve], bd.level);
  double const *  U = cur_buf(buf0, buf1, ctl[bd.solve].parity);
  constexpr int   NC = DIM == 2 ? 256 : 4096;
  constexpr int   NV = DIM == 2 ? 8 : 9;
  double          inv = 0.0;
  uint32_t        fail = 0u;
  for (int c = threadIdx.x; c < NC; c += 256)
  {
    const Cons   u = load_cons<DIM, 16>(U, static_cast<int64_t>(bd.slot) * NV * NC + c, -1);
    double       p;
    const double rate = cell_rate<DIM, T, FAST>(S, u, p);
    inv = fmax(inv, rate);
    fail |= cell_failure(S, u, rate, p);
  }
  report_failure(&ctl[bd.solve], fail);
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

template <int DIM>
static void
launch_inv_dt_dim(int precision, double * buf0, double * buf1, BlockDesc const * blocks, SolveDesc const * solves,
                  SolveCtl * ctl, int nblocks, hipStream_t stream)
{
  const dim3 grid(nblocks), block(256);
  if (precision == G5_PREC_MIXED)
    hipLaunchKernelGGL((k_inv_dt<DIM, float>), grid, block, 0, stream, buf0, buf1, blocks, solves, ctl, nblocks);
  else if (precision == G5_PREC_FP64_FAST)
    hipLaunchKernelGGL((k_inv_dt<DIM, double, true>), grid, block, 0, stream, buf0, buf1, blocks, solves, ctl,
                       nblocks);
  else
    hipLaunchKernelGGL((k_inv_dt<DIM, double>), grid, block, 0, stream, buf0, buf1, blocks, solves, ctl, nblocks);
}

void
launch_inv_dt(int dim, int precision, double * buf0, double * buf1, BlockDesc const * blocks, SolveDesc const * solves,
              SolveCtl * ctl, int nblocks, hipStream_t stream)
{
  if (nblocks <= 0)
    return;
  if (dim == 2)
    launch_inv_dt_dim<2>(precision, buf0, buf1, blocks, solves, ctl, nblocks, stream);
  else
    launch_inv_dt_dim<3>(precision, buf0, buf1, blocks, solves, ctl, nblocks, stream);
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

// the CFL maximum of the new state gives a finite, positive dt (0 would be an infinite dt)
G5D bool
rate_ok(SolveCtl const & c)
{
  const double inv = __longlong_as_double(static_cast<long long>(c.invdt_bits));
  return inv > 0.0 and inv <= 1.7976931348623157e308;
}

// first dt of the solves in ids[0..n) (all solves when ids is null)
__global__ void
k_start(SolveCtl * ctl, SolveDesc const * solves, int32_t const * ids, int n)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n)
    return;
  const int  s = ids ? ids[k] : k;
  SolveCtl & c = ctl[s];
  if (not rate_ok(c))
    c.failure |= static_cast<uint32_t>(G5_FAIL_NONFINITE);
  c.dt = next_dt(c, solves[s]);
  c.invdt_bits = 0ull;
  c.active = (c.failure == 0u and c.t < c.tend and c.steps < c.max_steps) ? 1 : 0;
}

// after a device-dt step: the active solves advance, flip buffers (unless the step wrote the new
// state over the current one: the device RK2_SSP step) and take their next dt
__global__ void
k_finalize(SolveCtl * ctl, SolveDesc const * solves, int nsolves, int flip)
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
  if (flip)
    c.parity ^= 1;
  // a failed new state (its cells checked by the step kernels) stops the solve where it failed
  if (not rate_ok(c))
    c.failure |= static_cast<uint32_t>(G5_FAIL_NONFINITE);
  if (c.failure != 0u)
  {
    c.active = 0;
    c.invdt_bits = 0ull;
    return;
  }
  c.dt = next_dt(c, solves[s]);
  c.invdt_bits = 0ull;
  if (not(c.t < c.tend and c.steps < c.max_steps))
    c.active = 0;
}

// after a host-dt step of every solve: flip buffers
__global__ void
k_flip(SolveCtl * ctl, int nsolves)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsolves)
  {
    ctl[s].parity ^= 1;
    ctl[s].invdt_bits = 0ull;
  }
}

__global__ void
k_reset_invdt(SolveCtl * ctl, int nsolves)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsolves)
    ctl[s].invdt_bits = 0ull;
}

void
launch_start(SolveCtl * ctl, SolveDesc const * solves, int32_t const * ids, int n, hipStream_t stream)
{
  if (n > 0)
    hipLaunchKernelGGL(k_start, dim3((n + 255) / 256), dim3(256), 0, stream, ctl, solves, ids, n);
}

void
launch_finalize(SolveCtl * ctl, SolveDesc const * solves, int nsolves, hipStream_t stream, bool flip)
{
  hipLaunchKernelGGL(k_finalize, dim3((nsolves + 255) / 256), dim3(256), 0, stream, ctl, solves, nsolves,
                     flip ? 1 : 0);
}

void
launch_flip(SolveCtl * ctl, int nsolves, hipStream_t stream)
{
  hipLaunchKernelGGL(k_flip, dim3((nsolves + 255) / 256), dim3(256), 0, stream, ctl, nsolves);
}

void
launch_reset_invdt(SolveCtl * ctl, int nsolves, hipStream_t stream)
{
  hipLaunchKernelGGL(k_reset_invdt, dim3((nsolves + 255) / 256), dim3(256), 0, stream, ctl, nsolves);
}

// copy the blocks of one solve between their slots (in buffer `parity`) and a contiguous
// [block][var][cell] array
__global__ void __launch_bounds__(256)
k_gather(double * buf0, double * buf1, SolveCtl const * ctl, int solve, int32_t const * slots, int base, int nblk,
         int block_doubles, double * out, int to_slots)
{
  const int b = blockIdx.x;
  if (b >= nblk)
    return;
  double *      U = ctl[solve].parity ? buf1 : buf0;
  const int64_t src = static_cast<int64_t>(slots[base + b]) * block_doubles;
  const int64_t dst = static_cast<int64_t>(b) * block_doubles;
  for (int k = threadIdx.x; k < block_doubles; k += 256)
  {
    if (to_slots)
      U[src + k] = out[dst + k];
    else
      out[dst + k] = U[src + k];
  }
}

void
launch_gather(double * buf0, double * buf1, SolveCtl const * ctl, int solve, int32_t const * slots, int base,
              int nblk, int block_doubles, double * out, bool to_slots, hipStream_t stream)
{
  if (nblk > 0)
    hipLaunchKernelGGL(k_gather, dim3(nblk), dim3(256), 0, stream, buf0, buf1, ctl, solve, slots, base, nblk,
                       block_doubles, out, to_slots ? 1 : 0);
}

// FillOutsideCellFunctor (border_conditions/FillOutside.cpp:69-170) as a gather: every cell of
// every outside block reads its source cell directly (no hash lookup).
template <int DIM>
__global__ void __launch_bounds__(256)
k_fill_outside(double * buf0, double * buf1, SolveCtl const * ctl, int solve, SolveDesc const * solves,
               int32_t const * slots, int32_t const * coords, int n_out, double * __restrict__ ring)
{
  constexpr int NC = DIM == 2 ? 256 : 4096;
  constexpr int NV = DIM == 2 ? 8 : 9;
  const int     o = blockIdx.x;
  if (o >= n_out)
    return;
  const SolveDesc S = solves[solve];
  double const *  U = cur_buf(buf0, buf1, ctl[solve].parity);
  const int       N[3] = { coords[3 * o], coords[3 * o + 1], coords[3 * o + 2] };
  for (int cell = threadIdx.x; cell < NC; cell += 256)
  {
    const int     c[3] = { cell & 15, (cell >> 4) & 15, DIM == 3 ? cell >> 8 : 0 };
    int           neg;
    const int64_t off = resolve_block_cell<DIM, 16>(S, slots, N, c, neg);
    Cons          u = load_cons<DIM, 16>(U, off, -1);
    apply_border<DIM>(S, u, neg);
    store_cons<DIM, 16>(ring, static_cast<int64_t>(o) * NV * NC + cell, u);
  }
}

void
launch_fill_outside(int dim, double * buf0, double * buf1, SolveCtl const * ctl, int solve, SolveDesc const * solves,
                    int32_t const * slots, int32_t const * coords, int n_out, double * ring, hipStream_t stream)
{
  if (n_out <= 0)
    return;
  if (dim == 2)
    hipLaunchKernelGGL(k_fill_outside<2>, dim3(n_out), dim3(256), 0, stream, buf0, buf1, ctl, solve, solves, slots,
                       coords, n_out, ring);
  else
    hipLaunchKernelGGL(k_fill_outside<3>, dim3(n_out), dim3(256), 0, stream, buf0, buf1, ctl, solve, solves, slots,
                       coords, n_out, ring);
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

// end of a device-dt RK2_SSP step, one workgroup per block of the launch list: U = 1/2 U2 + 1/2 U
// in place, U the block's slot in the buffer of its solve's parity and U2 the same slot in the
// other buffer (the same bits as 1/2 U + 1/2 U2: the two products are rounded alike and the sum
// is commutative); nothing for a solve that has finished
__global__ void __launch_bounds__(256)
k_rk2_combine(double * buf0, double * buf1, BlockDesc const * blocks, SolveCtl const * ctl, int nblocks,
              int block_doubles)
{
  const int b = blockIdx.x;
  if (b >= nblocks)
    return;
  const BlockDesc  bd = blocks[b];
  SolveCtl const & c = ctl[bd.solve];
  if (c.active == 0)
    return;
  const int64_t         base = static_cast<int64_t>(bd.slot) * block_doubles;
  double * __restrict__ U = (c.parity ? buf1 : buf0) + base;
  double const * __restrict__ U2 = (c.parity ? buf0 : buf1) + base;
  for (int k = threadIdx.x; k < block_doubles; k += 256)
    U[k] = 0.5 * U2[k] + 0.5 * U[k];
}

void
launch_rk2_combine(double * buf0, double * buf1, BlockDesc const * blocks, SolveCtl const * ctl, int nblocks,
                   int block_doubles, hipStream_t stream)
{
  if (nblocks > 0)
    hipLaunchKernelGGL(k_rk2_combine, dim3(nblocks), dim3(256), 0, stream, buf0, buf1, blocks, ctl, nblocks,
                       block_doubles);
}

// ---------------------------------------------------------------------------------------------
// Initial conditions on the device: InitShockBubbleDataFunctor (init/InitShockBubble.cpp:50-225)
// with a sphere or box inclusion, and InitTriplePointDataFunctor (init/InitTriplePoint.cpp:46-133).
// The sphere fraction of a cell is exact away from the surface and a midpoint quadrature on
// 16^2 / 8^3 sub-cells near it (the same arithmetic as the host generator g5_init_shock_bubble);
// the box fraction is the exact overlap.
// ---------------------------------------------------------------------------------------------
template <int DIM>
G5D double
sphere_fraction_d(double const * lo, double dx, double const * c, double R)
{
  double d2 = 0.0;
  for (int k = 0; k < DIM; ++k)
  {
    const double x = lo[k] + 0.5 * dx - c[k];
    d2 += x * x;
  }
  const double dist = sqrt(d2);
  const double half_diag = 0.5 * dx * sqrt(static_cast<double>(DIM));
  if (dist + half_diag <= R)
    return 1.0;
  if (dist - half_diag >= R)
    return 0.0;
  constexpr int n = DIM == 2 ? 16 : 8;
  int           inside = 0, total = 0;
  for (int kk = 0; kk < (DIM == 3 ? n : 1); ++kk)
    for (int jj = 0; jj < n; ++jj)
      for (int ii = 0; ii < n; ++ii)
      {
        const double p[3] = { lo[0] + (ii + 0.5) * dx / n, lo[1] + (jj + 0.5) * dx / n,
                              DIM == 3 ? lo[2] + (kk + 0.5) * dx / n : 0.0 };
        double       s = 0.0;
        for (int k = 0; k < DIM; ++k)
          s += (p[k] - c[k]) * (p[k] - c[k]);
        inside += s <= R * R ? 1 : 0;
        ++total;
      }
  return static_cast<double>(inside) / total;
}

template <int DIM>
G5D double
box_fraction_d(double const * lo, double dx, double const * c, double const * half)
{
  double f = 1.0;
  for (int k = 0; k < DIM; ++k)
  {
    const double a = fmax(lo[k], c[k] - half[k]);
    const double b = fmin(lo[k] + dx, c[k] + half[k]);
    f *= fmax(0.0, b - a) / dx;
  }
  return f;
}

template <int DIM>
__global__ void __launch_bounds__(256)
k_init_ic(double * buf0, double * buf1, SolveCtl const * ctl, int solve, SolveDesc const * solves,
          int32_t const * slots, IcDesc const * icp)
{
  constexpr int NC = DIM == 2 ? 256 : 4096;
  constexpr int NV = DIM == 2 ? 8 : 9;
  const int     b = blockIdx.x;
  const SolveDesc S = solves[solve];
  const int       nblk = S.nb[0] * S.nb[1] * S.nb[2];
  if (b >= nblk)
    return;
  IcDesc const & ic = *icp;
  double *       U = ctl[solve].parity ? buf1 : buf0;
  const int      bx = b % S.nb[0], by = (b / S.nb[0]) % S.nb[1], bz = b / (S.nb[0] * S.nb[1]);
  const int64_t  base = static_cast<int64_t>(slots[S.base + b]) * NV * NC;
  const int      IA0 = DIM == 2 ? 4 : 5, IAD0 = IA0 + 1, IA1 = IA0 + 2, IAD1 = IA0 + 3;
  const int      IU = 2, IV = 3, IW = 4, IE = 1, ID = 0;
  for (int cell = threadIdx.x; cell < NC; cell += 256)
  {
    const int    i = cell & 15, j = (cell >> 4) & 15, k = DIM == 3 ? cell >> 8 : 0;
    const double lo[3] = { ic.xmin[0] + (bx * 16 + i) * S.dx, ic.xmin[1] + (by * 16 + j) * S.dx,
                           ic.xmin[2] + (bz * 16 + k) * S.dx };
    const double xc = lo[0] + 0.5 * S.dx;
    double       u[9];
    if (ic.kind == 2)
    {
      int r = 0;
      if (xc > ic.x_front)
      {
        const double yc = lo[1] + 0.5 * S.dx;
        const double zc = lo[2] + 0.5 * S.dx;
        const bool   in1 = DIM == 2 ? (yc <= ic.yd) : (sqrt(yc * yc + zc * zc) <= ic.yd);
        r = in1 ? 1 : 2;
      }
      for (int v = 0; v < NV; ++v)
        u[v] = ic.S[r][v];
      u[ID] = u[IAD0] + u[IAD1];
      u[IA1] = 1.0 - u[IA0];
    }
    else
    {
      double const * st = xc <= ic.x_front ? ic.S[0] : ic.S[1];
      for (int v = 0; v < NV; ++v)
        u[v] = st[v];
      u[ID] = u[IAD0] + u[IAD1];
      u[IA1] = 1.0 - u[IA0];
      const double f = ic.kind == 0 ? sphere_fraction_d<DIM>(lo, S.dx, ic.center, ic.radius)
                                    : box_fraction_d<DIM>(lo, S.dx, ic.center, ic.half);
      if (f > 0.0)
      {
        const double phi0 = 1.0 - f;
        u[IAD0] = phi0 * ic.S[1][IAD0] + (1.0 - phi0) * ic.S[2][IAD0];
        u[IAD1] = phi0 * ic.S[1][IAD1] + (1.0 - phi0) * ic.S[2][IAD1];
        u[ID] = u[IAD0] + u[IAD1];
        u[IA0] = phi0 * ic.S[1][IA0] + (1.0 - phi0) * ic.S[2][IA0];
        u[IA1] = 1.0 - u[IA0];
        u[IU] = phi0 * ic.S[1][IU] + (1.0 - phi0) * ic.S[2][IU];
        u[IV] = phi0 * ic.S[1][IV] + (1.0 - phi0) * ic.S[2][IV];
        if (DIM == 3)
          u[IW] = phi0 * ic.S[1][IW] + (1.0 - phi0) * ic.S[2][IW];
        const double rho_mixed = u[IAD0] + u[IAD1];
        double       ekin = u[IU] * u[IU] + u[IV] * u[IV];
        if (DIM == 3)
          ekin += u[IW] * u[IW];
        ekin /= (2.0 * rho_mixed);
        const double eint_mixed = phi0 * ic.eint_pre + (1.0 - phi0) * ic.eint_bub;
        u[IE] = eint_mixed + ekin;
      }
    }
    for (int v = 0; v < NV; ++v)
      U[base + v * NC + cell] = u[v];
  }
}

void
launch_init_ic(int dim, double * buf0, double * buf1, SolveCtl const * ctl, int solve, SolveDesc const * solves,
               int32_t const * slots, IcDesc const * ic, int nblk, hipStream_t stream)
{
  if (nblk <= 0)
    return;
  if (dim == 2)
    hipLaunchKernelGGL(k_init_ic<2>, dim3(nblk), dim3(256), 0, stream, buf0, buf1, ctl, solve, solves, slots, ic);
  else
    hipLaunchKernelGGL(k_init_ic<3>, dim3(nblk), dim3(256), 0, stream, buf0, buf1, ctl, solve, solves, slots, ic);
}

} // namespace g5
