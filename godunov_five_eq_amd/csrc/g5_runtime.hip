// g5_runtime.hip -- the C ABI (include/godunov_five_eq.h): context, state transfer, step
// orchestration (the GodunovImplemV0::do_time_step / godunov_unsplit surface), device-resident
// time loop.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "g5_kernels.h"
#include "g5_types.h"

using namespace g5;

struct g5_ctx
{
  g5_config   cfg{};
  SolveDesc   sd{};
  int         nblocks = 0;
  int         nvar = 0;
  int         ncell = 0;
  size_t      nstate = 0; // doubles per state copy
  double *    dU[2] = { nullptr, nullptr };
  double *    dU3 = nullptr; // RK2_SSP stage buffer
  int         cur = 0;
  BlockDesc * d_blocks = nullptr;
  SolveDesc * d_solve = nullptr;
  SolveCtl *  d_ctl = nullptr;
  SolveCtl *  h_ctl = nullptr; // pinned
  int32_t *   d_outside = nullptr;
  double *    d_ring = nullptr;
  std::vector<int32_t> outside;
  hipStream_t stream = nullptr;
  bool        has_state = false;
  int         per_xcd = 0;
  double      last_run_ms = 0.0; // HIP-event time of the last g5_run step loop
  std::string err;
};

namespace
{

std::string g_create_error;

int
fail(g5_ctx * ctx, int code, std::string const & msg)
{
  if (ctx)
    ctx->err = msg;
  else
    g_create_error = msg;
  return code;
}

#define G5_HIP(ctx, call)                                                                          \
  do                                                                                               \
  {                                                                                                \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return fail(ctx, G5_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));            \
  } while (0)

int
validate(g5_config const & c, std::string & msg)
{
  if (c.dim != 2 and c.dim != 3)
    return msg = "dim must be 2 or 3", G5_ERR_INVALID;
  if (c.bsize != 16)
    return msg = "the HIP step kernels take bsize = 16", G5_ERR_UNSUPPORTED;
  for (int d = 0; d < c.dim; ++d)
  {
    if (c.nb[d] < 1)
      return msg = "nb must be >= 1", G5_ERR_INVALID;
    if (not c.periodic[d])
    {
      for (int s = 0; s < 2; ++s)
      {
        const int b = c.bc[2 * d + s];
        if (b != G5_BC_ZERO_GRADIENT and b != G5_BC_WALL)
          return msg = "non-periodic directions need ZERO_GRADIENT or WALL borders", G5_ERR_INVALID;
      }
    }
  }
  if (not(c.dx > 0.0) or not(c.cfl > 0.0))
    return msg = "dx and cfl must be > 0", G5_ERR_INVALID;
  if (not(c.gamma[0] > 1.0) or not(c.gamma[1] > 1.0))
    return msg = "gamma must be > 1", G5_ERR_INVALID;
  if (c.integrator != G5_TI_HANCOCK and c.integrator != G5_TI_RK1 and c.integrator != G5_TI_RK2_SSP)
    return msg = "unknown time integrator", G5_ERR_INVALID;
  if (c.precision != G5_PREC_FP64 and c.precision != G5_PREC_MIXED)
    return msg = "unknown precision", G5_ERR_INVALID;
  if (c.precision == G5_PREC_MIXED)
    return msg = "mixed precision is not available in this build", G5_ERR_UNSUPPORTED;
  return G5_OK;
}

void
launch_step(g5_ctx * ctx, double const * U, double * U2, double dt_fixed, bool hancock)
{
  StepArgs A;
  A.U = U;
  A.U2 = U2;
  A.blocks = ctx->d_blocks;
  A.solves = ctx->d_solve;
  A.ctl = ctx->d_ctl;
  A.nblocks = ctx->nblocks;
  A.per_xcd = ctx->per_xcd;
  A.dt_fixed = dt_fixed;
  if (ctx->cfg.dim == 2)
    launch_step2d(A, ctx->cfg.thinc_enabled != 0, hancock, ctx->stream);
  else
    launch_step3d(A, ctx->cfg.thinc_enabled != 0, hancock, ctx->stream);
}

// one godunov_unsplit iteration with a host dt
int
host_step(g5_ctx * ctx, double dt)
{
  const int c = ctx->cur;
  if (ctx->cfg.integrator == G5_TI_RK2_SSP)
  {
    // stage 1: U_RK = U + dt L(U); stage 2: U2 = U_RK + dt L(U_RK); U = 1/2 U + 1/2 U2
    launch_step(ctx, ctx->dU[c], ctx->dU3, dt, false);
    launch_step(ctx, ctx->dU3, ctx->dU[1 - c], dt, false);
    launch_lincomb(0.5, ctx->dU[c], 0.5, ctx->dU[1 - c], ctx->nstate, ctx->stream);
  }
  else
  {
    launch_step(ctx, ctx->dU[c], ctx->dU[1 - c], dt, ctx->cfg.integrator == G5_TI_HANCOCK);
  }
  ctx->cur = 1 - c;
  G5_HIP(ctx, hipGetLastError());
  return G5_OK;
}

} // namespace

extern "C"
{

const char *
g5_version(void)
{
  return "godunov_five_eq_amd 0.1 (gfx950, FP64 fused block step)";
}

const char *
g5_last_error(const g5_ctx * ctx)
{
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int
g5_create(const g5_config * cfg, g5_ctx ** out)
{
  if (not cfg or not out)
    return fail(nullptr, G5_ERR_INVALID, "null argument");
  *out = nullptr;
  std::string msg;
  const int   rc = validate(*cfg, msg);
  if (rc != G5_OK)
    return fail(nullptr, rc, msg);
  g5_ctx * ctx = new g5_ctx();
  ctx->cfg = *cfg;
  for (int d = cfg->dim; d < 3; ++d)
    ctx->cfg.nb[d] = 1;
  make_solve_desc(ctx->cfg, 0, ctx->sd);
  const int nx = ctx->cfg.nb[0], ny = ctx->cfg.nb[1], nz = ctx->cfg.nb[2];
  ctx->nblocks = nx * ny * nz;
  ctx->nvar = cfg->dim == 2 ? 8 : 9;
  ctx->ncell = cfg->dim == 2 ? 256 : 4096;
  ctx->nstate = static_cast<size_t>(ctx->nblocks) * ctx->nvar * ctx->ncell;
  ctx->per_xcd = (ctx->nblocks + 7) / 8;

  std::vector<BlockDesc> blocks(ctx->nblocks);
  for (int z = 0, b = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x, ++b)
        blocks[b] = BlockDesc{ 0, x, y, z };
  // outside ring: blocks leaving the brick along at least one non-periodic direction
  const int zlo = cfg->dim == 3 ? -1 : 0, zhi = cfg->dim == 3 ? nz : 0;
  for (int z = zlo; z <= zhi; ++z)
    for (int y = -1; y <= ny; ++y)
      for (int x = -1; x <= nx; ++x)
      {
        const int c[3] = { x, y, z };
        bool      outside = false, invalid = false;
        for (int d = 0; d < cfg->dim; ++d)
          if (c[d] < 0 or c[d] >= ctx->cfg.nb[d])
            (ctx->cfg.periodic[d] ? invalid : outside) = true;
        if (outside and not invalid)
        {
          ctx->outside.push_back(x);
          ctx->outside.push_back(y);
          ctx->outside.push_back(z);
        }
      }

  auto cleanup = [&](int code, std::string const & m) {
    g5_destroy(ctx);
    return fail(nullptr, code, m);
  };
  hipError_t e = hipStreamCreate(&ctx->stream);
  if (e != hipSuccess)
    return cleanup(G5_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  for (int k = 0; k < 2; ++k)
  {
    e = hipMalloc(&ctx->dU[k], ctx->nstate * sizeof(double));
    if (e != hipSuccess)
      return cleanup(G5_ERR_NOMEM, std::string("hipMalloc state: ") + hipGetErrorString(e));
  }
  if (cfg->integrator == G5_TI_RK2_SSP)
  {
    e = hipMalloc(&ctx->dU3, ctx->nstate * sizeof(double));
    if (e != hipSuccess)
      return cleanup(G5_ERR_NOMEM, std::string("hipMalloc stage: ") + hipGetErrorString(e));
  }
  if ((e = hipMalloc(&ctx->d_blocks, blocks.size() * sizeof(BlockDesc))) != hipSuccess or
      (e = hipMalloc(&ctx->d_solve, sizeof(SolveDesc))) != hipSuccess or
      (e = hipMalloc(&ctx->d_ctl, sizeof(SolveCtl))) != hipSuccess or
      (e = hipHostMalloc(&ctx->h_ctl, sizeof(SolveCtl), hipHostMallocDefault)) != hipSuccess)
    return cleanup(G5_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  if (not ctx->outside.empty())
  {
    if ((e = hipMalloc(&ctx->d_outside, ctx->outside.size() * sizeof(int32_t))) != hipSuccess)
      return cleanup(G5_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    hipMemcpy(ctx->d_outside, ctx->outside.data(), ctx->outside.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  }
  hipMemcpy(ctx->d_blocks, blocks.data(), blocks.size() * sizeof(BlockDesc), hipMemcpyHostToDevice);
  hipMemcpy(ctx->d_solve, &ctx->sd, sizeof(SolveDesc), hipMemcpyHostToDevice);
  SolveCtl c0{};
  hipMemcpy(ctx->d_ctl, &c0, sizeof(SolveCtl), hipMemcpyHostToDevice);
  e = hipGetLastError();
  if (e != hipSuccess)
    return cleanup(G5_ERR_HIP, std::string("setup: ") + hipGetErrorString(e));
  *out = ctx;
  return G5_OK;
}

int
g5_destroy(g5_ctx * ctx)
{
  if (not ctx)
    return G5_OK;
  if (ctx->stream)
    hipStreamSynchronize(ctx->stream);
  hipFree(ctx->dU[0]);
  hipFree(ctx->dU[1]);
  hipFree(ctx->dU3);
  hipFree(ctx->d_blocks);
  hipFree(ctx->d_solve);
  hipFree(ctx->d_ctl);
  hipFree(ctx->d_outside);
  hipFree(ctx->d_ring);
  if (ctx->h_ctl)
    hipHostFree(ctx->h_ctl);
  if (ctx->stream)
    hipStreamDestroy(ctx->stream);
  delete ctx;
  return G5_OK;
}

int
g5_shape(const g5_ctx * ctx, int32_t * n_blocks, int32_t * nvar, int32_t * cells_per_block)
{
  if (not ctx)
    return G5_ERR_INVALID;
  if (n_blocks)
    *n_blocks = ctx->nblocks;
  if (nvar)
    *nvar = ctx->nvar;
  if (cells_per_block)
    *cells_per_block = ctx->ncell;
  return G5_OK;
}

int
g5_mem_bytes(const g5_ctx * ctx, uint64_t * bytes)
{
  if (not ctx or not bytes)
    return G5_ERR_INVALID;
  uint64_t b = 2ull * ctx->nstate * sizeof(double);
  if (ctx->dU3)
    b += ctx->nstate * sizeof(double);
  b += ctx->nblocks * sizeof(BlockDesc) + sizeof(SolveDesc) + sizeof(SolveCtl);
  b += ctx->outside.size() * sizeof(int32_t);
  if (ctx->d_ring)
    b += ctx->outside.size() / 3 * ctx->nvar * ctx->ncell * sizeof(double);
  *bytes = b;
  return G5_OK;
}

int
g5_upload_state(g5_ctx * ctx, const double * U_host, size_t n)
{
  if (not ctx or not U_host)
    return fail(ctx, G5_ERR_INVALID, "null argument");
  if (n != ctx->nstate)
    return fail(ctx, G5_ERR_INVALID, "state size mismatch");
  G5_HIP(ctx, hipMemcpyAsync(ctx->dU[ctx->cur], U_host, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  G5_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->has_state = true;
  return G5_OK;
}

int
g5_download_state(g5_ctx * ctx, double * U_host, size_t n)
{
  if (not ctx or not U_host)
    return fail(ctx, G5_ERR_INVALID, "null argument");
  if (n != ctx->nstate)
    return fail(ctx, G5_ERR_INVALID, "state size mismatch");
  if (not ctx->has_state)
    return fail(ctx, G5_ERR_STATE, "no state uploaded");
  G5_HIP(ctx, hipMemcpyAsync(U_host, ctx->dU[ctx->cur], n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  G5_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return G5_OK;
}

int
g5_state_device_ptr(g5_ctx * ctx, double ** U_dev)
{
  if (not ctx or not U_dev)
    return fail(ctx, G5_ERR_INVALID, "null argument");
  *U_dev = ctx->dU[ctx->cur];
  return G5_OK;
}

int
g5_compute_dt(g5_ctx * ctx, double * dt)
{
  if (not ctx or not dt)
    return fail(ctx, G5_ERR_INVALID, "null argument");
  if (not ctx->has_state)
    return fail(ctx, G5_ERR_STATE, "no state uploaded");
  launch_reset_invdt(ctx->d_ctl, 1, ctx->stream);
  launch_inv_dt(ctx->cfg.dim, ctx->dU[ctx->cur], ctx->d_blocks, ctx->d_solve, ctx->d_ctl, ctx->nblocks, ctx->stream);
  G5_HIP(ctx, hipGetLastError());
  G5_HIP(ctx, hipMemcpyAsync(ctx->h_ctl, ctx->d_ctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, ctx->stream));
  G5_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double inv;
  std::memcpy(&inv, &ctx->h_ctl->invdt_bits, sizeof(double));
  *dt = ctx->cfg.cfl / inv;
  return G5_OK;
}

int
g5_fill_outside(g5_ctx * ctx, double * out_host, size_t n)
{
  if (not ctx)
    return fail(ctx, G5_ERR_INVALID, "null argument");
  if (not ctx->has_state)
    return fail(ctx, G5_ERR_STATE, "no state uploaded");
  const int    n_out = static_cast<int>(ctx->outside.size() / 3);
  const size_t need = static_cast<size_t>(n_out) * ctx->nvar * ctx->ncell;
  if (out_host and n != need)
    return fail(ctx, G5_ERR_INVALID, "ring size mismatch");
  if (n_out == 0)
    return G5_OK;
  if (not ctx->d_ring)
    G5_HIP(ctx, hipMalloc(&ctx->d_ring, need * sizeof(double)));
  launch_fill_outside(ctx->cfg.dim, ctx->dU[ctx->cur], ctx->d_solve, ctx->d_outside, n_out, ctx->d_ring, ctx->stream);
  G5_HIP(ctx, hipGetLastError());
  if (out_host)
    G5_HIP(ctx, hipMemcpyAsync(out_host, ctx->d_ring, need * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  G5_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return G5_OK;
}

int
g5_outside_coords(const g5_ctx * ctx, int32_t * n_outside, int32_t * coords, size_t ncoords)
{
  if (not ctx or not n_outside)
    return G5_ERR_INVALID;
  *n_outside = static_cast<int32_t>(ctx->outside.size() / 3);
  if (coords)
  {
    if (ncoords != ctx->outside.size())
      return G5_ERR_INVALID;
    std::memcpy(coords, ctx->outside.data(), ncoords * sizeof(int32_t));
  }
  return G5_OK;
}

int
g5_do_time_step(g5_ctx * ctx, double dt)
{
  if (not ctx)
    return fail(ctx, G5_ERR_INVALID, "null argument");
  if (not ctx->has_state)
    return fail(ctx, G5_ERR_STATE, "no state uploaded");
  if (not(dt > 0.0))
    return fail(ctx, G5_ERR_INVALID, "dt must be > 0");
  return host_step(ctx, dt);
}

int
g5_run(g5_ctx * ctx, int32_t nsteps, double * t, double tEnd, int32_t * steps_done)
{
  if (not ctx or not t)
    return fail(ctx, G5_ERR_INVALID, "null argument");
  if (not ctx->has_state)
    return fail(ctx, G5_ERR_STATE, "no state uploaded");
  if (nsteps < 0)
    return fail(ctx, G5_ERR_INVALID, "nsteps < 0");
  if (ctx->cfg.integrator == G5_TI_RK2_SSP)
  {
    // host-driven: the stage structure needs the dt of the step before the first stage
    int n = 0;
    for (; n < nsteps and *t < tEnd; ++n)
    {
      double dt;
      int    rc = g5_compute_dt(ctx, &dt);
      if (rc != G5_OK)
        return rc;
      if (*t + dt > tEnd)
        dt = tEnd - *t;
      rc = host_step(ctx, dt);
      if (rc != G5_OK)
        return rc;
      *t = *t + dt;
    }
    if (steps_done)
      *steps_done = n;
    return g5_synchronize(ctx);
  }
  SolveCtl c0{};
  c0.t = *t;
  c0.tend = tEnd;
  c0.max_steps = nsteps;
  c0.active = 1;
  G5_HIP(ctx, hipMemcpyAsync(ctx->d_ctl, &c0, sizeof(SolveCtl), hipMemcpyHostToDevice, ctx->stream));
  launch_inv_dt(ctx->cfg.dim, ctx->dU[ctx->cur], ctx->d_blocks, ctx->d_solve, ctx->d_ctl, ctx->nblocks, ctx->stream);
  launch_start(ctx->d_ctl, ctx->d_solve, 1, ctx->stream);
  const bool hancock = ctx->cfg.integrator == G5_TI_HANCOCK;
  hipEvent_t e0, e1;
  G5_HIP(ctx, hipEventCreate(&e0));
  G5_HIP(ctx, hipEventCreate(&e1));
  G5_HIP(ctx, hipEventRecord(e0, ctx->stream));
  for (int n = 0; n < nsteps; ++n)
  {
    launch_step(ctx, ctx->dU[ctx->cur], ctx->dU[1 - ctx->cur], 0.0, hancock);
    launch_finalize(ctx->d_ctl, ctx->d_solve, 1, ctx->stream);
    ctx->cur = 1 - ctx->cur;
  }
  G5_HIP(ctx, hipEventRecord(e1, ctx->stream));
  G5_HIP(ctx, hipGetLastError());
  G5_HIP(ctx, hipEventSynchronize(e1));
  float ms = 0.f;
  G5_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  ctx->last_run_ms = ms;
  G5_HIP(ctx, hipMemcpyAsync(ctx->h_ctl, ctx->d_ctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, ctx->stream));
  G5_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *t = ctx->h_ctl->t;
  if (steps_done)
    *steps_done = ctx->h_ctl->steps;
  return G5_OK;
}

int
g5_last_run_ms(const g5_ctx * ctx, double * ms)
{
  if (not ctx or not ms)
    return G5_ERR_INVALID;
  *ms = ctx->last_run_ms;
  return G5_OK;
}

int
g5_synchronize(g5_ctx * ctx)
{
  if (not ctx)
    return fail(ctx, G5_ERR_INVALID, "null argument");
  G5_HIP(ctx, hipStreamSynchronize(ctx->stream));
  G5_HIP(ctx, hipGetLastError());
  return G5_OK;
}

int
g5_time_step_kernel(g5_ctx * ctx, int32_t nsteps, double dt, double * mean_ms)
{
  if (not ctx or not mean_ms)
    return fail(ctx, G5_ERR_INVALID, "null argument");
  if (not ctx->has_state)
    return fail(ctx, G5_ERR_STATE, "no state uploaded");
  if (nsteps <= 0 or not(dt > 0.0))
    return fail(ctx, G5_ERR_INVALID, "nsteps and dt must be > 0");
  hipEvent_t e0, e1;
  G5_HIP(ctx, hipEventCreate(&e0));
  G5_HIP(ctx, hipEventCreate(&e1));
  const bool hancock = ctx->cfg.integrator == G5_TI_HANCOCK;
  G5_HIP(ctx, hipEventRecord(e0, ctx->stream));
  for (int n = 0; n < nsteps; ++n)
  {
    launch_step(ctx, ctx->dU[ctx->cur], ctx->dU[1 - ctx->cur], dt, hancock);
    ctx->cur = 1 - ctx->cur;
  }
  G5_HIP(ctx, hipEventRecord(e1, ctx->stream));
  G5_HIP(ctx, hipEventSynchronize(e1));
  float ms = 0.f;
  G5_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *mean_ms = static_cast<double>(ms) / nsteps;
  return G5_OK;
}

} // extern "C"
