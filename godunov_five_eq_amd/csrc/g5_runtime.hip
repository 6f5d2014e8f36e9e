// g5_runtime.hip -- the C ABI (include/godunov_five_eq.h).
//
// Everything is built on one internal object, the Pool: two state buffers of block slots sized by
// an HBM budget, a table of solves (SolveDesc + device-resident SolveCtl), a slot table per solve
// and the launch list of the blocks of all live solves.  A single solve (g5_ctx, the
// GodunovImplemV0 surface) is a pool of one; g5_pool_* exposes the pool itself for batched and
// served solves that are admitted, stepped together and released independently.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "g5_kernels.h"
#include "g5_types.h"

using namespace g5;

namespace
{

std::string g_create_error;

struct SolveRec
{
  bool                 live = false;
  bool                 started = false; // g5_pool_start ran for it
  g5_config            cfg{};
  int32_t              nblk = 0;
  int32_t              table_base = 0;
  int32_t              table_len = 0; // entries of its slot-table segment
  std::vector<int32_t> slots;
  // two-level static mesh (2-D): (x, y, level) of every block in the caller's block order --
  // the coarse leaves lexicographic, then the fine blocks lexicographic over the fine grid --
  // with slots[k] the slot of block k; border[k]: a fine block with a coarse neighbour.  The
  // slot-table segment holds the level-0 table (nb[0] x nb[1], -1 where refined) then the
  // level-1 table (2nb[0] x 2nb[1], -1 outside the refined blocks).  Empty: one level, whose
  // table is slots itself (indexed by the brick's lexicographic block index).
  std::vector<int32_t> two_level;
  std::vector<char>    border;
  std::vector<int32_t> table_host;
  int32_t              prolong = 0;
};

// first-fit allocator of contiguous index ranges (slot-table segments)
struct RangeAlloc
{
  std::vector<std::pair<int32_t, int32_t>> free; // (start, length), sorted by start
  int32_t                                  end = 0;
  int32_t
  take(int32_t n)
  {
    for (size_t k = 0; k < free.size(); ++k)
      if (free[k].second >= n)
      {
        const int32_t s = free[k].first;
        free[k].first += n;
        free[k].second -= n;
        if (free[k].second == 0)
          free.erase(free.begin() + static_cast<long>(k));
        return s;
      }
    const int32_t s = end;
    end += n;
    return s;
  }
  void
  give(int32_t s, int32_t n)
  {
    free.emplace_back(s, n);
    std::sort(free.begin(), free.end());
    std::vector<std::pair<int32_t, int32_t>> merged;
    for (auto const & r : free)
    {
      if (!merged.empty() and merged.back().first + merged.back().second == r.first)
        merged.back().second += r.second;
      else
        merged.push_back(r);
    }
    free.swap(merged);
    if (!free.empty() and free.back().first + free.back().second == end)
    {
      end = free.back().first;
      free.pop_back();
    }
  }
};

} // namespace

struct g5_pool
{
  int32_t     dim = 2, thinc = 0, precision = G5_PREC_FP64;
  int32_t     nvar = 8, ncell = 256, block_doubles = 2048;
  int64_t     nslots = 0;
  double *    buf[2] = { nullptr, nullptr };
  double *    stage = nullptr; // RK2_SSP stage buffer (single-solve contexts only)
  hipStream_t stream = nullptr;
  std::string err;

  std::vector<int32_t>  free_slots; // stack
  std::vector<SolveRec> solves;
  std::vector<int32_t>  free_ids;
  RangeAlloc            table;

  // device mirrors
  SolveDesc * d_solves = nullptr;
  SolveCtl *  d_ctl = nullptr;
  int32_t     cap_solves = 0;
  int32_t *   d_slots = nullptr;
  int32_t     cap_table = 0;
  BlockDesc * d_blocks = nullptr;
  int64_t     cap_blocks = 0;
  int32_t     nlist = 0;
  // 3-D one-level lists: the step kernels' list of z-columns of kMaxColumn blocks (padded with
  // solve = -1), ncols of them
  BlockDesc * d_cols = nullptr;
  int64_t     cap_cols = 0;
  int32_t     ncols = 0;
  bool        list_dirty = true;
  bool        any_inlet = false; // a live solve has an inflow border: the 3-D step kernels resolve it
  bool        any_mp = false;    // a live solve has the MILLER_PUCKETT source term
  // a live solve has two levels: the launch list starts with the nlist_first fine blocks that
  // border coarse blocks, and the step runs as two launches over the interface records d_iflux
  bool        any_two_level = false;
  int32_t     nlist_first = 0;
  double *    d_iflux = nullptr; // IFLUX per slot: [4 sides][16 cells][8 values]
  int64_t     cap_iflux = 0;
  SolveCtl *  h_ctl = nullptr; // pinned mirror of the control records
  int32_t     cap_hctl = 0;

  // HIP graphs of one and of kGraphSteps device-dt steps over the current launch list
  hipGraphExec_t graph = nullptr;
  hipGraphExec_t graph_k = nullptr;
  bool           graph_valid = false;   // both graphs are stale when false
  bool           graph1_ok = false;     // captured for the current list
  bool           graphk_ok = false;

  // staging of g5_pool_upload / g5_pool_download for solves whose slots are not one run
  double * d_xfer = nullptr;
  int64_t  cap_xfer = 0;
  int32_t * d_xslots = nullptr; // slots of a two-level solve in its block order, for those transfers
  int64_t   cap_xslots = 0;

  // scratch of g5_pool_start (block list and ids of the solves being started)
  BlockDesc * d_start_list = nullptr;
  int64_t     cap_start_list = 0;
  int32_t *   d_start_ids = nullptr;
  int64_t     cap_start_ids = 0;

  // launch geometry of the step kernels: fixed by the caller, or chosen by the autotuner for the
  // current launch-list size class
  LaunchCfg launch{ 0, 1, 1, 0, kMaxColumn };
  bool      launch_fixed = false;
  bool      autotune = true;
  int32_t   tuned_class = -1; // size class the current launch was tuned (or looked up) for
  double    tuned_ms = 0.0;
  int32_t   cus = 256;

  double     last_run_ms = 0.0;
  hipEvent_t ev[2] = { nullptr, nullptr }; // around the last stepping call
  SolveCtl * h_poll = nullptr;             // pinned: control record copies of run_solve's chunks
  hipEvent_t ev_poll[2] = { nullptr, nullptr };
  bool       timing_pending = false;
};

struct g5_ctx
{
  g5_pool *            pool = nullptr;
  g5_config            cfg{};
  int32_t              id = 0;
  bool                 has_state = false;
  std::vector<int32_t> outside;
  int32_t *            d_outside = nullptr;
  double *             d_ring = nullptr;
  double *             d_staging = nullptr; // contiguous [block][var][cell] buffer for transfers
  std::vector<int32_t> block_order;         // caller's block order (brick coordinates), empty: lexicographic
  int32_t *            d_flags = nullptr;   // g5_refine_flags outputs, per slot
  double *             d_imax = nullptr;
  std::string          err;
};

namespace
{

int
fail(std::string & err, int code, std::string const & msg)
{
  err = msg;
  return code;
}

#define G5_HIP(errstr, call)                                                                       \
  do                                                                                               \
  {                                                                                                \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return fail(errstr, G5_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));          \
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
    if (c.nb[d] < 1 or c.nb[d] > 32767)
      return msg = "nb must be in [1, 32767]", G5_ERR_INVALID;
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
  if (c.precision != G5_PREC_FP64 and c.precision != G5_PREC_MIXED and c.precision != G5_PREC_FP64_FAST)
    return msg = "unknown precision", G5_ERR_INVALID;
  if (c.source_term != G5_SOURCE_ALLAIRE_KOKH and c.source_term != G5_SOURCE_MILLER_PUCKETT)
    return msg = "unknown volume-fraction source term", G5_ERR_INVALID;
  if (c.inlet != G5_INLET_NONE and c.inlet != G5_INLET_SHOCK_BUBBLE and c.inlet != G5_INLET_JET)
    return msg = "unknown inlet", G5_ERR_INVALID;
  if (c.inlet != G5_INLET_NONE and c.periodic[0])
    return msg = "an inlet needs a non-periodic x direction (it overrides outside cells with x < xmin)", G5_ERR_INVALID;
  if (c.inlet == G5_INLET_JET and not(c.jet[2] > 0.0))
    return msg = "the jet width must be > 0", G5_ERR_INVALID;
  return G5_OK;
}

int64_t
nblocks_of(g5_config const & c)
{
  return static_cast<int64_t>(c.nb[0]) * c.nb[1] * (c.dim == 3 ? c.nb[2] : 1);
}

template <typename T>
int
grow(g5_pool * p, T *& ptr, int64_t & cap, int64_t need, hipStream_t s)
{
  if (need <= cap)
    return G5_OK;
  int64_t nc = std::max<int64_t>(need, cap * 2);
  nc = std::max<int64_t>(nc, 16);
  T *     np = nullptr;
  G5_HIP(p->err, hipMalloc(&np, nc * sizeof(T)));
  if (ptr)
  {
    G5_HIP(p->err, hipMemcpyAsync(np, ptr, cap * sizeof(T), hipMemcpyDeviceToDevice, s));
    G5_HIP(p->err, hipStreamSynchronize(s));
    hipFree(ptr);
  }
  ptr = np;
  cap = nc;
  return G5_OK;
}

int
pool_create(int dim, int thinc, int precision, uint64_t budget, bool rk2_stage, g5_pool ** out, std::string & err)
{
  g5_pool * p = new g5_pool();
  p->dim = dim;
  p->launch.col_len = dim == 3 ? kMaxColumn : 1;
  p->thinc = thinc;
  p->precision = precision;
  p->nvar = dim == 2 ? 8 : 9;
  p->ncell = dim == 2 ? 256 : 4096;
  p->block_doubles = p->nvar * p->ncell;
  const uint64_t per_slot = static_cast<uint64_t>(p->block_doubles) * sizeof(double) * (rk2_stage ? 3 : 2);
  p->nslots = static_cast<int64_t>(budget / per_slot);
  if (p->nslots < 1)
  {
    delete p;
    return fail(err, G5_ERR_NOMEM, "HBM budget smaller than one block");
  }
  auto bail = [&](int code, std::string const & m) {
    g5_pool_destroy(p);
    return fail(err, code, m);
  };
  hipError_t e = hipStreamCreate(&p->stream);
  if (e != hipSuccess)
    return bail(G5_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess and
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess and cus > 0)
    p->cus = cus;
  if (char const * at = std::getenv("G5_AUTOTUNE"))
    p->autotune = std::atoi(at) != 0;
  const size_t bytes = static_cast<size_t>(p->nslots) * p->block_doubles * sizeof(double);
  for (int k = 0; k < 2; ++k)
    if ((e = hipMalloc(&p->buf[k], bytes)) != hipSuccess)
      return bail(G5_ERR_NOMEM, std::string("hipMalloc state: ") + hipGetErrorString(e));
  if (rk2_stage and (e = hipMalloc(&p->stage, bytes)) != hipSuccess)
    return bail(G5_ERR_NOMEM, std::string("hipMalloc stage: ") + hipGetErrorString(e));
  p->free_slots.resize(p->nslots);
  for (int64_t k = 0; k < p->nslots; ++k)
    p->free_slots[k] = static_cast<int32_t>(p->nslots - 1 - k); // pop_back gives 0, 1, 2, ...
  *out = p;
  return G5_OK;
}

int
ensure_solve_capacity(g5_pool * p, int32_t need)
{
  if (need <= p->cap_solves)
    return G5_OK;
  int64_t cs = p->cap_solves, cc = p->cap_solves;
  int     rc = grow(p, p->d_solves, cs, need, p->stream);
  if (rc != G5_OK)
    return rc;
  rc = grow(p, p->d_ctl, cc, need, p->stream);
  if (rc != G5_OK)
    return rc;
  p->cap_solves = static_cast<int32_t>(std::min(cs, cc));
  p->graph_valid = false;
  return G5_OK;
}

// device side of an admission: descriptor, control record and slot-table segment of solve sid
int
admit_device(g5_pool * p, int32_t sid, double tend, int32_t max_steps)
{
  int rc = ensure_solve_capacity(p, static_cast<int32_t>(p->solves.size()));
  if (rc != G5_OK)
    return rc;
  SolveRec const & r = p->solves[sid];
  int64_t          ct = p->cap_table;
  rc = grow(p, p->d_slots, ct, p->table.end, p->stream);
  if (rc != G5_OK)
    return rc;
  p->cap_table = static_cast<int32_t>(ct);
  SolveDesc sd;
  make_solve_desc(r.cfg, r.table_base, sd);
  std::vector<int32_t> const & tab = r.two_level.empty() ? r.slots : r.table_host;
  if (!r.two_level.empty())
  {
    sd.base1 = r.table_base + r.cfg.nb[0] * r.cfg.nb[1] * (r.cfg.dim == 3 ? r.cfg.nb[2] : 1);
    sd.prolong = r.prolong;
    // interface records, indexed by slot
    rc = grow(p, p->d_iflux, p->cap_iflux, p->nslots * (r.cfg.dim == 3 ? kIflux3Doubles : kIfluxDoubles), p->stream);
    if (rc != G5_OK)
      return rc;
  }
  SolveCtl c{};
  c.tend = tend;
  c.max_steps = max_steps;
  G5_HIP(p->err, hipMemcpyAsync(p->d_slots + r.table_base, tab.data(), r.table_len * sizeof(int32_t),
                                hipMemcpyHostToDevice, p->stream));
  G5_HIP(p->err, hipMemcpyAsync(p->d_solves + sid, &sd, sizeof(SolveDesc), hipMemcpyHostToDevice, p->stream));
  G5_HIP(p->err, hipMemcpyAsync(p->d_ctl + sid, &c, sizeof(SolveCtl), hipMemcpyHostToDevice, p->stream));
  G5_HIP(p->err, hipStreamSynchronize(p->stream));
  return G5_OK;
}

// A two-level static mesh over the brick of cfg: the coarse blocks at refined[dim k .. dim k + dim)
// are each replaced by 2^dim blocks of half the cell length.  Fills r.two_level (x, y, z, level per
// block; order: coarse leaves lexicographic, then the fine blocks lexicographic over the fine grid)
// and r.border; r.nblk is the block count.
int
layout_two_level(g5_config const & cfg, std::vector<int32_t> const & refined, SolveRec & r, std::string & msg)
{
  const int dim = cfg.dim;
  const int nx = cfg.nb[0], ny = cfg.nb[1], nz = dim == 3 ? cfg.nb[2] : 1;
  if (cfg.precision == G5_PREC_MIXED)
    return msg = "two-level meshes exchange FP64 face records: FP64 or FP64_FAST precision", G5_ERR_UNSUPPORTED;
  if (refined.empty() or refined.size() % dim != 0)
    return msg = "no refined blocks", G5_ERR_INVALID;
  std::vector<char> ref(static_cast<size_t>(nx) * ny * nz, 0);
  for (size_t k = 0; k < refined.size(); k += dim)
  {
    const int c[3] = { refined[k], refined[k + 1], dim == 3 ? refined[k + 2] : 0 };
    for (int d = 0; d < dim; ++d)
      if (c[d] < 1 or c[d] > cfg.nb[d] - 2)
        return msg = "refined blocks must lie at least one block inside the brick's border", G5_ERR_INVALID;
    char & f = ref[c[0] + static_cast<size_t>(nx) * (c[1] + static_cast<size_t>(ny) * c[2])];
    if (f)
      return msg = "refined block repeated", G5_ERR_INVALID;
    f = 1;
  }
  r.two_level.clear();
  r.border.clear();
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x)
        if (!ref[x + static_cast<size_t>(nx) * (y + static_cast<size_t>(ny) * z)])
        {
          r.two_level.insert(r.two_level.end(), { x, y, z, 0 });
          r.border.push_back(0);
        }
  auto fine = [&](int x, int y, int z) {
    return ref[(x >> 1) + static_cast<size_t>(nx) * ((y >> 1) + static_cast<size_t>(ny) * (z >> 1))] != 0;
  };
  const int fz = dim == 3 ? 2 * nz : 1;
  for (int z = 0; z < fz; ++z)
    for (int y = 0; y < 2 * ny; ++y)
      for (int x = 0; x < 2 * nx; ++x)
        if (fine(x, y, z))
        {
          r.two_level.insert(r.two_level.end(), { x, y, z, 1 });
          bool b = !fine(x - 1, y, z) or !fine(x + 1, y, z) or !fine(x, y - 1, z) or !fine(x, y + 1, z);
          if (dim == 3)
            b = b or !fine(x, y, z - 1) or !fine(x, y, z + 1);
          r.border.push_back(b);
        }
  r.nblk = static_cast<int32_t>(r.border.size());
  return G5_OK;
}

// slot-table segment of a two-level solve from its slots: level 0 (nb[0] nb[1] nb[2]) then level 1
// (2nb[0] 2nb[1] 2nb[2], 2-D: 2nb[0] 2nb[1])
void
tables_two_level(SolveRec & r)
{
  const int    dim = r.cfg.dim;
  const int    nx = r.cfg.nb[0], ny = r.cfg.nb[1], nz = dim == 3 ? r.cfg.nb[2] : 1;
  const size_t n0 = static_cast<size_t>(nx) * ny * nz;
  r.table_host.assign(n0 * (1 + (dim == 3 ? 8 : 4)), -1);
  for (int32_t k = 0; k < r.nblk; ++k)
  {
    const int x = r.two_level[4 * k], y = r.two_level[4 * k + 1], z = r.two_level[4 * k + 2];
    if (r.two_level[4 * k + 3] == 0)
      r.table_host[x + static_cast<size_t>(nx) * (y + static_cast<size_t>(ny) * z)] = r.slots[k];
    else
      r.table_host[n0 + x + static_cast<size_t>(2 * nx) * (y + static_cast<size_t>(2 * ny) * z)] = r.slots[k];
  }
}

// admission of a solve; refined (two coordinates per refined coarse block), when given, makes it
// a two-level solve with the given prolongation (0 SIMPLE_COPY, 1 EXTRAPOLATE_LINEAR_MINMOD,
// 2 / 3 CONSERVATIVE_INTERPOLATION_ORDER_2 / _4)
int
pool_admit(g5_pool * p, g5_config const & cfg, double tend, int32_t max_steps, int32_t * id,
           std::vector<int32_t> const * refined = nullptr, int prolong = 0)
{
  std::string msg;
  int         rc = validate(cfg, msg);
  if (rc != G5_OK)
    return fail(p->err, rc, msg);
  SolveRec tl; // two-level layout
  if (refined)
  {
    if (prolong < 0 or prolong > 3)
      return fail(p->err, G5_ERR_INVALID, "unknown prolongation");
    if ((rc = layout_two_level(cfg, *refined, tl, msg)) != G5_OK)
      return fail(p->err, rc, msg);
  }
  if (cfg.dim != p->dim or (cfg.thinc_enabled != 0) != (p->thinc != 0) or cfg.precision != p->precision)
    return fail(p->err, G5_ERR_INVALID, "solve does not match the pool (dim, THINC, precision)");
  for (auto const & r : p->solves) // the step graph applies one integrator to the whole launch list
    if (r.live and r.cfg.integrator != cfg.integrator)
      return fail(p->err, G5_ERR_INVALID, "solve does not match the pool's integrator (one integrator per pool)");
  if (p->stage == nullptr and cfg.integrator == G5_TI_RK2_SSP)
    return fail(p->err, G5_ERR_UNSUPPORTED, "RK2_SSP needs the stage buffer (a single-solve context or g5_pool_create_rk2)");
  const int64_t nblk = refined ? tl.nblk : nblocks_of(cfg);
  if (nblk > p->nslots)
    return fail(p->err, G5_ERR_TOO_LARGE,
                "the solve has " + std::to_string(nblk) + " blocks, more than the pool's whole budget (" +
                  std::to_string(p->nslots) + ")");
  if (nblk > static_cast<int64_t>(p->free_slots.size()))
    return fail(p->err, G5_ERR_NOMEM, "HBM budget exhausted");
  int32_t sid;
  if (!p->free_ids.empty())
  {
    sid = p->free_ids.back();
    p->free_ids.pop_back();
  }
  else
  {
    sid = static_cast<int32_t>(p->solves.size());
    p->solves.emplace_back();
  }
  SolveRec & r = p->solves[sid];
  r.live = true;
  r.started = false;
  r.cfg = cfg;
  p->any_inlet = p->any_inlet or cfg.inlet != G5_INLET_NONE; // before any launch that lists it
  p->any_mp = p->any_mp or cfg.source_term == G5_SOURCE_MILLER_PUCKETT;
  for (int d = cfg.dim; d < 3; ++d)
    r.cfg.nb[d] = 1;
  r.nblk = static_cast<int32_t>(nblk);
  r.slots.resize(nblk);
  for (int64_t k = 0; k < nblk; ++k)
  {
    r.slots[k] = p->free_slots.back();
    p->free_slots.pop_back();
  }
  r.table_len = r.nblk;
  if (refined)
  {
    r.two_level = std::move(tl.two_level);
    r.border = std::move(tl.border);
    r.prolong = prolong;
    tables_two_level(r);
    r.table_len = static_cast<int32_t>(r.table_host.size());
    p->any_two_level = true;
  }
  r.table_base = p->table.take(r.table_len);
  rc = admit_device(p, sid, tend, max_steps);
  if (rc != G5_OK)
  {
    // give everything back: the solve was never admitted
    SolveRec & q = p->solves[sid];
    for (auto it = q.slots.rbegin(); it != q.slots.rend(); ++it)
      p->free_slots.push_back(*it);
    p->table.give(q.table_base, q.table_len);
    q = SolveRec();
    p->free_ids.push_back(sid);
    return rc;
  }
  p->list_dirty = true;
  *id = sid;
  return G5_OK;
}

int
pool_release(g5_pool * p, int32_t id)
{
  if (id < 0 or id >= static_cast<int32_t>(p->solves.size()) or !p->solves[id].live)
    return fail(p->err, G5_ERR_INVALID, "no such solve");
  G5_HIP(p->err, hipStreamSynchronize(p->stream));
  SolveRec & r = p->solves[id];
  for (auto it = r.slots.rbegin(); it != r.slots.rend(); ++it)
    p->free_slots.push_back(*it); // the next admission pops them in ascending order
  p->table.give(r.table_base, r.table_len);
  r = SolveRec();
  p->free_ids.push_back(id);
  SolveCtl c{}; // inactive
  G5_HIP(p->err, hipMemcpyAsync(p->d_ctl + id, &c, sizeof(SolveCtl), hipMemcpyHostToDevice, p->stream));
  G5_HIP(p->err, hipStreamSynchronize(p->stream));
  p->list_dirty = true;
  return G5_OK;
}

// part (two-level solves): 0 the fine blocks with a coarse neighbour, 1 the other blocks, 2 all
void
append_blocks(g5_pool const * p, int32_t sid, std::vector<BlockDesc> & out, int part = 2)
{
  SolveRec const & r = p->solves[sid];
  if (!r.two_level.empty())
  {
    for (int32_t k = 0; k < r.nblk; ++k)
      if (part == 2 or (part == 0) == (r.border[k] != 0))
        out.push_back(BlockDesc{ sid, r.slots[k], static_cast<int16_t>(r.two_level[4 * k]),
                                 static_cast<int16_t>(r.two_level[4 * k + 1]),
                                 static_cast<int16_t>(r.two_level[4 * k + 2]),
                                 static_cast<int16_t>(r.two_level[4 * k + 3]) });
    return;
  }
  if (part == 0)
    return;
  const int nx = r.cfg.nb[0], ny = r.cfg.nb[1], nz = r.cfg.nb[2];
  // 2-D: 4x4-block tiles in x-fastest tile order, so that a block's y neighbours (whose halo rows
  // it reads) run a few workgroups apart (measured 2.6% faster on a 4096-solve batch, although the
  // fetched HBM volume did not drop); 3-D: plain brick order
  const int T = r.cfg.dim == 2 ? 4 : 1;
  for (int z = 0; z < nz; ++z)
    for (int ty = 0; ty < ny; ty += T)
      for (int tx = 0; tx < nx; tx += T)
        for (int y = ty; y < std::min(ty + T, ny); ++y)
          for (int x = tx; x < std::min(tx + T, nx); ++x)
          {
            const int k = x + nx * (y + ny * z);
            out.push_back(BlockDesc{ sid, r.slots[k], static_cast<int16_t>(x), static_cast<int16_t>(y),
                                     static_cast<int16_t>(z), 0 });
          }
}

int
rebuild_list(g5_pool * p)
{
  if (!p->list_dirty)
    return G5_OK;
  std::vector<BlockDesc> list;
  bool                   inlet = false, mp = false, two = false;
  for (auto const & r : p->solves)
    two = two or (r.live and !r.two_level.empty());
  // with two-level solves: the fine blocks with a coarse neighbour first
  if (two)
    for (int32_t s = 0; s < static_cast<int32_t>(p->solves.size()); ++s)
      if (p->solves[s].live)
        append_blocks(p, s, list, 0);
  p->nlist_first = static_cast<int32_t>(list.size());
  for (int32_t s = 0; s < static_cast<int32_t>(p->solves.size()); ++s)
    if (p->solves[s].live)
    {
      append_blocks(p, s, list, two ? 1 : 2);
      inlet = inlet or p->solves[s].cfg.inlet != G5_INLET_NONE;
      mp = mp or p->solves[s].cfg.source_term == G5_SOURCE_MILLER_PUCKETT;
    }
  p->any_inlet = inlet;
  p->any_mp = mp;
  p->any_two_level = two;
  int64_t cb = p->cap_blocks;
  int     rc = grow(p, p->d_blocks, cb, static_cast<int64_t>(list.size()), p->stream);
  if (rc != G5_OK)
    return rc;
  p->cap_blocks = cb;
  if (!list.empty())
    G5_HIP(p->err, hipMemcpyAsync(p->d_blocks, list.data(), list.size() * sizeof(BlockDesc), hipMemcpyHostToDevice,
                                  p->stream));
  G5_HIP(p->err, hipStreamSynchronize(p->stream));
  p->nlist = static_cast<int32_t>(list.size());
  // 3-D one-level lists: columns of kMaxColumn consecutive z blocks per solve, in the order
  // (column z range, y, x), padded with solve = -1 where nb[2] is not a multiple
  std::vector<BlockDesc> cols;
  if (p->dim == 3 and !two)
    for (int32_t s = 0; s < static_cast<int32_t>(p->solves.size()); ++s)
    {
      SolveRec const & r = p->solves[s];
      if (!r.live)
        continue;
      const int nx = r.cfg.nb[0], ny = r.cfg.nb[1], nz = r.cfg.nb[2];
      for (int z0 = 0; z0 < nz; z0 += kMaxColumn)
        for (int y = 0; y < ny; ++y)
          for (int x = 0; x < nx; ++x)
            for (int z = z0; z < z0 + kMaxColumn; ++z)
              cols.push_back(z < nz ? BlockDesc{ s, r.slots[x + nx * (y + ny * z)], static_cast<int16_t>(x),
                                                 static_cast<int16_t>(y), static_cast<int16_t>(z), 0 }
                                    : BlockDesc{ -1, -1, 0, 0, 0, 0 });
    }
  if (!cols.empty())
  {
    int64_t cc = p->cap_cols;
    rc = grow(p, p->d_cols, cc, static_cast<int64_t>(cols.size()), p->stream);
    if (rc != G5_OK)
      return rc;
    p->cap_cols = cc;
    G5_HIP(p->err, hipMemcpyAsync(p->d_cols, cols.data(), cols.size() * sizeof(BlockDesc), hipMemcpyHostToDevice,
                                  p->stream));
    G5_HIP(p->err, hipStreamSynchronize(p->stream));
  }
  p->ncols = static_cast<int32_t>(cols.size() / kMaxColumn);
  // without tuning: the longest z-columns that still leave two work items per CU (a short list
  // marched in long columns would leave CUs idle)
  if (p->dim == 3 and !p->launch_fixed and !p->autotune)
  {
    int cl = kMaxColumn;
    while (cl > 1 and p->nlist / cl < 2 * p->cus)
      cl /= 2;
    if (cl != p->launch.col_len)
      p->graph_valid = false;
    p->launch.col_len = cl;
  }
  p->list_dirty = false;
  p->graph_valid = false;
  return G5_OK;
}

StepArgs
step_args(g5_pool * p, BlockDesc const * blocks, int32_t n, double dt_fixed)
{
  StepArgs A;
  A.buf0 = p->buf[0];
  A.buf1 = p->buf[1];
  A.U_in = nullptr;
  A.U_out = nullptr;
  A.blocks = blocks;
  A.solves = p->d_solves;
  A.ctl = p->d_ctl;
  A.slots = p->d_slots;
  A.nblocks = n;
  A.per_xcd = (n + 7) / 8;
  A.wg_per_xcd = (p->dim == 2 and p->launch.wg_per_cu > 0) ? p->launch.wg_per_cu * std::max(1, p->cus / 8) : 0;
  A.remap = p->launch.remap;
  A.dt_fixed = dt_fixed;
  A.iflux = p->d_iflux;
  A.col_stride = 1;
  A.col_len = 1;
  if (p->dim == 3 and !p->any_two_level and blocks == p->d_blocks and p->ncols > 0 and p->launch.col_len > 1)
  {
    // the 3-D step marches z-columns: items of col_len blocks of the column list (with one block
    // per item the plain list serves, without the columns' padding entries)
    A.blocks = p->d_cols;
    A.col_stride = kMaxColumn;
    A.col_len = p->launch.col_len;
    A.nblocks = p->ncols * (kMaxColumn / p->launch.col_len);
    A.per_xcd = (A.nblocks + 7) / 8;
  }
  return A;
}

void
launch_step(g5_pool * p, StepArgs const & A, bool hancock)
{
  if (A.nblocks <= 0)
    return;
  if (p->any_two_level) // every caller passes the whole launch list
  {
    if (p->dim == 2)
      launch_step2d_two_level(A, p->nlist_first, p->thinc != 0, hancock, p->precision, p->any_mp, p->stream);
    else
      launch_step3d_two_level(A, p->nlist_first, p->thinc != 0, hancock, p->precision, p->stream);
  }
  else if (p->dim == 2)
    launch_step2d(A, p->thinc != 0, hancock, p->precision, p->launch.ring_wave != 0, p->any_mp, p->stream);
  else
    launch_step3d(A, p->thinc != 0, hancock, p->precision, p->launch.zsplit != 0, p->any_inlet, p->any_mp, p->stream);
}

// ---------------------------------------------------------------------------------------------
// Autotuner of the launch geometry.  Candidates: one workgroup per block with XCD-contiguous
// block ranges or plain order, in 2-D the five-wavefront ring form, in 3-D the two z-solve
// placements.  Each candidate runs the
// device-dt step kernel of the live launch list a few times; a step kernel only writes the next
// state buffer and max-accumulates the next CFL rate, both functions of the current state alone,
// so the repeated launches leave exactly what the real step will write.  The best candidate is
// kept per (dim, THINC, precision, integrator, list-size class) for the life of the process.
// ---------------------------------------------------------------------------------------------
struct TuneKey
{
  int32_t dim, thinc, precision, hancock, size_class;
  bool
  operator<(TuneKey const & o) const
  {
    return std::tie(dim, thinc, precision, hancock, size_class) <
           std::tie(o.dim, o.thinc, o.precision, o.hancock, o.size_class);
  }
};
struct TuneVal
{
  LaunchCfg cfg;
  double    ms;
};
std::mutex                    g_tune_mutex;
std::map<TuneKey, TuneVal>    g_tune_cache;

int32_t
size_class(int32_t nblocks)
{
  int32_t c = 0;
  while ((1 << (c + 1)) <= nblocks)
    ++c;
  return c; // floor(log2(nblocks))
}

// The persistent 2-D form (workgroups looping over a block range) is not a candidate: it measured
// 1.7-2.4x slower everywhere, and it shares the looping block_range branch with the persistent
// 3-D form, which faulted; g5_pool_set_launch still selects it explicitly.
std::vector<LaunchCfg>
tune_candidates(g5_pool const * p)
{
  std::vector<LaunchCfg> c;
  for (int z = (p->dim == 3 ? 0 : 1); z <= 1; ++z)
  {
    c.push_back({ 0, 1, z, 0, 1 });
    c.push_back({ 0, 0, z, 0, 1 });
    if (p->dim == 2) // a fifth wavefront for the ring work: shorter per-block latency, fewer registers per lane
      c.push_back({ 0, 1, z, 1, 1 });
    else // z-columns of 2, 4 and 8 blocks per workgroup (fewer column ends and prologues per plane)
      for (int cl = 2; cl <= kMaxColumn; cl *= 2)
        c.push_back({ 0, 1, z, 0, cl });
  }
  return c;
}

int
autotune(g5_pool * p, bool hancock)
{
  const TuneKey key{ p->dim, p->thinc, p->precision, hancock ? 1 : 0, size_class(p->nlist) };
  {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end())
    {
      p->launch = it->second.cfg;
      p->tuned_ms = it->second.ms;
      p->tuned_class = key.size_class;
      return G5_OK;
    }
  }
  hipEvent_t e0, e1;
  G5_HIP(p->err, hipEventCreate(&e0));
  G5_HIP(p->err, hipEventCreate(&e1));
  LaunchCfg best = p->launch;
  double    best_ms = 1e300;
  for (LaunchCfg const & c : tune_candidates(p))
  {
    p->launch = c;
    const StepArgs A = step_args(p, p->d_blocks, p->nlist, 0.0);
    launch_step(p, A, hancock); // warm-up
    G5_HIP(p->err, hipEventRecord(e0, p->stream));
    constexpr int reps = 3;
    for (int r = 0; r < reps; ++r)
      launch_step(p, A, hancock);
    G5_HIP(p->err, hipEventRecord(e1, p->stream));
    G5_HIP(p->err, hipEventSynchronize(e1));
    float ms = 0.f;
    G5_HIP(p->err, hipEventElapsedTime(&ms, e0, e1));
    if (ms / reps < best_ms)
    {
      best_ms = ms / reps;
      best = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  G5_HIP(p->err, hipGetLastError());
  p->launch = best;
  p->tuned_ms = best_ms;
  p->tuned_class = key.size_class;
  std::lock_guard<std::mutex> lk(g_tune_mutex);
  g_tune_cache[key] = TuneVal{ best, best_ms };
  return G5_OK;
}

// before a device-dt step over a new launch list: tune (or look up) its launch geometry
int
ensure_tuned(g5_pool * p, bool hancock)
{
  if (p->launch_fixed or !p->autotune or p->nlist == 0 or p->tuned_class == size_class(p->nlist) or
      p->any_two_level) // the two-level step has one launch form
    return G5_OK;
  const LaunchCfg before = p->launch;
  const int       rc = autotune(p, hancock);
  if (rc == G5_OK and std::memcmp(&before, &p->launch, sizeof(LaunchCfg)) != 0)
    p->graph_valid = false;
  return rc;
}

bool
pool_rk2(g5_pool const * p)
{
  for (auto const & r : p->solves)
    if (r.live)
      return r.cfg.integrator == G5_TI_RK2_SSP;
  return false;
}

bool
pool_hancock(g5_pool const * p)
{
  for (auto const & r : p->solves)
    if (r.live)
      return r.cfg.integrator == G5_TI_HANCOCK;
  return true;
}

// steps per launch of the multi-step graph (fewer host launches for small, latency-bound lists)
constexpr int kGraphSteps = 8;

// capture nsteps device-dt steps (step kernel + finalise each) of the current launch list
int
capture_steps(g5_pool * p, int nsteps, hipGraphExec_t * out)
{
  if (*out)
  {
    (void)hipGraphExecDestroy(*out);
    *out = nullptr;
  }
  // RK2_SSP (pools with the stage buffer: single-solve contexts, g5_pool_create_rk2): per step, stage 1
  // U_RK = U + dt L(U) into the stage buffer, stage 2 U2 = U_RK + dt L(U_RK) into the other
  // buffer, the CFL rates of the stages discarded, U = 1/2 U2 + 1/2 U in place (the same bits as
  // 1/2 U + 1/2 U2), the CFL rate of the new U, and the finalisation without a buffer flip; every
  // buffer is chosen on the device from the solve's parity (SolverGodunovFiveEq.cpp:496-524)
  const bool rk2 = p->stage != nullptr and pool_rk2(p);
  hipGraph_t g;
  G5_HIP(p->err, hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  const StepArgs A = step_args(p, p->d_blocks, p->nlist, 0.0);
  const bool     hancock = pool_hancock(p);
  const int      nsolves = static_cast<int>(p->solves.size());
  for (int k = 0; k < nsteps; ++k)
  {
    if (rk2)
    {
      StepArgs S1 = A; // reads the current buffer (by the device parity)
      S1.U_out = p->stage;
      launch_step(p, S1, false);
      StepArgs S2 = A; // writes the other buffer
      S2.U_in = p->stage;
      launch_step(p, S2, false);
      launch_reset_invdt(p->d_ctl, nsolves, p->stream);
      launch_rk2_combine(p->buf[0], p->buf[1], p->d_blocks, p->d_ctl, p->nlist, p->block_doubles, p->stream);
      launch_inv_dt(p->dim, p->precision, p->buf[0], p->buf[1], p->d_blocks, p->d_solves, p->d_ctl,
                    p->nlist, p->stream);
      launch_finalize(p->d_ctl, p->d_solves, nsolves, p->stream, false);
    }
    else
    {
      launch_step(p, A, hancock);
      launch_finalize(p->d_ctl, p->d_solves, nsolves, p->stream);
    }
  }
  G5_HIP(p->err, hipStreamEndCapture(p->stream, &g));
  G5_HIP(p->err, hipGraphInstantiate(out, g, nullptr, nullptr, 0));
  (void)hipGraphDestroy(g);
  return G5_OK;
}

// (re)capture, when stale, the step graph of 1 or of kGraphSteps steps over the current list
int
pool_capture(g5_pool * p, bool multi)
{
  if (p->nlist == 0)
    return G5_OK;
  if (!p->graph_valid)
  {
    p->graph1_ok = p->graphk_ok = false;
    p->graph_valid = true;
  }
  bool & ok = multi ? p->graphk_ok : p->graph1_ok;
  if (ok)
    return G5_OK;
  const int rc = capture_steps(p, multi ? kGraphSteps : 1, multi ? &p->graph_k : &p->graph);
  if (rc != G5_OK)
  {
    p->graph_valid = false;
    return rc;
  }
  ok = true;
  return G5_OK;
}

int pool_timing(g5_pool * p);

// n device-dt steps of every active solve, as launches of the captured step graphs
int
pool_launch_steps(g5_pool * p, int32_t n)
{
  if (p->nlist == 0 or n <= 0)
    return G5_OK;
  if (n >= kGraphSteps)
  {
    const int rc = pool_capture(p, true);
    if (rc != G5_OK)
      return rc;
  }
  if (n % kGraphSteps != 0)
  {
    const int rc = pool_capture(p, false);
    if (rc != G5_OK)
      return rc;
  }
  for (; n >= kGraphSteps; n -= kGraphSteps)
    G5_HIP(p->err, hipGraphLaunch(p->graph_k, p->stream));
  for (; n > 0; --n)
    G5_HIP(p->err, hipGraphLaunch(p->graph, p->stream));
  return G5_OK;
}

int
pool_step(g5_pool * p, int32_t nsteps, bool sync)
{
  int rc = rebuild_list(p);
  if (rc != G5_OK)
    return rc;
  rc = ensure_tuned(p, pool_hancock(p));
  if (rc != G5_OK)
    return rc;
  if (p->ev[0] == nullptr)
  {
    G5_HIP(p->err, hipEventCreate(&p->ev[0]));
    G5_HIP(p->err, hipEventCreate(&p->ev[1]));
  }
  G5_HIP(p->err, hipEventRecord(p->ev[0], p->stream));
  rc = pool_launch_steps(p, nsteps);
  if (rc != G5_OK)
    return rc;
  G5_HIP(p->err, hipEventRecord(p->ev[1], p->stream));
  G5_HIP(p->err, hipGetLastError());
  p->timing_pending = true;
  return sync ? pool_timing(p) : G5_OK;
}

// Up to nsteps device-dt steps of solve `id`, stopping once it has finished: the steps go out in
// chunks (64 steps, doubling up to 4096), each followed by an asynchronous copy of the solve's
// control record into pinned memory.  The host waits for a chunk's record only after the next
// chunk is queued, so the device never waits for the host, and stops queueing once a record shows
// the solve inactive (the chunk still in flight then runs as no-op steps: every block returns at
// once).  A run to tEnd with a large step budget thus costs the steps it takes.
int
pool_run_solve(g5_pool * p, int32_t id, int32_t nsteps)
{
  int rc = rebuild_list(p);
  if (rc != G5_OK)
    return rc;
  rc = ensure_tuned(p, pool_hancock(p));
  if (rc != G5_OK)
    return rc;
  if (p->ev[0] == nullptr)
  {
    G5_HIP(p->err, hipEventCreate(&p->ev[0]));
    G5_HIP(p->err, hipEventCreate(&p->ev[1]));
  }
  if (p->h_poll == nullptr)
  {
    G5_HIP(p->err, hipHostMalloc(&p->h_poll, 2 * sizeof(SolveCtl), hipHostMallocDefault));
    G5_HIP(p->err, hipEventCreateWithFlags(&p->ev_poll[0], hipEventDisableTiming));
    G5_HIP(p->err, hipEventCreateWithFlags(&p->ev_poll[1], hipEventDisableTiming));
  }
  G5_HIP(p->err, hipEventRecord(p->ev[0], p->stream));
  int32_t launched = 0, chunk = 64;
  int     slot = 0;
  bool    previous = false, done = false;
  while (not done and launched < nsteps)
  {
    const int32_t n = std::min(chunk, nsteps - launched);
    rc = pool_launch_steps(p, n);
    if (rc != G5_OK)
      return rc;
    launched += n;
    G5_HIP(p->err, hipMemcpyAsync(p->h_poll + slot, p->d_ctl + id, sizeof(SolveCtl), hipMemcpyDeviceToHost, p->stream));
    G5_HIP(p->err, hipEventRecord(p->ev_poll[slot], p->stream));
    if (previous) // the chunk before this one: finished?
    {
      G5_HIP(p->err, hipEventSynchronize(p->ev_poll[slot ^ 1]));
      done = p->h_poll[slot ^ 1].active == 0;
    }
    previous = true;
    slot ^= 1;
    chunk = std::min(2 * chunk, 4096);
  }
  G5_HIP(p->err, hipEventRecord(p->ev[1], p->stream));
  G5_HIP(p->err, hipGetLastError());
  p->timing_pending = true;
  return pool_timing(p);
}

// wait for the last stepping call and read its HIP-event duration
int
pool_timing(g5_pool * p)
{
  if (!p->timing_pending)
    return G5_OK;
  G5_HIP(p->err, hipEventSynchronize(p->ev[1]));
  float ms = 0.f;
  G5_HIP(p->err, hipEventElapsedTime(&ms, p->ev[0], p->ev[1]));
  p->last_run_ms = ms;
  p->timing_pending = false;
  return G5_OK;
}

// first dt of the given solves from their current state
int
pool_start(g5_pool * p, int32_t const * ids, int32_t n)
{
  std::vector<BlockDesc> list;
  for (int32_t k = 0; k < n; ++k)
  {
    if (ids[k] < 0 or ids[k] >= static_cast<int32_t>(p->solves.size()) or !p->solves[ids[k]].live)
      return fail(p->err, G5_ERR_INVALID, "no such solve");
    append_blocks(p, ids[k], list);
  }
  int rc = grow(p, p->d_start_list, p->cap_start_list, std::max<int64_t>(1, list.size()), p->stream);
  if (rc == G5_OK)
    rc = grow(p, p->d_start_ids, p->cap_start_ids, std::max<int64_t>(1, n), p->stream);
  if (rc != G5_OK)
    return rc;
  G5_HIP(p->err, hipMemcpyAsync(p->d_start_list, list.data(), list.size() * sizeof(BlockDesc), hipMemcpyHostToDevice,
                                p->stream));
  G5_HIP(p->err, hipMemcpyAsync(p->d_start_ids, ids, n * sizeof(int32_t), hipMemcpyHostToDevice, p->stream));
  // invdt of these solves is 0 (admit and finalize leave it cleared)
  launch_inv_dt(p->dim, p->precision, p->buf[0], p->buf[1], p->d_start_list, p->d_solves, p->d_ctl,
                static_cast<int>(list.size()), p->stream);
  launch_start(p->d_ctl, p->d_solves, p->d_start_ids, n, p->stream);
  G5_HIP(p->err, hipGetLastError());
  G5_HIP(p->err, hipStreamSynchronize(p->stream));
  for (int32_t k = 0; k < n; ++k)
    p->solves[ids[k]].started = true;
  return G5_OK;
}

int read_ctl(g5_pool * p);

// a new state of solve id (upload, device initial condition): clear its failure bits
int
clear_failure(g5_pool * p, int32_t id)
{
  static const uint32_t zero = 0u;
  G5_HIP(p->err, hipMemcpyAsync(reinterpret_cast<char *>(p->d_ctl + id) + offsetof(SolveCtl, failure), &zero,
                                sizeof(uint32_t), hipMemcpyHostToDevice, p->stream));
  G5_HIP(p->err, hipStreamSynchronize(p->stream));
  return G5_OK;
}

// message of a failed solve
std::string
failure_message(int32_t id, SolveCtl const & c)
{
  std::string m = "solve " + std::to_string(id) + " stopped after " + std::to_string(c.steps) + " steps (t = " +
                  std::to_string(c.t) + "):";
  if (c.failure & G5_FAIL_NONFINITE)
    m += " non-finite state or time step";
  if (c.failure & G5_FAIL_NEGATIVE_PRESSURE)
    m += " negative pressure";
  return m;
}

int
pool_transfer(g5_pool * p, int32_t id, double * host, size_t n, bool upload, double ** staging)
{
  if (id < 0 or id >= static_cast<int32_t>(p->solves.size()) or !p->solves[id].live)
    return fail(p->err, G5_ERR_INVALID, "no such solve");
  SolveRec const & r = p->solves[id];
  const size_t     need = static_cast<size_t>(r.nblk) * p->block_doubles;
  if (n != need)
    return fail(p->err, G5_ERR_INVALID, "state size mismatch");
  // slots in one ascending run (the usual case in a pool that is not fragmented): the state is
  // already contiguous in the current buffer, one copy
  bool contiguous = true;
  for (int32_t k = 1; k < r.nblk and contiguous; ++k)
    contiguous = r.slots[k] == r.slots[0] + k;
  if (contiguous)
  {
    int rc = read_ctl(p);
    if (rc != G5_OK)
      return rc;
    double * base = p->buf[p->h_ctl[id].parity] + static_cast<int64_t>(r.slots[0]) * p->block_doubles;
    if (upload)
      G5_HIP(p->err, hipMemcpyAsync(base, host, need * sizeof(double), hipMemcpyHostToDevice, p->stream));
    else
      G5_HIP(p->err, hipMemcpyAsync(host, base, need * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    G5_HIP(p->err, hipStreamSynchronize(p->stream));
    return upload ? clear_failure(p, id) : G5_OK;
  }
  double * st = staging ? *staging : nullptr;
  if (!st)
  {
    int rc = grow(p, p->d_xfer, p->cap_xfer, static_cast<int64_t>(need), p->stream);
    if (rc != G5_OK)
      return rc;
    st = p->d_xfer;
  }
  // block k's slot: the solve's slot-table segment, or for a two-level solve (whose segment holds
  // its level tables) its slot list uploaded for the transfer
  int32_t const * slot_list = p->d_slots;
  int32_t         base = r.table_base;
  if (!r.two_level.empty())
  {
    int rc = grow(p, p->d_xslots, p->cap_xslots, static_cast<int64_t>(r.nblk), p->stream);
    if (rc != G5_OK)
      return rc;
    G5_HIP(p->err, hipMemcpyAsync(p->d_xslots, r.slots.data(), r.nblk * sizeof(int32_t), hipMemcpyHostToDevice,
                                  p->stream));
    slot_list = p->d_xslots;
    base = 0;
  }
  if (upload)
  {
    G5_HIP(p->err, hipMemcpyAsync(st, host, need * sizeof(double), hipMemcpyHostToDevice, p->stream));
    launch_gather(p->buf[0], p->buf[1], p->d_ctl, id, slot_list, base, r.nblk, p->block_doubles, st, true, p->stream);
  }
  else
  {
    launch_gather(p->buf[0], p->buf[1], p->d_ctl, id, slot_list, base, r.nblk, p->block_doubles, st, false,
                  p->stream);
    G5_HIP(p->err, hipMemcpyAsync(host, st, need * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  }
  G5_HIP(p->err, hipGetLastError());
  G5_HIP(p->err, hipStreamSynchronize(p->stream));
  return upload ? clear_failure(p, id) : G5_OK;
}

// host-side description of the shock-bubble / triple-point initial conditions
void
region_cons(g5_config const & c, g5_region const & r, double * U)
{
  double g1[2], gp[2];
  for (int k = 0; k < 2; ++k)
  {
    g1[k] = 1.0 / (c.gamma[k] - 1.0);
    gp[k] = c.gamma[k] * c.pinf[k] / (c.gamma[k] - 1.0);
  }
  const int    dim = c.dim;
  const double rho = r.alpha_rho[0] + r.alpha_rho[1];
  const double a0 = r.alpha0, a1 = 1.0 - r.alpha0;
  const double e = (r.p * (a0 * g1[0] + a1 * g1[1]) + (a0 * gp[0] + a1 * gp[1])) / rho;
  const double ekin = dim == 2 ? 0.5 * (r.u * r.u + r.v * r.v) : 0.5 * (r.u * r.u + r.v * r.v + r.w * r.w);
  const int    IA0 = dim == 2 ? 4 : 5;
  for (int v = 0; v < 9; ++v)
    U[v] = 0.0;
  U[0] = rho;
  U[1] = (e + ekin) * rho;
  U[2] = r.u * rho;
  U[3] = r.v * rho;
  if (dim == 3)
    U[4] = r.w * rho;
  U[IA0] = r.alpha0;
  U[IA0 + 1] = r.alpha_rho[0];
  U[IA0 + 2] = 1.0 - r.alpha0;
  U[IA0 + 3] = r.alpha_rho[1];
}

void
fill_ic(g5_config const & c, g5_region const * regions3, IcDesc & ic)
{
  for (int r = 0; r < 3; ++r)
    region_cons(c, regions3[r], ic.S[r]);
  for (int d = 0; d < 3; ++d)
    ic.xmin[d] = c.xmin[d];
  const int dim = c.dim;
  auto      eint_of = [&](double const * s) {
    double ekin = s[2] * s[2] + s[3] * s[3];
    if (dim == 3)
      ekin += s[4] * s[4];
    ekin /= (2.0 * s[0]);
    return s[1] - ekin;
  };
  ic.eint_pre = eint_of(ic.S[1]);
  ic.eint_bub = eint_of(ic.S[2]);
}

int
pool_init_ic(g5_pool * p, int32_t id, IcDesc const & ic)
{
  if (id < 0 or id >= static_cast<int32_t>(p->solves.size()) or !p->solves[id].live)
    return fail(p->err, G5_ERR_INVALID, "no such solve");
  if (!p->solves[id].two_level.empty())
    return fail(p->err, G5_ERR_UNSUPPORTED, "device initial conditions fill one-level solves (upload the state)");
  IcDesc * d_ic = nullptr;
  G5_HIP(p->err, hipMalloc(&d_ic, sizeof(IcDesc)));
  G5_HIP(p->err, hipMemcpyAsync(d_ic, &ic, sizeof(IcDesc), hipMemcpyHostToDevice, p->stream));
  launch_init_ic(p->dim, p->buf[0], p->buf[1], p->d_ctl, id, p->d_solves, p->d_slots, d_ic, p->solves[id].nblk,
                 p->stream);
  G5_HIP(p->err, hipGetLastError());
  G5_HIP(p->err, hipStreamSynchronize(p->stream));
  hipFree(d_ic);
  return clear_failure(p, id);
}

int
read_ctl(g5_pool * p)
{
  const int32_t n = static_cast<int32_t>(p->solves.size());
  if (n > p->cap_hctl)
  {
    if (p->h_ctl)
      hipHostFree(p->h_ctl);
    p->h_ctl = nullptr;
    G5_HIP(p->err, hipHostMalloc(&p->h_ctl, std::max(n, 1) * sizeof(SolveCtl), hipHostMallocDefault));
    p->cap_hctl = std::max(n, 1);
  }
  if (n > 0)
    G5_HIP(p->err, hipMemcpyAsync(p->h_ctl, p->d_ctl, n * sizeof(SolveCtl), hipMemcpyDeviceToHost, p->stream));
  G5_HIP(p->err, hipStreamSynchronize(p->stream));
  return G5_OK;
}

// ---------------------------------------------------------------------------------------------
// Scratch of g5_step_noalloc: one solve's descriptor, control record (CFL max and failure bits),
// launch list and slot table, in a file-scope device array instead of hipMalloc'd memory.  Each
// GPU holds its own copy at its own address, found once per device with hipGetSymbolAddress;
// host code uses only those addresses (the array's name is a host-side stand-in).
// ---------------------------------------------------------------------------------------------
constexpr int64_t kNoallocMaxBlocks = 32768;
constexpr size_t  kNoallocCtlOff = (sizeof(SolveDesc) + 255) / 256 * 256;
constexpr size_t  kNoallocBlocksOff = kNoallocCtlOff + 256;
constexpr size_t  kNoallocSlotsOff = kNoallocBlocksOff + kNoallocMaxBlocks * sizeof(BlockDesc);
constexpr size_t  kNoallocBytes = kNoallocSlotsOff + kNoallocMaxBlocks * sizeof(int32_t);
static_assert(sizeof(SolveCtl) <= 256, "control record fits its 256-byte slot");

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

constexpr int kNoallocMaxDevices = 64;
std::mutex    g_noalloc_mutex;
void *        g_noalloc_ptr[kNoallocMaxDevices] = {};
// host staging of the launch list and the slot table (no heap use per call)
BlockDesc     g_noalloc_list[kNoallocMaxBlocks];
int32_t       g_noalloc_slots[kNoallocMaxBlocks];

// device address of this device's scratch (looked up on the device's first call)
int
noalloc_scratch(unsigned char ** out, std::string & err)
{
  int dev = 0;
  G5_HIP(err, hipGetDevice(&dev));
  if (dev < 0 or dev >= kNoallocMaxDevices)
    return fail(err, G5_ERR_UNSUPPORTED, "device index beyond the scratch table");
  if (g_noalloc_ptr[dev] == nullptr)
  {
    void * p = nullptr;
    G5_HIP(err, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr)
      return fail(err, G5_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_noalloc_ptr[dev] = p;
  }
  *out = static_cast<unsigned char *>(g_noalloc_ptr[dev]);
  return G5_OK;
}

} // namespace

extern "C"
{

const char *
g5_version(void)
{
  return "godunov_five_eq_amd 0.2 (gfx950, fused block step, pooled solves)";
}

const char *
g5_last_error(const g5_ctx * ctx)
{
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

// =============================================================================================
// pools
// =============================================================================================
int
g5_pool_create_rk2(int32_t dim, int32_t thinc_enabled, int32_t precision, uint64_t hbm_budget_bytes, g5_pool ** out)
{
  if (not out)
    return fail(g_create_error, G5_ERR_INVALID, "null argument");
  *out = nullptr;
  if (dim != 2 and dim != 3)
    return fail(g_create_error, G5_ERR_INVALID, "dim must be 2 or 3");
  if (precision != G5_PREC_FP64 and precision != G5_PREC_MIXED and precision != G5_PREC_FP64_FAST)
    return fail(g_create_error, G5_ERR_INVALID, "unknown precision");
  return pool_create(dim, thinc_enabled ? 1 : 0, precision, hbm_budget_bytes, true, out, g_create_error);
}

int
g5_pool_create(int32_t dim, int32_t thinc_enabled, int32_t precision, uint64_t hbm_budget_bytes, g5_pool ** out)
{
  if (not out)
    return fail(g_create_error, G5_ERR_INVALID, "null argument");
  *out = nullptr;
  if (dim != 2 and dim != 3)
    return fail(g_create_error, G5_ERR_INVALID, "dim must be 2 or 3");
  if (precision != G5_PREC_FP64 and precision != G5_PREC_MIXED and precision != G5_PREC_FP64_FAST)
    return fail(g_create_error, G5_ERR_INVALID, "unknown precision");
  return pool_create(dim, thinc_enabled ? 1 : 0, precision, hbm_budget_bytes, false, out, g_create_error);
}

int
g5_pool_destroy(g5_pool * p)
{
  if (not p)
    return G5_OK;
  if (p->stream)
    hipStreamSynchronize(p->stream);
  if (p->graph)
    (void)hipGraphExecDestroy(p->graph);
  if (p->graph_k)
    (void)hipGraphExecDestroy(p->graph_k);
  for (hipEvent_t e : p->ev)
    if (e)
      hipEventDestroy(e);
  for (hipEvent_t e : p->ev_poll)
    if (e)
      hipEventDestroy(e);
  if (p->h_poll)
    hipHostFree(p->h_poll);
  hipFree(p->buf[0]);
  hipFree(p->buf[1]);
  hipFree(p->stage);
  hipFree(p->d_solves);
  hipFree(p->d_ctl);
  hipFree(p->d_slots);
  hipFree(p->d_blocks);
  hipFree(p->d_cols);
  hipFree(p->d_start_list);
  hipFree(p->d_start_ids);
  hipFree(p->d_xfer);
  hipFree(p->d_iflux);
  hipFree(p->d_xslots);
  if (p->h_ctl)
    hipHostFree(p->h_ctl);
  if (p->stream)
    hipStreamDestroy(p->stream);
  delete p;
  return G5_OK;
}

const char *
g5_pool_last_error(const g5_pool * p)
{
  return p ? p->err.c_str() : g_create_error.c_str();
}

int
g5_pool_admit(g5_pool * p, const g5_config * cfg, double t_end, int32_t max_steps, int32_t * solve_id)
{
  if (not p or not cfg or not solve_id)
    return G5_ERR_INVALID;
  return pool_admit(p, *cfg, t_end, max_steps, solve_id);
}

int
g5_pool_admit_two_level(g5_pool * p, const g5_config * cfg, const int32_t * refined_coords, int32_t n_refined,
                        int32_t prolongation, double t_end, int32_t max_steps, int32_t * solve_id)
{
  if (not p or not cfg or not solve_id)
    return G5_ERR_INVALID;
  if (not refined_coords or n_refined < 1)
    return fail(p->err, G5_ERR_INVALID, "no refined blocks");
  const std::vector<int32_t> refined(refined_coords, refined_coords + cfg->dim * static_cast<size_t>(n_refined));
  return pool_admit(p, *cfg, t_end, max_steps, solve_id, &refined, prolongation);
}

int
g5_pool_block_layout(const g5_pool * p, int32_t id, int32_t * coords, size_t n, int32_t * n_blocks)
{
  if (not p or id < 0 or id >= static_cast<int32_t>(p->solves.size()) or !p->solves[id].live)
    return G5_ERR_INVALID;
  SolveRec const & r = p->solves[id];
  if (n_blocks)
    *n_blocks = r.nblk;
  if (not coords)
    return G5_OK;
  if (n != 4 * static_cast<size_t>(r.nblk))
    return G5_ERR_INVALID;
  const int nx = r.cfg.nb[0], ny = r.cfg.nb[1];
  for (int32_t k = 0; k < r.nblk; ++k)
  {
    int32_t * c = coords + 4 * static_cast<size_t>(k);
    if (!r.two_level.empty())
    {
      for (int d = 0; d < 4; ++d)
        c[d] = r.two_level[4 * k + d];
    }
    else
    {
      c[0] = k % nx;
      c[1] = (k / nx) % ny;
      c[2] = k / (nx * ny);
      c[3] = 0;
    }
  }
  return G5_OK;
}

int
g5_pool_release(g5_pool * p, int32_t id)
{
  if (not p)
    return G5_ERR_INVALID;
  return pool_release(p, id);
}

int
g5_pool_upload(g5_pool * p, int32_t id, const double * U, size_t n)
{
  if (not p or not U)
    return G5_ERR_INVALID;
  return pool_transfer(p, id, const_cast<double *>(U), n, true, nullptr);
}

int
g5_pool_download(g5_pool * p, int32_t id, double * U, size_t n)
{
  if (not p or not U)
    return G5_ERR_INVALID;
  return pool_transfer(p, id, U, n, false, nullptr);
}

int
g5_pool_init_shock_bubble(g5_pool * p, int32_t id, const g5_region * regions3, int32_t shape, double x_front,
                          const double * center3, const double * size3)
{
  if (not p or not regions3 or not center3 or not size3)
    return G5_ERR_INVALID;
  if (id < 0 or id >= static_cast<int32_t>(p->solves.size()) or !p->solves[id].live)
    return fail(p->err, G5_ERR_INVALID, "no such solve");
  IcDesc ic{};
  ic.kind = shape == 0 ? 0 : 1;
  ic.x_front = x_front;
  for (int d = 0; d < 3; ++d)
  {
    ic.center[d] = center3[d];
    ic.half[d] = size3[d];
  }
  ic.radius = size3[0];
  fill_ic(p->solves[id].cfg, regions3, ic);
  return pool_init_ic(p, id, ic);
}

int
g5_pool_init_triple_point(g5_pool * p, int32_t id, const g5_region * regions3, double xd, double yd)
{
  if (not p or not regions3)
    return G5_ERR_INVALID;
  if (id < 0 or id >= static_cast<int32_t>(p->solves.size()) or !p->solves[id].live)
    return fail(p->err, G5_ERR_INVALID, "no such solve");
  IcDesc ic{};
  ic.kind = 2;
  ic.x_front = xd;
  ic.yd = yd;
  fill_ic(p->solves[id].cfg, regions3, ic);
  return pool_init_ic(p, id, ic);
}

int
g5_pool_start(g5_pool * p, const int32_t * ids, int32_t n)
{
  if (not p or (n > 0 and not ids))
    return G5_ERR_INVALID;
  return pool_start(p, ids, n);
}

int
g5_pool_step(g5_pool * p, int32_t nsteps)
{
  if (not p or nsteps < 0)
    return G5_ERR_INVALID;
  return pool_step(p, nsteps, true);
}

int
g5_pool_step_async(g5_pool * p, int32_t nsteps)
{
  if (not p or nsteps < 0)
    return G5_ERR_INVALID;
  return pool_step(p, nsteps, false);
}

int
g5_pool_set_clock(g5_pool * p, int32_t id, double t, int32_t steps_done)
{
  if (not p)
    return G5_ERR_INVALID;
  if (id < 0 or id >= static_cast<int32_t>(p->solves.size()) or !p->solves[id].live)
    return fail(p->err, G5_ERR_INVALID, "no such solve");
  if (p->solves[id].started)
    return fail(p->err, G5_ERR_STATE, "the clock of a started solve cannot be set");
  if (steps_done < 0 or not std::isfinite(t))
    return fail(p->err, G5_ERR_INVALID, "t must be finite and steps_done >= 0");
  int rc = read_ctl(p);
  if (rc != G5_OK)
    return rc;
  SolveCtl c = p->h_ctl[id];
  c.t = t;
  c.steps = steps_done;
  G5_HIP(p->err, hipMemcpyAsync(p->d_ctl + id, &c, sizeof(SolveCtl), hipMemcpyHostToDevice, p->stream));
  G5_HIP(p->err, hipStreamSynchronize(p->stream));
  return G5_OK;
}

int
g5_pool_status(g5_pool * p, int32_t id, double * t, int32_t * steps, int32_t * active)
{
  if (not p)
    return G5_ERR_INVALID;
  if (id < 0 or id >= static_cast<int32_t>(p->solves.size()) or !p->solves[id].live)
    return fail(p->err, G5_ERR_INVALID, "no such solve");
  const int rc = read_ctl(p);
  if (rc != G5_OK)
    return rc;
  if (t)
    *t = p->h_ctl[id].t;
  if (steps)
    *steps = p->h_ctl[id].steps;
  if (active)
    *active = p->h_ctl[id].active;
  if (p->h_ctl[id].failure != 0u)
    return fail(p->err, G5_ERR_DIVERGED, failure_message(id, p->h_ctl[id]));
  return G5_OK;
}

int
g5_pool_failure(g5_pool * p, int32_t id, int32_t * failure)
{
  if (not p or not failure)
    return G5_ERR_INVALID;
  if (id < 0 or id >= static_cast<int32_t>(p->solves.size()) or !p->solves[id].live)
    return fail(p->err, G5_ERR_INVALID, "no such solve");
  const int rc = read_ctl(p);
  if (rc != G5_OK)
    return rc;
  *failure = static_cast<int32_t>(p->h_ctl[id].failure);
  return G5_OK;
}

int
g5_pool_finished(g5_pool * p, int32_t * ids, int32_t cap, int32_t * n_done)
{
  if (not p or not n_done)
    return G5_ERR_INVALID;
  const int rc = read_ctl(p);
  if (rc != G5_OK)
    return rc;
  int32_t n = 0;
  for (int32_t s = 0; s < static_cast<int32_t>(p->solves.size()); ++s)
    if (p->solves[s].live and p->solves[s].started and p->h_ctl[s].active == 0)
    {
      if (ids and n < cap)
        ids[n] = s;
      ++n;
    }
  *n_done = n;
  return G5_OK;
}

int
g5_pool_usage(const g5_pool * p, int64_t * used_blocks, int64_t * capacity_blocks, int32_t * live_solves)
{
  if (not p)
    return G5_ERR_INVALID;
  if (used_blocks)
    *used_blocks = p->nslots - static_cast<int64_t>(p->free_slots.size());
  if (capacity_blocks)
    *capacity_blocks = p->nslots;
  if (live_solves)
  {
    int32_t n = 0;
    for (auto const & r : p->solves)
      n += r.live ? 1 : 0;
    *live_solves = n;
  }
  return G5_OK;
}

int
g5_pool_last_step_ms(g5_pool * p, double * ms)
{
  if (not p or not ms)
    return G5_ERR_INVALID;
  const int rc = pool_timing(p);
  if (rc != G5_OK)
    return rc;
  *ms = p->last_run_ms;
  return G5_OK;
}

namespace
{
void
to_launch(LaunchCfg const & c, g5_pool const * p, g5_launch * out)
{
  out->wg_per_cu = c.wg_per_cu;
  out->xcd_remap = c.remap;
  out->zsplit = c.zsplit;
  out->ring_wave = c.ring_wave;
  out->col_len = c.col_len;
  out->autotuned = (!p->launch_fixed and p->autotune and p->tuned_class >= 0) ? 1 : 0;
  out->tuned_ms = p->tuned_ms;
}
int
set_launch(g5_pool * p, g5_launch const * l)
{
  if (l == nullptr)
  {
    p->launch_fixed = false;
    p->tuned_class = -1; // re-tune (or look up) at the next step
    return G5_OK;
  }
  if (l->wg_per_cu < 0 or l->wg_per_cu > 16 or (l->xcd_remap != 0 and l->xcd_remap != 1) or
      (l->zsplit != 0 and l->zsplit != 1) or (l->ring_wave != 0 and l->ring_wave != 1))
    return fail(p->err, G5_ERR_INVALID, "launch: wg_per_cu in [0,16], xcd_remap, zsplit and ring_wave in {0,1}");
  if (p->dim == 3 and l->wg_per_cu != 0)
    return fail(p->err, G5_ERR_UNSUPPORTED, "the 3-D step has no persistent form (wg_per_cu must be 0)");
  if (l->ring_wave != 0 and (p->dim != 2 or l->wg_per_cu != 0))
    return fail(p->err, G5_ERR_UNSUPPORTED, "ring_wave is a 2-D form with one workgroup per block");
  if (l->col_len < 1 or l->col_len > kMaxColumn or (l->col_len & (l->col_len - 1)) != 0)
    return fail(p->err, G5_ERR_INVALID, "launch: col_len in {1, 2, 4, 8}");
  if (l->col_len != 1 and p->dim != 3)
    return fail(p->err, G5_ERR_UNSUPPORTED, "col_len > 1 is a form of the 3-D step");
  p->launch = LaunchCfg{ l->wg_per_cu, l->xcd_remap, l->zsplit, l->ring_wave, p->dim == 3 ? l->col_len : 1 };
  p->launch_fixed = true;
  p->graph_valid = false;
  return G5_OK;
}
} // namespace

int
g5_pool_set_launch(g5_pool * p, const g5_launch * launch)
{
  if (not p)
    return G5_ERR_INVALID;
  return set_launch(p, launch);
}

int
g5_pool_get_launch(const g5_pool * p, g5_launch * launch)
{
  if (not p or not launch)
    return G5_ERR_INVALID;
  to_launch(p->launch, p, launch);
  return G5_OK;
}

int
g5_host_alloc(size_t bytes, void ** ptr)
{
  if (not ptr)
    return G5_ERR_INVALID;
  *ptr = nullptr;
  const hipError_t e = hipHostMalloc(ptr, bytes == 0 ? 1 : bytes, hipHostMallocDefault);
  if (e != hipSuccess)
    return fail(g_create_error, G5_ERR_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  return G5_OK;
}

int
g5_host_free(void * ptr)
{
  if (ptr)
    hipHostFree(ptr);
  return G5_OK;
}

int
g5_pool_prepare(g5_pool * p)
{
  if (not p)
    return G5_ERR_INVALID;
  int rc = rebuild_list(p);
  if (rc == G5_OK)
    rc = ensure_tuned(p, pool_hancock(p));
  if (rc == G5_OK)
    rc = pool_capture(p, true);
  if (rc == G5_OK)
    rc = pool_capture(p, false);
  return rc;
}

int
g5_pool_properties(const g5_pool * p, int32_t * dim, int32_t * thinc_enabled, int32_t * precision)
{
  if (not p)
    return G5_ERR_INVALID;
  if (dim)
    *dim = p->dim;
  if (thinc_enabled)
    *thinc_enabled = p->thinc;
  if (precision)
    *precision = p->precision;
  return G5_OK;
}

int
g5_pool_set_autotune(g5_pool * p, int32_t enable)
{
  if (not p)
    return G5_ERR_INVALID;
  p->autotune = enable != 0;
  p->tuned_class = -1;
  return G5_OK;
}

// =============================================================================================
// single solve: a pool of one whose slots are the lexicographic block order
// =============================================================================================
static int
create_ctx(const g5_config * cfg, std::vector<int32_t> const * refined, int prolong, g5_ctx ** out)
{
  if (not cfg or not out)
    return fail(g_create_error, G5_ERR_INVALID, "null argument");
  *out = nullptr;
  std::string msg;
  int         rc = validate(*cfg, msg);
  if (rc != G5_OK)
    return fail(g_create_error, rc, msg);
  g5_ctx * ctx = new g5_ctx();
  ctx->cfg = *cfg;
  for (int d = cfg->dim; d < 3; ++d)
    ctx->cfg.nb[d] = 1;
  int64_t nblk = nblocks_of(ctx->cfg);
  if (refined)
  {
    SolveRec tl;
    rc = layout_two_level(ctx->cfg, *refined, tl, msg);
    if (rc != G5_OK)
    {
      delete ctx;
      return fail(g_create_error, rc, msg);
    }
    nblk = tl.nblk;
  }
  const int      bd = (cfg->dim == 2 ? 8 * 256 : 9 * 4096);
  const bool     rk2 = cfg->integrator == G5_TI_RK2_SSP;
  const uint64_t budget = static_cast<uint64_t>(nblk) * bd * sizeof(double) * (rk2 ? 3 : 2);
  rc = pool_create(cfg->dim, cfg->thinc_enabled ? 1 : 0, cfg->precision, budget, rk2, &ctx->pool, g_create_error);
  if (rc != G5_OK)
  {
    delete ctx;
    return rc;
  }
  rc = pool_admit(ctx->pool, ctx->cfg, 1e300, 0x7fffffff, &ctx->id, refined, prolong);
  if (rc != G5_OK)
  {
    g_create_error = ctx->pool->err;
    g5_destroy(ctx);
    return rc;
  }
  // outside ring: blocks leaving the brick along at least one non-periodic direction
  const int nx = ctx->cfg.nb[0], ny = ctx->cfg.nb[1], nz = ctx->cfg.nb[2];
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
  *out = ctx;
  return G5_OK;
}

int
g5_create(const g5_config * cfg, g5_ctx ** out)
{
  return create_ctx(cfg, nullptr, 0, out);
}

int
g5_create_two_level(const g5_config * cfg, const int32_t * refined_coords, int32_t n_refined, int32_t prolongation,
                    g5_ctx ** out)
{
  if (not refined_coords or n_refined < 1)
    return fail(g_create_error, G5_ERR_INVALID, "no refined blocks");
  if (prolongation < 0 or prolongation > 3)
    return fail(g_create_error, G5_ERR_INVALID, "prolongation must be 0 (SIMPLE_COPY), 1 (EXTRAPOLATE_LINEAR_MINMOD), "
                                                "2 or 3 (CONSERVATIVE_INTERPOLATION_ORDER_2 / _4)");
  if (not cfg)
    return fail(g_create_error, G5_ERR_INVALID, "null argument");
  const std::vector<int32_t> refined(refined_coords, refined_coords + cfg->dim * static_cast<size_t>(n_refined));
  return create_ctx(cfg, &refined, prolongation, out);
}

int
g5_block_layout(const g5_ctx * ctx, int32_t * coords, size_t n)
{
  if (not ctx or not coords)
    return G5_ERR_INVALID;
  SolveRec const & r = ctx->pool->solves[ctx->id];
  if (n != 4 * static_cast<size_t>(r.nblk))
    return G5_ERR_INVALID;
  for (int32_t k = 0; k < r.nblk; ++k)
  {
    int32_t * c = coords + 4 * static_cast<size_t>(k);
    if (!r.two_level.empty())
    {
      for (int d = 0; d < 4; ++d)
        c[d] = r.two_level[4 * k + d];
    }
    else if (!ctx->block_order.empty())
    {
      for (int d = 0; d < 3; ++d)
        c[d] = ctx->block_order[3 * static_cast<size_t>(k) + d];
      c[3] = 0;
    }
    else
    {
      const int nx = ctx->cfg.nb[0], ny = ctx->cfg.nb[1];
      c[0] = k % nx;
      c[1] = (k / nx) % ny;
      c[2] = k / (nx * ny);
      c[3] = 0;
    }
  }
  return G5_OK;
}

int
g5_destroy(g5_ctx * ctx)
{
  if (not ctx)
    return G5_OK;
  if (ctx->pool)
  {
    hipStreamSynchronize(ctx->pool->stream);
    hipFree(ctx->d_outside);
    hipFree(ctx->d_ring);
    hipFree(ctx->d_staging);
    hipFree(ctx->d_flags);
    hipFree(ctx->d_imax);
    g5_pool_destroy(ctx->pool);
  }
  delete ctx;
  return G5_OK;
}

int
g5_shape(const g5_ctx * ctx, int32_t * n_blocks, int32_t * nvar, int32_t * cells_per_block)
{
  if (not ctx)
    return G5_ERR_INVALID;
  if (n_blocks)
    *n_blocks = ctx->pool->solves[ctx->id].nblk;
  if (nvar)
    *nvar = ctx->pool->nvar;
  if (cells_per_block)
    *cells_per_block = ctx->pool->ncell;
  return G5_OK;
}

int
g5_mem_bytes(const g5_ctx * ctx, uint64_t * bytes)
{
  if (not ctx or not bytes)
    return G5_ERR_INVALID;
  g5_pool const * p = ctx->pool;
  uint64_t        b = static_cast<uint64_t>(p->nslots) * p->block_doubles * sizeof(double) * (p->stage ? 3 : 2);
  b += p->cap_solves * (sizeof(SolveDesc) + sizeof(SolveCtl)) + p->cap_table * sizeof(int32_t) +
       p->cap_blocks * sizeof(BlockDesc);
  b += ctx->outside.size() * sizeof(int32_t);
  if (ctx->d_ring)
    b += ctx->outside.size() / 3 * p->block_doubles * sizeof(double);
  if (ctx->d_staging)
    b += static_cast<uint64_t>(p->solves[ctx->id].nblk) * p->block_doubles * sizeof(double);
  b += p->cap_iflux * sizeof(double);
  *bytes = b;
  return G5_OK;
}

// a context with the caller's block order: block k of the caller lives in slot k, so the state is
// the current buffer itself
static int
ctx_copy_slots(g5_ctx * ctx, double * host, size_t n, bool upload)
{
  g5_pool *    p = ctx->pool;
  const size_t need = static_cast<size_t>(p->solves[ctx->id].nblk) * p->block_doubles;
  if (n != need)
    return fail(p->err, G5_ERR_INVALID, "state size mismatch");
  int rc = read_ctl(p);
  if (rc != G5_OK)
    return rc;
  double * base = p->buf[p->h_ctl[ctx->id].parity];
  if (upload)
    G5_HIP(p->err, hipMemcpyAsync(base, host, need * sizeof(double), hipMemcpyHostToDevice, p->stream));
  else
    G5_HIP(p->err, hipMemcpyAsync(host, base, need * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  G5_HIP(p->err, hipStreamSynchronize(p->stream));
  return upload ? clear_failure(p, ctx->id) : G5_OK;
}

int
g5_upload_state(g5_ctx * ctx, const double * U_host, size_t n)
{
  if (not ctx or not U_host)
    return ctx ? fail(ctx->err, G5_ERR_INVALID, "null argument") : G5_ERR_INVALID;
  const int rc = ctx->block_order.empty()
                   ? pool_transfer(ctx->pool, ctx->id, const_cast<double *>(U_host), n, true, &ctx->d_staging)
                   : ctx_copy_slots(ctx, const_cast<double *>(U_host), n, true);
  if (rc != G5_OK)
    return fail(ctx->err, rc, ctx->pool->err);
  ctx->has_state = true;
  return G5_OK;
}

int
g5_download_state(g5_ctx * ctx, double * U_host, size_t n)
{
  if (not ctx or not U_host)
    return ctx ? fail(ctx->err, G5_ERR_INVALID, "null argument") : G5_ERR_INVALID;
  if (not ctx->has_state)
    return fail(ctx->err, G5_ERR_STATE, "no state uploaded");
  const int rc = ctx->block_order.empty() ? pool_transfer(ctx->pool, ctx->id, U_host, n, false, &ctx->d_staging)
                                          : ctx_copy_slots(ctx, U_host, n, false);
  return rc == G5_OK ? G5_OK : fail(ctx->err, rc, ctx->pool->err);
}

int
g5_state_device_ptr(g5_ctx * ctx, double ** U_dev)
{
  if (not ctx or not U_dev)
    return ctx ? fail(ctx->err, G5_ERR_INVALID, "null argument") : G5_ERR_INVALID;
  const int rc = read_ctl(ctx->pool);
  if (rc != G5_OK)
    return fail(ctx->err, rc, ctx->pool->err);
  // a pool of one uses slots 0..n-1, in lexicographic order or in the caller's block order
  // (g5_set_block_order): the buffer is the brick's U array in that order
  *U_dev = ctx->pool->buf[ctx->pool->h_ctl[ctx->id].parity];
  return G5_OK;
}

int
g5_compute_dt(g5_ctx * ctx, double * dt)
{
  if (not ctx or not dt)
    return ctx ? fail(ctx->err, G5_ERR_INVALID, "null argument") : G5_ERR_INVALID;
  if (not ctx->has_state)
    return fail(ctx->err, G5_ERR_STATE, "no state uploaded");
  g5_pool * p = ctx->pool;
  int       rc = rebuild_list(p);
  if (rc != G5_OK)
    return fail(ctx->err, rc, p->err);
  launch_reset_invdt(p->d_ctl, 1, p->stream);
  launch_inv_dt(p->dim, p->precision, p->buf[0], p->buf[1], p->d_blocks, p->d_solves, p->d_ctl, p->nlist, p->stream);
  G5_HIP(ctx->err, hipGetLastError());
  rc = read_ctl(p);
  if (rc != G5_OK)
    return fail(ctx->err, rc, p->err);
  double inv;
  std::memcpy(&inv, &p->h_ctl[0].invdt_bits, sizeof(double));
  *dt = ctx->cfg.cfl / inv;
  launch_reset_invdt(p->d_ctl, 1, p->stream);
  if (p->h_ctl[0].failure != 0u or not(inv > 0.0 and std::isfinite(inv)))
  {
    SolveCtl c = p->h_ctl[0];
    c.failure |= c.failure == 0u ? static_cast<uint32_t>(G5_FAIL_NONFINITE) : 0u;
    return fail(ctx->err, G5_ERR_DIVERGED, failure_message(ctx->id, c));
  }
  return G5_OK;
}

int
g5_fill_outside(g5_ctx * ctx, double * out_host, size_t n)
{
  if (not ctx)
    return G5_ERR_INVALID;
  if (not ctx->has_state)
    return fail(ctx->err, G5_ERR_STATE, "no state uploaded");
  g5_pool *    p = ctx->pool;
  const int    n_out = static_cast<int>(ctx->outside.size() / 3);
  const size_t need = static_cast<size_t>(n_out) * p->block_doubles;
  if (out_host and n != need)
    return fail(ctx->err, G5_ERR_INVALID, "ring size mismatch");
  if (n_out == 0)
    return G5_OK;
  if (not ctx->d_outside)
  {
    G5_HIP(ctx->err, hipMalloc(&ctx->d_outside, ctx->outside.size() * sizeof(int32_t)));
    G5_HIP(ctx->err, hipMemcpy(ctx->d_outside, ctx->outside.data(), ctx->outside.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice));
  }
  if (not ctx->d_ring)
    G5_HIP(ctx->err, hipMalloc(&ctx->d_ring, need * sizeof(double)));
  launch_fill_outside(p->dim, p->buf[0], p->buf[1], p->d_ctl, ctx->id, p->d_solves, p->d_slots, ctx->d_outside, n_out,
                      ctx->d_ring, p->stream);
  G5_HIP(ctx->err, hipGetLastError());
  if (out_host)
    G5_HIP(ctx->err, hipMemcpyAsync(out_host, ctx->d_ring, need * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  G5_HIP(ctx->err, hipStreamSynchronize(p->stream));
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
g5_set_block_order(g5_ctx * ctx, const int32_t * coords, int32_t n_blocks)
{
  if (not ctx or not coords)
    return ctx ? fail(ctx->err, G5_ERR_INVALID, "null argument") : G5_ERR_INVALID;
  g5_pool *       p = ctx->pool;
  SolveRec &      r = p->solves[ctx->id];
  const int       nx = ctx->cfg.nb[0], ny = ctx->cfg.nb[1], nz = ctx->cfg.nb[2];
  if (!r.two_level.empty())
    return fail(ctx->err, G5_ERR_UNSUPPORTED, "a two-level context keeps its own block order (g5_block_layout)");
  if (n_blocks != r.nblk)
    return fail(ctx->err, G5_ERR_INVALID, "n_blocks must equal the number of owned blocks");
  // caller block k goes to slot k: slots[brick index of block k] = k
  std::vector<int32_t> slots(r.nblk, -1);
  for (int32_t k = 0; k < n_blocks; ++k)
  {
    const int x = coords[3 * k], y = coords[3 * k + 1], z = ctx->cfg.dim == 3 ? coords[3 * k + 2] : 0;
    if (x < 0 or x >= nx or y < 0 or y >= ny or z < 0 or z >= nz)
      return fail(ctx->err, G5_ERR_INVALID, "block coordinates outside the brick");
    const int32_t lex = x + nx * (y + ny * z);
    if (slots[lex] >= 0)
      return fail(ctx->err, G5_ERR_INVALID, "block coordinates repeated");
    slots[lex] = k;
  }
  // the state moves with its blocks: gather the current state in the old order, scatter it back
  std::vector<double> U;
  if (ctx->has_state)
  {
    U.resize(static_cast<size_t>(r.nblk) * p->block_doubles);
    const int rc = pool_transfer(p, ctx->id, U.data(), U.size(), false, &ctx->d_staging);
    if (rc != G5_OK)
      return fail(ctx->err, rc, p->err);
  }
  // a context's slots are 0..n-1 (its budget holds exactly its blocks): a permutation of them
  r.slots = slots;
  G5_HIP(ctx->err, hipMemcpyAsync(p->d_slots + r.table_base, r.slots.data(), r.nblk * sizeof(int32_t),
                                  hipMemcpyHostToDevice, p->stream));
  G5_HIP(ctx->err, hipStreamSynchronize(p->stream));
  p->list_dirty = true;
  if (ctx->has_state)
  {
    const int rc = pool_transfer(p, ctx->id, U.data(), U.size(), true, &ctx->d_staging);
    if (rc != G5_OK)
      return fail(ctx->err, rc, p->err);
  }
  ctx->block_order.assign(coords, coords + 3 * static_cast<size_t>(n_blocks));
  return G5_OK;
}

// one godunov_unsplit iteration with a host dt
static int
host_step(g5_ctx * ctx, double dt)
{
  g5_pool * p = ctx->pool;
  int       rc = rebuild_list(p);
  if (rc != G5_OK)
    return fail(ctx->err, rc, p->err);
  if (ctx->cfg.integrator == G5_TI_RK2_SSP)
  {
    // stage 1: U_RK = U + dt L(U); stage 2: U2 = U_RK + dt L(U_RK); U = 1/2 U + 1/2 U2
    rc = read_ctl(p);
    if (rc != G5_OK)
      return fail(ctx->err, rc, p->err);
    const int par = p->h_ctl[ctx->id].parity;
    StepArgs  A = step_args(p, p->d_blocks, p->nlist, dt);
    A.U_out = p->stage;
    launch_step(p, A, false);
    A.U_in = p->stage;
    A.U_out = p->buf[1 - par];
    launch_step(p, A, false);
    launch_lincomb(0.5, p->buf[par], 0.5, p->buf[1 - par], static_cast<size_t>(p->nslots) * p->block_doubles,
                   p->stream);
  }
  else
  {
    launch_step(p, step_args(p, p->d_blocks, p->nlist, dt), ctx->cfg.integrator == G5_TI_HANCOCK);
  }
  launch_flip(p->d_ctl, 1, p->stream);
  G5_HIP(ctx->err, hipGetLastError());
  return G5_OK;
}

int
g5_do_time_step(g5_ctx * ctx, double dt)
{
  if (not ctx)
    return G5_ERR_INVALID;
  if (not ctx->has_state)
    return fail(ctx->err, G5_ERR_STATE, "no state uploaded");
  if (not(dt > 0.0))
    return fail(ctx->err, G5_ERR_INVALID, "dt must be > 0");
  return host_step(ctx, dt);
}

int
g5_do_time_step_buffers(g5_ctx * ctx, const double * U_dev, double * U2_dev, double dt)
{
  if (not ctx)
    return G5_ERR_INVALID;
  if (not U_dev or not U2_dev or U_dev == U2_dev)
    return fail(ctx->err, G5_ERR_INVALID, "U and U2 must be two distinct device arrays");
  if (not(dt > 0.0))
    return fail(ctx->err, G5_ERR_INVALID, "dt must be > 0");
  g5_pool * p = ctx->pool;
  int       rc = rebuild_list(p);
  if (rc != G5_OK)
    return fail(ctx->err, rc, p->err);
  StepArgs A = step_args(p, p->d_blocks, p->nlist, dt);
  A.U_in = U_dev;
  if (ctx->cfg.integrator == G5_TI_RK2_SSP)
  {
    // stage 1 into the stage buffer, stage 2 into U2, then U2 = 1/2 U + 1/2 U2
    A.U_out = p->stage;
    launch_step(p, A, false);
    A.U_in = p->stage;
    A.U_out = U2_dev;
    launch_step(p, A, false);
    launch_lincomb(0.5, U_dev, 0.5, U2_dev, static_cast<size_t>(p->nslots) * p->block_doubles, p->stream);
  }
  else
  {
    A.U_out = U2_dev;
    launch_step(p, A, ctx->cfg.integrator == G5_TI_HANCOCK);
  }
  G5_HIP(ctx->err, hipGetLastError());
  return G5_OK;
}

int
g5_compute_dt_buffer(g5_ctx * ctx, const double * U_dev, double * dt)
{
  if (not ctx or not dt)
    return ctx ? fail(ctx->err, G5_ERR_INVALID, "null argument") : G5_ERR_INVALID;
  if (not U_dev)
    return fail(ctx->err, G5_ERR_INVALID, "null device array");
  g5_pool * p = ctx->pool;
  int       rc = rebuild_list(p);
  if (rc != G5_OK)
    return fail(ctx->err, rc, p->err);
  static const uint32_t zero = 0u;
  launch_reset_invdt(p->d_ctl, 1, p->stream);
  G5_HIP(ctx->err, hipMemcpyAsync(reinterpret_cast<char *>(p->d_ctl) + offsetof(SolveCtl, failure), &zero,
                                  sizeof(uint32_t), hipMemcpyHostToDevice, p->stream));
  double * Ud = const_cast<double *>(U_dev); // read only: both buffer arguments are the caller's array
  launch_inv_dt(p->dim, p->precision, Ud, Ud, p->d_blocks, p->d_solves, p->d_ctl, p->nlist, p->stream);
  G5_HIP(ctx->err, hipGetLastError());
  rc = read_ctl(p);
  if (rc != G5_OK)
    return fail(ctx->err, rc, p->err);
  double inv;
  std::memcpy(&inv, &p->h_ctl[0].invdt_bits, sizeof(double));
  *dt = ctx->cfg.cfl / inv;
  SolveCtl c = p->h_ctl[0];
  launch_reset_invdt(p->d_ctl, 1, p->stream);
  G5_HIP(ctx->err, hipMemcpyAsync(reinterpret_cast<char *>(p->d_ctl) + offsetof(SolveCtl, failure), &zero,
                                  sizeof(uint32_t), hipMemcpyHostToDevice, p->stream));
  G5_HIP(ctx->err, hipStreamSynchronize(p->stream));
  if (c.failure != 0u or not(inv > 0.0 and std::isfinite(inv)))
  {
    c.failure |= c.failure == 0u ? static_cast<uint32_t>(G5_FAIL_NONFINITE) : 0u;
    return fail(ctx->err, G5_ERR_DIVERGED, failure_message(ctx->id, c));
  }
  return G5_OK;
}

int
g5_refine_flags(g5_ctx * ctx, const g5_refine_params * prm, int32_t * flags, double * indicator_max, size_t n)
{
  if (not ctx or not prm or not flags)
    return ctx ? fail(ctx->err, G5_ERR_INVALID, "null argument") : G5_ERR_INVALID;
  if (not ctx->has_state)
    return fail(ctx->err, G5_ERR_STATE, "no state uploaded");
  g5_pool *     p = ctx->pool;
  const int32_t nblk = p->solves[ctx->id].nblk;
  if (n != static_cast<size_t>(nblk))
    return fail(ctx->err, G5_ERR_INVALID, "n must equal the number of blocks");
  if (prm->indicator < G5_INDICATOR_LOHNER_SPLIT or prm->indicator > G5_INDICATOR_SIMPLE_GRADIENT)
    return fail(ctx->err, G5_ERR_INVALID, "unknown refine indicator");
  int rc = rebuild_list(p);
  if (rc != G5_OK)
    return fail(ctx->err, rc, p->err);
  // a context's slots are its block order: block k lives in slot k
  if (not ctx->d_flags)
  {
    G5_HIP(ctx->err, hipMalloc(&ctx->d_flags, p->nslots * sizeof(int32_t)));
    G5_HIP(ctx->err, hipMalloc(&ctx->d_imax, p->nslots * sizeof(double)));
  }
  G5_HIP(ctx->err, hipMemcpyAsync(ctx->d_flags, flags, nblk * sizeof(int32_t), hipMemcpyHostToDevice, p->stream));
  G5_HIP(ctx->err, hipMemsetAsync(ctx->d_imax, 0, nblk * sizeof(double), p->stream));
  RefineArgs R;
  R.indicator = prm->indicator;
  R.level_min = prm->level_min;
  R.level_max = prm->level_max;
  R.base_level = prm->base_level;
  R.refine_th = prm->refine_threshold;
  R.coarsen_th = prm->coarsen_threshold;
  R.eps = prm->epsilon_lohner;
  R.flags = ctx->d_flags;
  R.imax = ctx->d_imax;
  launch_refine_flags(p->dim, p->buf[0], p->buf[1], p->d_blocks, p->d_solves, p->d_ctl, p->d_slots, R, p->nlist,
                      p->stream);
  G5_HIP(ctx->err, hipGetLastError());
  G5_HIP(ctx->err, hipMemcpyAsync(flags, ctx->d_flags, nblk * sizeof(int32_t), hipMemcpyDeviceToHost, p->stream));
  if (indicator_max)
    G5_HIP(ctx->err,
           hipMemcpyAsync(indicator_max, ctx->d_imax, nblk * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  G5_HIP(ctx->err, hipStreamSynchronize(p->stream));
  return G5_OK;
}

int
g5_run(g5_ctx * ctx, int32_t nsteps, double * t, double tEnd, int32_t * steps_done)
{
  if (not ctx or not t)
    return ctx ? fail(ctx->err, G5_ERR_INVALID, "null argument") : G5_ERR_INVALID;
  if (not ctx->has_state)
    return fail(ctx->err, G5_ERR_STATE, "no state uploaded");
  if (nsteps < 0)
    return fail(ctx->err, G5_ERR_INVALID, "nsteps < 0");
  g5_pool * p = ctx->pool;
  int       rc = read_ctl(p);
  if (rc != G5_OK)
    return fail(ctx->err, rc, p->err);
  SolveCtl c0{};
  c0.t = *t;
  c0.tend = tEnd;
  c0.max_steps = nsteps;
  c0.parity = p->h_ctl[ctx->id].parity;
  G5_HIP(ctx->err, hipMemcpyAsync(p->d_ctl + ctx->id, &c0, sizeof(SolveCtl), hipMemcpyHostToDevice, p->stream));
  rc = pool_start(p, &ctx->id, 1);
  if (rc == G5_OK)
    rc = pool_run_solve(p, ctx->id, nsteps);
  if (rc == G5_OK)
    rc = read_ctl(p);
  if (rc != G5_OK)
    return fail(ctx->err, rc, p->err);
  *t = p->h_ctl[ctx->id].t;
  if (steps_done)
    *steps_done = p->h_ctl[ctx->id].steps;
  if (p->h_ctl[ctx->id].failure != 0u)
    return fail(ctx->err, G5_ERR_DIVERGED, failure_message(ctx->id, p->h_ctl[ctx->id]));
  return G5_OK;
}

int
g5_last_run_ms(const g5_ctx * ctx, double * ms)
{
  if (not ctx or not ms)
    return G5_ERR_INVALID;
  *ms = ctx->pool->last_run_ms;
  return G5_OK;
}

int
g5_synchronize(g5_ctx * ctx)
{
  if (not ctx)
    return G5_ERR_INVALID;
  G5_HIP(ctx->err, hipStreamSynchronize(ctx->pool->stream));
  G5_HIP(ctx->err, hipGetLastError());
  return G5_OK;
}

int
g5_time_step_kernel(g5_ctx * ctx, int32_t nsteps, double dt, double * mean_ms)
{
  if (not ctx or not mean_ms)
    return ctx ? fail(ctx->err, G5_ERR_INVALID, "null argument") : G5_ERR_INVALID;
  if (not ctx->has_state)
    return fail(ctx->err, G5_ERR_STATE, "no state uploaded");
  if (nsteps <= 0 or not(dt > 0.0))
    return fail(ctx->err, G5_ERR_INVALID, "nsteps and dt must be > 0");
  g5_pool * p = ctx->pool;
  int       rc = rebuild_list(p);
  if (rc != G5_OK)
    return fail(ctx->err, rc, p->err);
  hipEvent_t e0, e1;
  G5_HIP(ctx->err, hipEventCreate(&e0));
  G5_HIP(ctx->err, hipEventCreate(&e1));
  const bool hancock = ctx->cfg.integrator == G5_TI_HANCOCK;
  G5_HIP(ctx->err, hipEventRecord(e0, p->stream));
  for (int n = 0; n < nsteps; ++n)
  {
    launch_step(p, step_args(p, p->d_blocks, p->nlist, dt), hancock);
    launch_flip(p->d_ctl, 1, p->stream);
  }
  G5_HIP(ctx->err, hipEventRecord(e1, p->stream));
  G5_HIP(ctx->err, hipEventSynchronize(e1));
  float ms = 0.f;
  G5_HIP(ctx->err, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *mean_ms = static_cast<double>(ms) / nsteps;
  return G5_OK;
}


int
g5_step_noalloc(const g5_config * cfg, const double * U_dev, double * U2_dev, double dt, double * dt_next,
                int32_t * failure)
{
  std::lock_guard<std::mutex> lk(g_noalloc_mutex);
  if (not cfg or not U_dev or not U2_dev or not dt_next or U_dev == U2_dev)
    return fail(g_create_error, G5_ERR_INVALID, "g5_step_noalloc: null or aliased argument");
  std::string msg;
  int         rc = validate(*cfg, msg);
  if (rc != G5_OK)
    return fail(g_create_error, rc, msg);
  if (cfg->integrator == G5_TI_RK2_SSP)
    return fail(g_create_error, G5_ERR_UNSUPPORTED,
                "g5_step_noalloc: RK2_SSP needs a stage buffer of the whole state (use g5_do_time_step_buffers)");
  if (not(dt > 0.0))
    return fail(g_create_error, G5_ERR_INVALID, "dt must be > 0");
  g5_config c = *cfg;
  for (int d = c.dim; d < 3; ++d)
    c.nb[d] = 1;
  const int64_t nblk = nblocks_of(c);
  if (nblk > kNoallocMaxBlocks)
    return fail(g_create_error, G5_ERR_TOO_LARGE,
                "g5_step_noalloc: " + std::to_string(nblk) + " blocks, beyond the scratch maximum of " +
                  std::to_string(kNoallocMaxBlocks));
  unsigned char * base = nullptr;
  rc = noalloc_scratch(&base, g_create_error);
  if (rc != G5_OK)
    return rc;
  // descriptor, cleared control record (CFL max 0, no failure), brick-ordered launch list and
  // identity slot table (block k of U in slot k)
  SolveDesc sd;
  make_solve_desc(c, 0, sd);
  SolveCtl ctl{};
  int64_t  k = 0;
  for (int z = 0; z < c.nb[2]; ++z)
    for (int y = 0; y < c.nb[1]; ++y)
      for (int x = 0; x < c.nb[0]; ++x, ++k)
      {
        g_noalloc_list[k] = BlockDesc{ 0, static_cast<int32_t>(k), static_cast<int16_t>(x), static_cast<int16_t>(y),
                                       static_cast<int16_t>(z), 0 };
        g_noalloc_slots[k] = static_cast<int32_t>(k);
      }
  hipStream_t st = nullptr; // the null stream: the call is ordered after the caller's blocking work
  G5_HIP(g_create_error, hipMemcpyAsync(base, &sd, sizeof(SolveDesc), hipMemcpyHostToDevice, st));
  G5_HIP(g_create_error, hipMemcpyAsync(base + kNoallocCtlOff, &ctl, sizeof(SolveCtl), hipMemcpyHostToDevice, st));
  G5_HIP(g_create_error, hipMemcpyAsync(base + kNoallocBlocksOff, g_noalloc_list, nblk * sizeof(BlockDesc),
                                        hipMemcpyHostToDevice, st));
  G5_HIP(g_create_error, hipMemcpyAsync(base + kNoallocSlotsOff, g_noalloc_slots, nblk * sizeof(int32_t),
                                        hipMemcpyHostToDevice, st));
  StepArgs A;
  A.buf0 = nullptr;
  A.buf1 = nullptr;
  A.U_in = U_dev;
  A.U_out = U2_dev;
  A.blocks = reinterpret_cast<BlockDesc const *>(base + kNoallocBlocksOff);
  A.solves = reinterpret_cast<SolveDesc const *>(base);
  A.ctl = reinterpret_cast<SolveCtl *>(base + kNoallocCtlOff);
  A.slots = reinterpret_cast<int32_t const *>(base + kNoallocSlotsOff);
  A.nblocks = static_cast<int32_t>(nblk);
  A.per_xcd = static_cast<int32_t>((nblk + 7) / 8);
  A.wg_per_xcd = 0;
  A.remap = 1;
  A.dt_fixed = dt;
  A.iflux = nullptr; // one level
  A.col_stride = 1;  // one block per workgroup
  A.col_len = 1;
  const bool hancock = c.integrator == G5_TI_HANCOCK;
  const bool mp = c.source_term == G5_SOURCE_MILLER_PUCKETT;
  if (c.dim == 2)
    launch_step2d(A, c.thinc_enabled != 0, hancock, c.precision, false, mp, st);
  else
    launch_step3d(A, c.thinc_enabled != 0, hancock, c.precision, true, c.inlet != G5_INLET_NONE, mp, st);
  G5_HIP(g_create_error, hipGetLastError());
  G5_HIP(g_create_error, hipMemcpy(&ctl, base + kNoallocCtlOff, sizeof(SolveCtl), hipMemcpyDeviceToHost));
  double inv;
  std::memcpy(&inv, &ctl.invdt_bits, sizeof(double));
  *dt_next = c.cfl / inv;
  if (failure)
    *failure = static_cast<int32_t>(ctl.failure);
  if (ctl.failure != 0u or not(inv > 0.0 and std::isfinite(inv)))
  {
    ctl.failure |= ctl.failure == 0u ? static_cast<uint32_t>(G5_FAIL_NONFINITE) : 0u;
    if (failure)
      *failure = static_cast<int32_t>(ctl.failure);
    return fail(g_create_error, G5_ERR_DIVERGED, failure_message(0, ctl));
  }
  return G5_OK;
}

} // extern "C"
