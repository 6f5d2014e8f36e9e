// g5_serve.cpp -- admission of a workload spec into a pool (device initial condition) and the
// serving request loop (continuous batching over a 2-D and a 3-D pool).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/godunov_five_eq.h"

extern "C" {

int
g5_pool_admit_spec(g5_pool * pool, const g5_solve_spec * s, int32_t * solve_id)
{
  if (not pool or not s or not solve_id)
    return G5_ERR_INVALID;
  int32_t dim = 0, thinc = 0, precision = 0;
  int     rc = g5_pool_properties(pool, &dim, &thinc, &precision);
  if (rc != G5_OK)
    return rc;
  if (s->dim != dim)
    return G5_ERR_INVALID;
  g5_config cfg;
  g5_region reg[3];
  if ((rc = g5_spec_config(s, thinc, precision, &cfg)) != G5_OK or (rc = g5_spec_regions(s, reg)) != G5_OK)
    return rc;
  int32_t id = -1;
  rc = g5_pool_admit(pool, &cfg, 1e300, s->steps, &id);
  if (rc != G5_OK)
    return rc;
  if (s->kind == 1)
    rc = g5_pool_init_triple_point(pool, id, reg, 1.0 / 3.0, 0.5);
  else
  {
    // inclusion centred at (0.55, 0.5, 0.5): circle/sphere of radius 0.25 or box of half-sizes
    // (0.2, 0.3, 0.3); the shock starts at x = 0.2
    const double center[3] = { 0.55, 0.5, 0.5 };
    const double size_circle[3] = { 0.25, 0.25, 0.25 };
    const double size_box[3] = { 0.2, 0.3, 0.3 };
    rc = g5_pool_init_shock_bubble(pool, id, reg, s->shape, 0.2, center, s->shape ? size_box : size_circle);
  }
  if (rc != G5_OK)
  {
    g5_pool_release(pool, id);
    return rc;
  }
  *solve_id = id;
  return G5_OK;
}

// ---------------------------------------------------------------------------------------------
// Serving loop: requests arrive on a Poisson clock; each is admitted into the pool of its
// dimension as soon as the pool's HBM share holds it (first come, first served), its initial
// condition is generated on the device, and it is stepped together with every other solve in
// flight; a finished solve is downloaded (the result) and released, which frees its memory for
// the queue.  Latency = result on the host - arrival.  A request with more blocks than its
// pool's whole budget is rejected at once (it could never be admitted and would block its
// dimension's queue); a solve stopped by a failure is collected like a finished one and counted
// as failed.
// ---------------------------------------------------------------------------------------------
int
g5_serve_run(const g5_serve_config * c, const g5_solve_spec * given, g5_serve_report * r, g5_serve_result * results,
             double * const * states)
{
  if (not c or not r or c->n_requests <= 0 or c->steps_per_poll <= 0 or c->share_3d < 0.0 or c->share_3d > 1.0)
    return G5_ERR_INVALID;
  std::memset(r, 0, sizeof(*r));
  std::vector<g5_solve_spec> specs(c->n_requests);
  int                        rc = G5_OK;
  if (given)
  {
    for (int32_t k = 0; k < c->n_requests; ++k)
    {
      specs[k] = given[k];
      if ((specs[k].dim != 2 and specs[k].dim != 3) or (k > 0 and specs[k].arrival < specs[k - 1].arrival))
        return G5_ERR_INVALID; // dimension 2 or 3, arrivals in order
    }
  }
  else
    rc = g5_serve_specs(c->seed, c->n_requests, c->rate_hz, c->frac_3d, specs.data());
  if (rc != G5_OK)
    return rc;
  bool need[2] = { false, false };
  for (auto const & s : specs)
    need[s.dim == 2 ? 0 : 1] = true;
  g5_pool *      pools[2] = { nullptr, nullptr };
  const uint64_t budget3 = need[0] ? static_cast<uint64_t>(static_cast<double>(c->hbm_budget_bytes) * c->share_3d)
                                   : c->hbm_budget_bytes;
  const uint64_t budget2 = c->hbm_budget_bytes - (need[1] ? budget3 : 0);
  if (need[0])
    rc = g5_pool_create(2, c->thinc_enabled, c->precision, budget2, &pools[0]);
  if (rc == G5_OK and need[1])
    rc = g5_pool_create(3, c->thinc_enabled, c->precision, budget3, &pools[1]);
  if (rc != G5_OK)
  {
    g5_pool_destroy(pools[0]);
    g5_pool_destroy(pools[1]);
    return rc;
  }
  for (g5_pool * p : pools)
    if (p)
      g5_pool_set_autotune(p, c->autotune);
  struct Live
  {
    int32_t req;
    int32_t id;
  };
  std::vector<Live>            live[2];
  std::vector<int32_t>         queue; // arrived, not admitted
  std::vector<g5_serve_result> res(c->n_requests);
  double *                     host_result = nullptr; // pinned: results come back at full PCIe rate
  size_t                       host_cap = 0;
  size_t                       next = 0;
  int32_t                      done = 0; // completed, failed or rejected
  double                       cell_updates = 0.0;
  int32_t                      max_inflight = 0;
  double                       step_ms[2] = { 0.0, 0.0 };
  double                       host_s[4] = { 0.0, 0.0, 0.0, 0.0 }; // admission, start, wait for steps, collection
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  auto       now = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };

  while (done < c->n_requests and rc == G5_OK)
  {
    const double tn = now();
    while (next < specs.size() and specs[next].arrival <= tn)
      queue.push_back(static_cast<int32_t>(next++));
    double ph = now();
    // admission, first come first served per dimension
    std::vector<int32_t> started[2];
    bool                 blocked[2] = { false, false };
    std::vector<int32_t> rest;
    for (size_t q = 0; q < queue.size(); ++q)
    {
      const int32_t req = queue[q];
      const int     k = specs[req].dim == 2 ? 0 : 1;
      if (blocked[k])
      {
        rest.push_back(req);
        continue;
      }
      int32_t   id = -1;
      const int a = g5_pool_admit_spec(pools[k], &specs[req], &id);
      if (a == G5_ERR_TOO_LARGE)
      {
        res[req].outcome = G5_SERVE_REJECTED;
        res[req].latency_ms = (now() - specs[req].arrival) * 1e3;
        ++r->rejected;
        ++done;
        continue;
      }
      if (a == G5_ERR_NOMEM)
      {
        blocked[k] = true; // keep the arrival order within a dimension
        rest.push_back(req);
        continue;
      }
      if (a != G5_OK)
      {
        rc = a;
        break;
      }
      live[k].push_back({ req, id });
      started[k].push_back(id);
    }
    queue.swap(rest);
    host_s[0] += now() - ph;
    ph = now();
    for (int k = 0; k < 2 and rc == G5_OK; ++k)
      if (!started[k].empty())
        rc = g5_pool_start(pools[k], started[k].data(), static_cast<int32_t>(started[k].size()));
    host_s[1] += now() - ph;
    const int32_t inflight = static_cast<int32_t>(live[0].size() + live[1].size());
    max_inflight = std::max(max_inflight, inflight);
    if (inflight == 0)
    {
      // nothing in flight: a blocked queue would be a request too large for an empty pool, which
      // admission rejects, so only future arrivals are left
      if (next < specs.size())
      {
        const double wait = specs[next].arrival - now();
        if (wait > 0.0)
          std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
      }
      continue;
    }
    // step everything in flight: both pools' steps are enqueued before either is waited for
    bool stepped[2] = { false, false };
    for (int k = 0; k < 2 and rc == G5_OK; ++k)
      if (!live[k].empty())
      {
        rc = g5_pool_step_async(pools[k], c->steps_per_poll);
        stepped[k] = rc == G5_OK;
      }
    ph = now();
    for (int k = 0; k < 2 and rc == G5_OK; ++k)
      if (stepped[k])
      {
        double ms = 0.0;
        rc = g5_pool_last_step_ms(pools[k], &ms);
        step_ms[k] += ms;
      }
    host_s[2] += now() - ph;
    ph = now();
    // collect finished (or failed) solves
    for (int k = 0; k < 2 and rc == G5_OK; ++k)
    {
      if (live[k].empty())
        continue;
      std::vector<int32_t> fin(live[k].size());
      int32_t              nf = 0;
      rc = g5_pool_finished(pools[k], fin.data(), static_cast<int32_t>(fin.size()), &nf);
      for (int32_t f = 0; f < nf and rc == G5_OK; ++f)
      {
        auto it = std::find_if(live[k].begin(), live[k].end(), [&](Live const & l) { return l.id == fin[f]; });
        if (it == live[k].end())
          continue;
        g5_solve_spec const & s = specs[it->req];
        g5_serve_result &     out = res[it->req];
        size_t                n = 0;
        g5_spec_state_size(&s, &n);
        if (n > host_cap)
        {
          g5_host_free(host_result);
          host_result = nullptr;
          host_cap = n;
          rc = g5_host_alloc(host_cap * sizeof(double), reinterpret_cast<void **>(&host_result));
          if (rc != G5_OK)
            break;
        }
        rc = g5_pool_download(pools[k], it->id, host_result, n);
        if (rc != G5_OK)
          break;
        out.latency_ms = (now() - s.arrival) * 1e3;
        int32_t active = 0;
        rc = g5_pool_status(pools[k], it->id, &out.t, &out.steps, &active);
        if (rc == G5_ERR_DIVERGED)
        {
          rc = g5_pool_failure(pools[k], it->id, &out.failure);
          out.outcome = G5_SERVE_FAILED;
          ++r->failed;
        }
        else if (rc == G5_OK)
        {
          out.outcome = G5_SERVE_DONE;
          cell_updates += static_cast<double>(n / (s.dim == 2 ? 8 : 9)) * out.steps;
        }
        if (rc != G5_OK)
          break;
        if (states and states[it->req])
          std::memcpy(states[it->req], host_result, n * sizeof(double));
        rc = g5_pool_release(pools[k], it->id);
        live[k].erase(it);
        ++done;
      }
    }
    host_s[3] += now() - ph;
  }
  const double wall = now();
  g5_host_free(host_result);
  for (int k = 0; k < 2; ++k)
    g5_pool_destroy(pools[k]);
  if (rc != G5_OK)
    return rc;
  // latency percentiles over the completed requests
  std::vector<double> sorted;
  double              sum = 0.0;
  for (auto const & o : res)
    if (o.outcome == G5_SERVE_DONE)
    {
      sorted.push_back(o.latency_ms);
      sum += o.latency_ms;
    }
  std::sort(sorted.begin(), sorted.end());
  auto pct = [&](double q) {
    if (sorted.empty())
      return 0.0;
    const size_t i = static_cast<size_t>(std::ceil(q * sorted.size())) - 1;
    return sorted[std::min(i, sorted.size() - 1)];
  };
  r->completed = static_cast<int32_t>(sorted.size());
  r->max_inflight = max_inflight;
  r->wall_s = wall;
  r->cell_updates = cell_updates;
  r->cell_updates_per_s = cell_updates / wall;
  r->p50_ms = pct(0.50);
  r->p99_ms = pct(0.99);
  r->max_ms = sorted.empty() ? 0.0 : sorted.back();
  r->mean_ms = sorted.empty() ? 0.0 : sum / sorted.size();
  r->last_arrival_s = specs.back().arrival;
  r->step_s_2d = step_ms[0] * 1e-3;
  r->step_s_3d = step_ms[1] * 1e-3;
  r->admit_s = host_s[0];
  r->start_s = host_s[1];
  r->wait_s = host_s[2];
  r->collect_s = host_s[3];
  if (results)
    std::copy(res.begin(), res.end(), results);
  return G5_OK;
}

int
g5_serve(const g5_serve_config * c, g5_serve_report * r, double * latencies_ms, size_t cap)
{
  if (not c)
    return G5_ERR_INVALID;
  std::vector<g5_serve_result> res(c->n_requests > 0 ? c->n_requests : 0);
  const int                    rc = g5_serve_run(c, nullptr, r, res.data(), nullptr);
  if (rc == G5_OK and latencies_ms)
    for (size_t k = 0; k < std::min(cap, res.size()); ++k)
      latencies_ms[k] = res[k].latency_ms;
  return rc;
}

} // extern "C"
