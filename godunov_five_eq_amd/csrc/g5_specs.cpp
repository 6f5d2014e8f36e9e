// g5_specs.cpp -- seeded workloads of the batching and serving forms (host code only).
//
// A solve spec describes one independent 2-D or 3-D run: grid, initial-condition family
// (shock-bubble with a circle/sphere or rectangle/box inclusion, or triple point), the two
// stiffened-gas materials, the shock Mach number and the step budget.  Specs are drawn from
// std::mt19937_64 (whose output sequence the C++ standard fixes) through this file's own
// uniform mappings, so a (seed, index) pair names the same solve on every platform.
//
// Initial conditions follow the reference's generators: init/InitShockBubble.cpp:50-225
// (post-shock region for x <= x_front, pre-shock region, inclusion of material 0) and
// init/InitTriplePoint.cpp:46-133 (left region for x <= xd, right regions split at y = yd).
// The post-shock state comes from the stiffened-gas Rankine-Hugoniot relations
// (g5_rankine_hugoniot, app/tools/moving_shock.py:79-163 for the ideal-gas form).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/godunov_five_eq.h"

namespace
{

// uniform double in [0, 1) from the top 53 bits
double
u01(std::mt19937_64 & g)
{
  return static_cast<double>(g() >> 11) * 0x1.0p-53;
}
double
uniform(std::mt19937_64 & g, double a, double b)
{
  return a + (b - a) * u01(g);
}
// uniform integer in [0, n)
int64_t
below(std::mt19937_64 & g, int64_t n)
{
  return static_cast<int64_t>(u01(g) * static_cast<double>(n));
}
// uniform multiple of 16 in [lo, hi] (both multiples of 16)
int32_t
side16(std::mt19937_64 & g, int32_t lo, int32_t hi)
{
  return lo + 16 * static_cast<int32_t>(below(g, (hi - lo) / 16 + 1));
}

std::mt19937_64
stream_of(uint64_t seed, uint64_t index)
{
  // one generator per (seed, index): splitmix64 of the pair
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (index + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return std::mt19937_64(z);
}

// materials, Mach number and IC family shared by both workloads
void
draw_physics(std::mt19937_64 & g, g5_solve_spec & s)
{
  s.kind = static_cast<int32_t>(below(g, 2));
  s.shape = static_cast<int32_t>(below(g, 2));
  s.gamma[1] = uniform(g, 1.1, 4.4); // surrounding material (carries the shock)
  s.pinf[1] = below(g, 2) ? 6e8 : 0.0;
  s.gamma[0] = uniform(g, 1.1, 4.4); // inclusion (gas)
  s.pinf[0] = 0.0;
  s.mach = uniform(g, 1.1, 3.0);
  if (s.pinf[1] > 0.0)
  {
    // liquid-like surrounding material (water: rho 1000 kg/m^3, 1 bar) with a gas inclusion
    s.rho1 = 1000.0;
    s.p1 = 1e5;
    s.rho0 = uniform(g, 1.0, 10.0);
  }
  else
  {
    // gas-gas (normalised units, as the air-helium shock-bubble)
    s.rho1 = 1.0;
    s.p1 = 1.0;
    s.rho0 = uniform(g, 0.1, 5.0);
  }
}

} // namespace

extern "C" {

int
g5_batch_spec(uint64_t seed, int32_t index, g5_solve_spec * out)
{
  if (not out or index < 0)
    return G5_ERR_INVALID;
  std::memset(out, 0, sizeof(*out));
  std::mt19937_64 g = stream_of(seed, static_cast<uint64_t>(index));
  static const int32_t sides[4] = { 64, 128, 256, 512 };
  const int32_t        side = sides[below(g, 4)];
  const double         aspect = uniform(g, 1.0, 4.0);
  const int32_t        lng = std::max<int32_t>(side, 16 * static_cast<int32_t>(std::lround(side * aspect / 16.0)));
  out->dim = 2;
  out->nb[0] = lng / 16;
  out->nb[1] = side / 16;
  out->nb[2] = 1;
  out->steps = 100;
  draw_physics(g, *out);
  out->arrival = 0.0;
  return G5_OK;
}

int
g5_serve_specs(uint64_t seed, int32_t n, double rate_hz, double frac_3d, g5_solve_spec * out)
{
  if (not out or n < 0 or not(rate_hz > 0.0) or frac_3d < 0.0 or frac_3d > 1.0)
    return G5_ERR_INVALID;
  std::mt19937_64 arrivals = stream_of(seed, 0xA11ull << 40);
  double          t = 0.0;
  for (int32_t k = 0; k < n; ++k)
  {
    g5_solve_spec & s = out[k];
    std::memset(&s, 0, sizeof(s));
    // Poisson process: exponential inter-arrival times of mean 1/rate
    t += -std::log(1.0 - u01(arrivals)) / rate_hz;
    s.arrival = t;
    std::mt19937_64 g = stream_of(seed, static_cast<uint64_t>(k));
    if (u01(g) < frac_3d)
    {
      s.dim = 3;
      for (int d = 0; d < 3; ++d)
        s.nb[d] = side16(g, 64, 128) / 16;
    }
    else
    {
      s.dim = 2;
      s.nb[0] = side16(g, 64, 1024) / 16;
      s.nb[1] = side16(g, 64, 1024) / 16;
      s.nb[2] = 1;
    }
    s.steps = 50 + static_cast<int32_t>(below(g, 451));
    draw_physics(g, s);
  }
  return G5_OK;
}

int
g5_spec_config(const g5_solve_spec * s, int32_t thinc, int32_t precision, g5_config * cfg)
{
  if (not s or not cfg or (s->dim != 2 and s->dim != 3))
    return G5_ERR_INVALID;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->dim = s->dim;
  cfg->bsize = 16;
  for (int d = 0; d < 3; ++d)
    cfg->nb[d] = d < s->dim ? s->nb[d] : 1;
  // unit height: the shortest transverse direction holds 1.0
  int32_t nshort = s->nb[1];
  if (s->dim == 3)
    nshort = std::min(nshort, s->nb[2]);
  cfg->dx = 1.0 / (16.0 * nshort);
  const bool triple = s->kind == 1;
  for (int f = 0; f < 6; ++f)
  {
    const int d = f / 2;
    if (d >= s->dim)
      cfg->bc[f] = G5_BC_NONE;
    else if (d == 0 or triple)
      cfg->bc[f] = G5_BC_ZERO_GRADIENT; // inflow/outflow along the shock (all sides: triple point)
    else
      cfg->bc[f] = G5_BC_WALL; // shock tube walls of the shock-bubble
  }
  cfg->integrator = G5_TI_HANCOCK;
  cfg->thinc_enabled = thinc;
  cfg->precision = precision;
  cfg->cfl = 0.5;
  cfg->slope_type = triple ? 2.0 : 1.0; // MC for the triple point (test_triple_point_2d.ini), else minmod
  cfg->smallr = 1e-10;
  cfg->smallc = 1e-10;
  for (int k = 0; k < 2; ++k)
  {
    cfg->gamma[k] = s->gamma[k];
    cfg->pinf[k] = s->pinf[k];
  }
  cfg->thinc_beta = 1.6;
  cfg->thinc_epsilon = 1e-3;
  return G5_OK;
}

int
g5_spec_regions(const g5_solve_spec * s, g5_region * regions3)
{
  if (not s or not regions3)
    return G5_ERR_INVALID;
  double    rho2, u2, p2, ws;
  const int rc = g5_rankine_hugoniot(s->gamma[1], s->pinf[1], s->rho1, s->p1, s->mach, 0.0, &rho2, &u2, &p2, &ws);
  if (rc != G5_OK)
    return rc;
  std::memset(regions3, 0, 3 * sizeof(g5_region));
  // region 0: post-shock material 1; region 1: pre-shock material 1; region 2: material 0
  regions3[0].alpha0 = 0.0;
  regions3[0].alpha_rho[1] = rho2;
  regions3[0].p = p2;
  regions3[0].u = u2;
  regions3[1].alpha0 = 0.0;
  regions3[1].alpha_rho[1] = s->rho1;
  regions3[1].p = s->p1;
  regions3[2].alpha0 = 1.0;
  regions3[2].alpha_rho[0] = s->rho0;
  regions3[2].p = s->p1;
  return G5_OK;
}

int
g5_spec_state_size(const g5_solve_spec * s, size_t * n)
{
  if (not s or not n or (s->dim != 2 and s->dim != 3))
    return G5_ERR_INVALID;
  const size_t blocks = static_cast<size_t>(s->nb[0]) * s->nb[1] * (s->dim == 3 ? s->nb[2] : 1);
  *n = blocks * (s->dim == 2 ? 8 * 256 : 9 * 4096);
  return G5_OK;
}

} // extern "C"
