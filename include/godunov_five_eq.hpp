// godunov_five_eq.hpp -- C++ host layer above the C ABI (header only).
//
// Mirrors the reference's plugin surface so that a C++ caller written against
// GodunovImplemBase / SolverGodunovFiveEq finds the same names and meanings:
//   - ConfigMap: .ini reader with the getReal/getInteger/getBool/getString/getRealVector accessors
//     the reference reads its parameters with (same sections and keys);
//   - HydroSettings (models/HydroSettings.{h,cpp}): cfl, slope_type, smallr, smallc, smallp, ...;
//   - config_from_ini: the .ini keys of a fixed single-level brick mesh -> g5_config;
//   - GodunovImplemHip<dim>: resize_auxiliary_data, total_mem_size_in_bytes, do_time_step(U, U2, dt)
//     and do_time_step(U, U2, t, dt) of scheme/GodunovImplemBase.h:93-107;
//   - SolverGodunovFiveEq<dim>: compute_dt_local (SolverGodunovFiveEq.cpp:368-390),
//     godunov_unsplit (:481-526), next_iteration_impl (:395-476) and the run loop with the
//     Mcell-updates/s accounting of :474 and :585-590; save_solution_hdf5 (:773-917) and
//     init_restart (init/init.cpp:132-207) over libg5h5.so (godunov_five_eq_io.h).
// Errors are reported with std::runtime_error carrying g5_last_error().
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "godunov_five_eq.h"
#include "godunov_five_eq_io.h" // save_solution_hdf5 / init_restart need libg5h5.so at link time

namespace g5cpp
{

// ---------------------------------------------------------------------------------------------
// .ini reader: "[section]" headers, "key=value" lines, '#' or ';' comments
// ---------------------------------------------------------------------------------------------
class ConfigMap
{
public:
  ConfigMap() = default;

  static ConfigMap
  from_string(std::string const & text)
  {
    ConfigMap          m;
    std::istringstream in(text);
    std::string        line, section;
    while (std::getline(in, line))
    {
      const auto hash = line.find_first_of("#;");
      if (hash != std::string::npos)
        line = line.substr(0, hash);
      line = trim(line);
      if (line.empty())
        continue;
      if (line.front() == '[' and line.back() == ']')
      {
        section = trim(line.substr(1, line.size() - 2));
        continue;
      }
      const auto eq = line.find('=');
      if (eq == std::string::npos)
        continue;
      m.m_values[section + "/" + trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
    }
    return m;
  }

  static ConfigMap
  from_file(std::string const & path)
  {
    std::ifstream f(path);
    if (not f)
      throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return from_string(ss.str());
  }

  void
  setString(std::string const & section, std::string const & key, std::string const & value)
  {
    m_values[section + "/" + key] = value;
  }

  bool
  has(std::string const & section, std::string const & key) const
  {
    return m_values.count(section + "/" + key) != 0;
  }

  std::string
  getString(std::string const & section, std::string const & key, std::string const & def) const
  {
    const auto it = m_values.find(section + "/" + key);
    return it == m_values.end() ? def : it->second;
  }

  double
  getReal(std::string const & section, std::string const & key, double def) const
  {
    const auto it = m_values.find(section + "/" + key);
    return it == m_values.end() ? def : std::stod(it->second);
  }

  int
  getInteger(std::string const & section, std::string const & key, int def) const
  {
    const auto it = m_values.find(section + "/" + key);
    return it == m_values.end() ? def : std::stoi(it->second);
  }

  bool
  getBool(std::string const & section, std::string const & key, bool def) const
  {
    const auto it = m_values.find(section + "/" + key);
    if (it == m_values.end())
      return def;
    std::string v = it->second;
    std::transform(v.begin(), v.end(), v.begin(), ::tolower);
    return v == "1" or v == "yes" or v == "true" or v == "on";
  }

  // comma-separated list, blanks around the items removed ([output] write_variables=a,b,c)
  std::vector<std::string>
  getStringVector(std::string const & section, std::string const & key, std::vector<std::string> const & def) const
  {
    const auto it = m_values.find(section + "/" + key);
    if (it == m_values.end())
      return def;
    std::vector<std::string> out;
    std::stringstream        ss(it->second);
    std::string              item;
    while (std::getline(ss, item, ','))
    {
      const auto a = item.find_first_not_of(" \t");
      const auto b = item.find_last_not_of(" \t");
      if (a != std::string::npos)
        out.push_back(item.substr(a, b - a + 1));
    }
    return out;
  }

  std::vector<double>
  getRealVector(std::string const & section, std::string const & key, std::vector<double> const & def) const
  {
    const auto it = m_values.find(section + "/" + key);
    if (it == m_values.end())
      return def;
    std::vector<double> out;
    std::stringstream   ss(it->second);
    std::string         item;
    while (std::getline(ss, item, ','))
      if (not trim(item).empty())
        out.push_back(std::stod(trim(item)));
    return out;
  }

private:
  static std::string
  trim(std::string const & s)
  {
    const auto b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos)
      return "";
    const auto e = s.find_last_not_of(" \t\r\n");
    return s.substr(b, e - b + 1);
  }
  std::map<std::string, std::string> m_values;
};

// models/HydroSettings.cpp:18-29 (same keys and defaults)
struct HydroSettings
{
  double      cfl, slope_type, smallr, smallc, smallp, smallpp;
  int         niter_riemann;
  std::string riemann;
  bool        abort_when_negative_eint, abort_when_negative_pressure;
  explicit HydroSettings(ConfigMap const & c)
    : cfl(c.getReal("hydro", "cfl", 0.5))
    , slope_type(c.getReal("hydro", "slope_type", 1.0))
    , smallr(c.getReal("hydro", "smallr", 1e-10))
    , smallc(c.getReal("hydro", "smallc", 1e-10))
    , smallp(smallc * smallc)
    , smallpp(smallr * smallp)
    , niter_riemann(c.getInteger("hydro", "niter_riemann", 10))
    , riemann(c.getString("hydro", "riemann", "UNKNOWN"))
    , abort_when_negative_eint(c.getBool("hydro", "abort_when_negative_eint", false))
    , abort_when_negative_pressure(c.getBool("hydro", "abort_when_negative_pressure", true))
  {}
};

inline int
bc_from_name(std::string const & s)
{
  if (s == "ZERO_GRADIENT")
    return G5_BC_ZERO_GRADIENT;
  if (s == "WALL")
    return G5_BC_WALL;
  if (s == "PERIODIC")
    return G5_BC_PERIODIC;
  return G5_BC_NONE;
}

// The .ini keys of a fixed single-level brick mesh (level_min == level_max) as a g5_config:
// [run] dimension, [amr] bx (=by=bz) level_min level_max time_integrator,
// [p4est_connectivity] nbrick_{x,y,z} periodic_{x,y,z}, [mesh] xmin ymin zmin scaling_factor
// boundary_type_*, [hydro] cfl slope_type smallr smallc riemann, [THINC] enabled beta epsilon,
// [material{0,1}] mg_eos_name with [material{k}_mie_gruneisen_{ig,sg}] gamma pinf.
inline g5_region region_from_ini(ConfigMap const & c, int i);

inline g5_config
config_from_ini(ConfigMap const & c)
{
  g5_config cfg{};
  cfg.dim = c.getInteger("run", "dimension", 2);
  cfg.bsize = c.getInteger("amr", "bx", 16);
  // cubic blocks only: by and bz, when given, must equal bx
  if (c.getInteger("amr", "by", cfg.bsize) != cfg.bsize or (cfg.dim == 3 and c.getInteger("amr", "bz", cfg.bsize) != cfg.bsize))
    throw std::runtime_error("[amr] by and bz must equal bx (cubic blocks)");
  // the gravity source (Hancock predictor term, ComputeFluxesAndStoreFunctor.cpp:197-202, and the
  // gravity CFL, ComputeDtFiveEqFunctor.cpp:138-201) is not on this path
  if (c.getBool("gravity", "enabled", false))
    throw std::runtime_error("[gravity] enabled=true is not available on this path");
  const int lmin = c.getInteger("amr", "level_min", 0);
  const int lmax = c.getInteger("amr", "level_max", lmin);
  if (lmin != lmax)
    throw std::runtime_error("only a fixed single-level mesh is supported (level_min == level_max)");
  const char * ax[3] = { "x", "y", "z" };
  for (int d = 0; d < 3; ++d)
  {
    const int nbrick = c.getInteger("p4est_connectivity", std::string("nbrick_") + ax[d], 1);
    cfg.nb[d] = d < cfg.dim ? nbrick * (1 << lmin) : 1;
    cfg.periodic[d] = d < cfg.dim ? c.getInteger("p4est_connectivity", std::string("periodic_") + ax[d], 0) : 0;
    cfg.xmin[d] = c.getReal("mesh", std::string(ax[d]) + "min", 0.0);
    for (int s = 0; s < 2; ++s)
    {
      const std::string key = std::string("boundary_type_") + ax[d] + (s == 0 ? "min" : "max");
      cfg.bc[2 * d + s] = cfg.periodic[d] ? G5_BC_PERIODIC : bc_from_name(c.getString("mesh", key, "ZERO_GRADIENT"));
    }
  }
  const double scaling = c.getReal("mesh", "scaling_factor", 1.0);
  cfg.dx = scaling / (static_cast<double>(cfg.bsize) * static_cast<double>(1 << lmin));
  const HydroSettings hs(c);
  if (hs.riemann != "HLLC")
    throw std::runtime_error("only riemann=HLLC is supported (models/RiemannSolvers.h:489-499)");
  cfg.cfl = hs.cfl;
  cfg.slope_type = hs.slope_type;
  cfg.smallr = hs.smallr;
  cfg.smallc = hs.smallc;
  cfg.abort_negative_pressure = hs.abort_when_negative_pressure ? 1 : 0;
  const std::string ti = c.getString("amr", "time_integrator", "HANCOCK");
  if (ti != "HANCOCK" and ti != "RK1" and ti != "RK2_SSP")
    throw std::runtime_error("unknown time_integrator " + ti + " (HANCOCK, RK1, RK2_SSP)");
  cfg.integrator = ti == "RK2_SSP" ? G5_TI_RK2_SSP : (ti == "RK1" ? G5_TI_RK1 : G5_TI_HANCOCK);
  cfg.thinc_enabled = c.getBool("THINC", "enabled", false) ? 1 : 0;
  cfg.thinc_beta = c.getReal("THINC", "beta", 1.5);
  cfg.thinc_epsilon = c.getReal("THINC", "epsilon", 1e-3);
  for (int k = 0; k < 2; ++k)
  {
    const std::string mat = "material" + std::to_string(k);
    const std::string eos = c.getString(mat, "mg_eos_name", "MG_IDEAL_GAS");
    // the Mie-Grueneisen closures beyond these two (MG_JWL, MG_SHOCKWAVE2 of the reference's
    // app files) are defined by the reference's EOS class, which is not in its tree
    if (eos != "MG_IDEAL_GAS" and eos != "MG_STIFFENED_GAS")
      throw std::runtime_error("[" + mat + "] mg_eos_name=" + eos +
                               " is not available on this path (MG_IDEAL_GAS, MG_STIFFENED_GAS)");
    const std::string sec = mat + (eos == "MG_STIFFENED_GAS" ? "_mie_gruneisen_sg" : "_mie_gruneisen_ig");
    cfg.gamma[k] = c.getReal(sec, "gamma", 1.4);
    cfg.pinf[k] = eos == "MG_STIFFENED_GAS" ? c.getReal(sec, "pinf", 0.0) : 0.0;
  }
  // [hydro] precision (this layer's key): FP64 (bit-exact), FP64_FAST (reassociated) or MIXED
  const std::string prec = c.getString("hydro", "precision", "FP64");
  if (prec != "FP64" and prec != "FP64_FAST" and prec != "MIXED")
    throw std::runtime_error("unknown [hydro] precision " + prec + " (FP64, FP64_FAST, MIXED)");
  cfg.precision = prec == "MIXED" ? G5_PREC_MIXED : (prec == "FP64_FAST" ? G5_PREC_FP64_FAST : G5_PREC_FP64);
  // [volume_fraction_advection] source_term_type (app/test_shock_bubble_air_water_2d.ini:106-108)
  const std::string st = c.getString("volume_fraction_advection", "source_term_type", "ALLAIRE_KOKH");
  if (st == "MILLER_PUCKETT")
    cfg.source_term = G5_SOURCE_MILLER_PUCKETT;
  else if (st == "ALLAIRE_KOKH")
    cfg.source_term = G5_SOURCE_ALLAIRE_KOKH;
  else
    throw std::runtime_error("unknown [volume_fraction_advection] source_term_type " + st);
  // inflow overrides of fill_outside_quadrants (SolverGodunovFiveEq.cpp:997-1018): the jet takes
  // region 1 ([border] enable_jet, [jet] y, z, width: border_conditions/BCJet.h:62-69,
  // BCJet.cpp:55), the shock-bubble inlet region 0 ([shock_bubble] use_inlet_bc, a
  // ShockBubbleParams key of the absent core: the key name is this layer's choice)
  cfg.jet[0] = c.getReal("jet", "y", 0.5);
  cfg.jet[1] = c.getReal("jet", "z", 0.5);
  cfg.jet[2] = c.getReal("jet", "width", 0.05);
  int inlet_region = -1;
  if (c.getBool("border", "enable_jet", false))
  {
    cfg.inlet = G5_INLET_JET;
    inlet_region = 1;
  }
  else if (c.getString("hydro", "problem", "") == "shock_bubble" and c.getBool("shock_bubble", "use_inlet_bc", false))
  {
    cfg.inlet = G5_INLET_SHOCK_BUBBLE;
    inlet_region = 0;
  }
  if (inlet_region >= 0)
  {
    const g5_region r = region_from_ini(c, inlet_region);
    if (g5_region_conservative(&cfg, &r, cfg.inlet_u, cfg.dim == 2 ? 8 : 9) != G5_OK)
      throw std::runtime_error("inlet region state: " + std::string(g5_last_error(nullptr)));
  }
  return cfg;
}

// [regionN] keys of common.h:80-109
inline g5_region
region_from_ini(ConfigMap const & c, int i)
{
  const std::string s = "region" + std::to_string(i);
  g5_region         r{};
  const auto        ar = c.getRealVector(s, "alpha_rho", { 0.0, 0.0 });
  if (ar.size() != 2)
    throw std::runtime_error("alpha_rho needs two values");
  r.alpha0 = c.getReal(s, "alpha0", 1.0);
  r.alpha_rho[0] = ar[0];
  r.alpha_rho[1] = ar[1];
  r.p = c.getReal(s, "p", 1.0);
  r.u = c.getReal(s, "u", 0.0);
  r.v = c.getReal(s, "v", 0.0);
  r.w = c.getReal(s, "w", 0.0);
  return r;
}

inline void
check(int rc, g5_ctx const * ctx)
{
  if (rc != G5_OK)
    throw std::runtime_error(std::string("g5: ") + g5_last_error(ctx));
}

// The initial state of [hydro] problem on the owned blocks, U[block][var][cell] (the data functors
// dispatched by init/init.cpp:37-118, host side):
//   shock_bubble  InitShockBubble.cpp:50-225, [shock_bubble] x_front, bubble_x, bubble_y,
//                 bubble_z, bubble_radius, regions 0 (post-shock), 1 (pre-shock), 2 (bubble);
//   triple_point  InitTriplePoint.cpp:30-90, [triple_point] xd, yd (defaults 1 and 1.5, the
//                 7x3 domain of app/test_triple_point_2d.ini), regions 0, 1, 2;
//   jet           InitJet.cpp:45-72, region 0 everywhere (the jet enters through the inlet).
// The parameter structs of the other problems are in the kalypsso core, which is not part of the
// reference tree: they raise.
inline std::vector<double>
initial_state(ConfigMap const & c, g5_config const & cfg)
{
  const std::string problem = c.getString("hydro", "problem", "");
  const int         nvar = cfg.dim == 2 ? 8 : 9;
  const size_t      ncell = static_cast<size_t>(cfg.dim == 2 ? cfg.bsize * cfg.bsize : cfg.bsize * cfg.bsize * cfg.bsize);
  size_t            n = static_cast<size_t>(nvar) * ncell;
  for (int d = 0; d < cfg.dim; ++d)
    n *= static_cast<size_t>(cfg.nb[d]);
  std::vector<double> U(n);
  g5_region           regs[3];
  if (problem == "shock_bubble" or problem == "triple_point")
    for (int r = 0; r < 3; ++r)
      regs[r] = region_from_ini(c, r);
  if (problem == "shock_bubble")
  {
    const double ctr[3] = { c.getReal("shock_bubble", "bubble_x", 0.5), c.getReal("shock_bubble", "bubble_y", 0.5),
                            c.getReal("shock_bubble", "bubble_z", 0.5) };
    check(g5_init_shock_bubble(&cfg, regs, c.getReal("shock_bubble", "x_front", 0.2), ctr,
                               c.getReal("shock_bubble", "bubble_radius", 0.1), U.data(), U.size()),
          nullptr);
  }
  else if (problem == "triple_point")
    check(g5_init_triple_point(&cfg, regs, c.getReal("triple_point", "xd", 1.0), c.getReal("triple_point", "yd", 1.5),
                               U.data(), U.size()),
          nullptr);
  else if (problem == "jet")
  {
    const g5_region r0 = region_from_ini(c, 0);
    double          u[9];
    check(g5_region_conservative(&cfg, &r0, u, nvar), nullptr);
    for (size_t k = 0; k < n; ++k)
      U[k] = u[(k / ncell) % nvar];
  }
  else
    throw std::runtime_error("problem '" + problem + "' is not available on this path (shock_bubble, triple_point, jet)");
  return U;
}

// ---------------------------------------------------------------------------------------------
// State handle: the device-resident conservative state of a context (the solver's U / U2).
// ---------------------------------------------------------------------------------------------
struct DataArrayBlock
{
  g5_ctx * ctx = nullptr;  // a context's own state ...
  double * data = nullptr; // ... or the caller's device array U[block][var][cell], blocks in the
                           // context's block order (GodunovImplemHip::set_block_order)
};

// scheme/GodunovImplemBase.h:59-198 surface, implemented by the HIP kernels
template <int dim>
class GodunovImplemHip
{
public:
  explicit GodunovImplemHip(g5_config const & cfg)
  {
    if (cfg.dim != dim)
      throw std::runtime_error("dimension mismatch");
    g5_ctx * c = nullptr;
    check(g5_create(&cfg, &c), nullptr);
    m_ctx.reset(c);
  }

  // a two-level static mesh: the coarse blocks at refined[dim k .. dim k + dim) refined into 2^dim
  // blocks of half the cell length; prolongation 0 SIMPLE_COPY, 1 EXTRAPOLATE_LINEAR_MINMOD,
  // 2 / 3 CONSERVATIVE_INTERPOLATION_ORDER_2 / _4
  // (g5_create_two_level; the state's block order is block_layout())
  GodunovImplemHip(g5_config const & cfg, std::vector<int32_t> const & refined, int prolongation)
  {
    if (cfg.dim != dim)
      throw std::runtime_error("dimension mismatch");
    g5_ctx * c = nullptr;
    check(g5_create_two_level(&cfg, refined.data(), static_cast<int32_t>(refined.size() / dim), prolongation, &c),
          nullptr);
    m_ctx.reset(c);
  }

  // static factory, scheme/GodunovImplemV0.h:103-114
  static GodunovImplemHip *
  create(ConfigMap const & config_map)
  {
    return new GodunovImplemHip(config_from_ini(config_map));
  }

  // (x, y, z, level) of each block of the state, in order
  std::vector<int32_t>
  block_layout() const
  {
    int32_t n = 0, nv = 0, nc = 0;
    check(g5_shape(m_ctx.get(), &n, &nv, &nc), m_ctx.get());
    std::vector<int32_t> c(4 * static_cast<size_t>(n));
    check(g5_block_layout(m_ctx.get(), c.data(), c.size()), m_ctx.get());
    return c;
  }

  // ComputeRefineFlags::run (scheme/ComputeRefineFlags.cpp:18-62) on the current state: flags
  // (one per block, G5_AMR_*) max-accumulated, as into the AMR context's reset flag array
  void
  compute_refine_flags(g5_refine_params const & params, std::vector<int32_t> & flags)
  {
    check(g5_refine_flags(m_ctx.get(), &params, flags.data(), nullptr, flags.size()), m_ctx.get());
  }

  // scratch is sized at creation for the fixed mesh: nothing to resize
  void
  resize_auxiliary_data()
  {}

  uint64_t
  total_mem_size_in_bytes()
  {
    uint64_t b = 0;
    check(g5_mem_bytes(m_ctx.get(), &b), m_ctx.get());
    return b;
  }

  // With device arrays (U.data, U2.data): U2 <- one step of U on the caller's arrays, which the
  // caller owns (its U <- U2 copy stays its own, SolverGodunovFiveEq.cpp:494).  With context
  // handles (or none): U -> U2 is kept inside the context (U2 becomes the current state, as after
  // the reference's deep_copy U <- U2 in godunov_unsplit); a handle must then be this context's.
  void
  do_time_step(DataArrayBlock const & U, DataArrayBlock const & U2, double dt)
  {
    if (U.data or U2.data)
    {
      if (not U.data or not U2.data)
        throw std::invalid_argument("do_time_step: U and U2 must both be device arrays");
      check(g5_do_time_step_buffers(m_ctx.get(), U.data, U2.data, dt), m_ctx.get());
      return;
    }
    if ((U.ctx and U.ctx != m_ctx.get()) or (U2.ctx and U2.ctx != m_ctx.get()))
      throw std::invalid_argument("do_time_step: state handle of another context (the state stays in its context)");
    check(g5_do_time_step(m_ctx.get(), dt), m_ctx.get());
  }

  // the caller's block order (brick coordinates of its blocks, 3 per block): the order of its
  // DataArrayBlock device arrays and of the context's own state from then on
  void
  set_block_order(std::vector<int32_t> const & coords)
  {
    check(g5_set_block_order(m_ctx.get(), coords.data(), static_cast<int32_t>(coords.size() / 3)), m_ctx.get());
  }

  // compute_dt_local on the caller's device array (SolverGodunovFiveEq.cpp:368-390)
  double
  compute_dt_local(DataArrayBlock const & U)
  {
    double dt = 0.0;
    if (U.data)
      check(g5_compute_dt_buffer(m_ctx.get(), U.data, &dt), m_ctx.get());
    else
      check(g5_compute_dt(m_ctx.get(), &dt), m_ctx.get());
    return dt;
  }

  void
  do_time_step(DataArrayBlock const & U, DataArrayBlock const & U2, double /*t*/, double dt)
  {
    do_time_step(U, U2, dt);
  }

  g5_ctx *
  ctx() const
  {
    return m_ctx.get();
  }

private:
  struct Del
  {
    void
    operator()(g5_ctx * c) const
    {
      g5_destroy(c);
    }
  };
  std::unique_ptr<g5_ctx, Del> m_ctx;
};

// SolverGodunovFiveEq: time loop around the implementation (fixed mesh, one GPU)
template <int dim>
class SolverGodunovFiveEq
{
public:
  explicit SolverGodunovFiveEq(ConfigMap const & config_map)
    : m_config_map(config_map)
    , m_cfg(config_from_ini(config_map))
    , m_impl(m_cfg)
    , m_tEnd(config_map.getReal("run", "tEnd", 1.0))
    , m_nStepmax(config_map.getInteger("run", "nStepmax", 1000))
    , m_outputDir(config_map.getString("output", "outputDir", "./"))
    , m_outputPrefix(config_map.getString("output", "outputPrefix", "output"))
    , m_write_variables(config_map.getStringVector("output", "write_variables", {}))
    , m_restart_prefix(config_map.getString("restart", "filePrefix", ""))
    , m_hdf5_enabled(config_map.getBool("output", "hdf5_enabled", false))
    , m_nOutput(config_map.getInteger("run", "nOutput", 0))
  {}

  // HDF5 + XDMF output of the state (save_solution_hdf5, SolverGodunovFiveEq.cpp:773-917):
  // <outputDir>/<outputPrefix>_<iteration, 7 digits>.h5 and .xmf.  A checkpoint holds every
  // conservative variable; otherwise the conservative variables named in [output]
  // write_variables (derived quantities are not written; with none named, every variable).
  // Returns the .h5 path.
  std::string
  save_solution_hdf5(bool checkpoint = true)
  {
    uint32_t mask = 0;
    if (!checkpoint)
      for (int v = 0; v < (dim == 2 ? 8 : 9); ++v)
        if (std::find(m_write_variables.begin(), m_write_variables.end(), g5_h5_var_name(dim, v)) !=
            m_write_variables.end())
          mask |= 1u << v;
    char it[16];
    std::snprintf(it, sizeof(it), "%07d", m_iteration);
    const std::string base = m_outputPrefix + "_" + it;
    const auto        U = get_state();
    if (g5_h5_write(m_outputDir.c_str(), base.c_str(), &m_cfg, U.data(), U.size(), m_t, m_iteration, mask) != G5_OK)
      throw std::runtime_error(std::string("save_solution_hdf5: ") + g5_h5_last_error());
    return m_outputDir + "/" + base + ".h5";
  }

  // init (init/init.cpp:37-118): the restart when [restart] enabled, else the initial state of
  // [hydro] problem (initial_state)
  void
  init()
  {
    if (m_config_map.getBool("restart", "enabled", false))
      init_restart();
    else
      set_state(initial_state(m_config_map, m_cfg));
  }

  // restart (init_restart, init/init.cpp:132-207): state, time and iteration from
  // <outputDir>/<[restart] filePrefix>.h5, a checkpoint of the same mesh
  void
  init_restart()
  {
    if (m_restart_prefix.empty())
      throw std::runtime_error("init_restart: [restart] filePrefix is not set");
    const std::string path = m_outputDir + "/" + m_restart_prefix + ".h5";
    std::vector<double> U(num_cells() * (dim == 2 ? 8 : 9));
    double              t = 0.0;
    int64_t             it = 0;
    if (g5_h5_read(path.c_str(), &m_cfg, U.data(), U.size(), &t, &it) != G5_OK)
      throw std::runtime_error(std::string("init_restart: ") + g5_h5_last_error());
    set_state(U);
    m_t = t;
    m_iteration = static_cast<int>(it);
  }

  void
  set_state(std::vector<double> const & U)
  {
    check(g5_upload_state(m_impl.ctx(), U.data(), U.size()), m_impl.ctx());
  }

  std::vector<double>
  get_state()
  {
    int32_t nb, nv, nc;
    check(g5_shape(m_impl.ctx(), &nb, &nv, &nc), m_impl.ctx());
    std::vector<double> U(static_cast<size_t>(nb) * nv * nc);
    check(g5_download_state(m_impl.ctx(), U.data(), U.size()), m_impl.ctx());
    return U;
  }

  double
  compute_dt_local()
  {
    double dt = 0.0;
    check(g5_compute_dt(m_impl.ctx(), &dt), m_impl.ctx());
    return dt;
  }

  void
  godunov_unsplit(double dt)
  {
    m_impl.do_time_step(DataArrayBlock{ m_impl.ctx() }, DataArrayBlock{ m_impl.ctx() }, m_t, dt);
  }

  // one iteration: dt (clamped to tEnd), step, cell-update accounting
  void
  next_iteration_impl()
  {
    double dt = compute_dt_local();
    if (m_t + dt > m_tEnd)
      dt = m_tEnd - m_t;
    godunov_unsplit(dt);
    m_t += dt;
    ++m_iteration;
    m_cell_updates += num_cells();
  }

  // The whole run with the device-resident time loop, to tEnd or nStepmax; returns the
  // Mcell-updates/s of the stepping.  With [output] hdf5_enabled and [run] nOutput > 0 the state
  // is written (save_solution_hdf5, the [output] write_variables) at every iteration that is a
  // multiple of nOutput and after the last step; nOutput < 0 writes after every step (the
  // reference's .ini convention).  The scheduling of the reference's absent SolverBase
  // (should_save_solution) is restated this way.
  double
  run()
  {
    const bool out = m_hdf5_enabled and m_nOutput != 0;
    const int  every = m_nOutput < 0 ? 1 : m_nOutput;
    double     s = 0.0;
    uint64_t   stepped = 0;
    int        last_saved = -1;
    while (m_iteration < m_nStepmax and m_t < m_tEnd)
    {
      if (out and m_iteration % every == 0)
      {
        save_solution_hdf5(false);
        last_saved = m_iteration;
      }
      const int n = out ? std::min(every - m_iteration % every, m_nStepmax - m_iteration) : m_nStepmax - m_iteration;
      int32_t   done = 0;
      const auto t0 = std::chrono::steady_clock::now();
      check(g5_run(m_impl.ctx(), n, &m_t, m_tEnd, &done), m_impl.ctx());
      s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      m_iteration += done;
      stepped += static_cast<uint64_t>(done);
      if (done == 0)
        break;
    }
    if (out and last_saved != m_iteration)
      save_solution_hdf5(false);
    m_cell_updates += stepped * num_cells();
    return s > 0.0 ? static_cast<double>(stepped) * num_cells() / s / 1e6 : 0.0;
  }

  uint64_t
  num_cells() const
  {
    uint64_t n = 1;
    for (int d = 0; d < dim; ++d)
      n *= static_cast<uint64_t>(m_cfg.nb[d]) * m_cfg.bsize;
    return n;
  }

  double
  time() const
  {
    return m_t;
  }
  int
  iteration() const
  {
    return m_iteration;
  }
  uint64_t
  total_num_cell_updates() const
  {
    return m_cell_updates;
  }
  g5_config const &
  config() const
  {
    return m_cfg;
  }

private:
  ConfigMap              m_config_map;
  g5_config              m_cfg;
  GodunovImplemHip<dim>  m_impl;
  double                 m_tEnd;
  int                    m_nStepmax;
  std::string              m_outputDir, m_outputPrefix;
  std::vector<std::string> m_write_variables;
  std::string              m_restart_prefix;
  bool                     m_hdf5_enabled;
  int                      m_nOutput;
  double                 m_t = 0.0;
  int                    m_iteration = 0;
  uint64_t               m_cell_updates = 0;
};

// Many independent solves stepped together (the batching and serving forms): RAII over g5_pool.
class PoolHip
{
public:
  // rk2: also hold the RK2_SSP stage buffer (g5_pool_create_rk2), for RK2_SSP solves
  PoolHip(int dim, bool thinc, int precision, uint64_t hbm_budget_bytes, bool rk2 = false)
  {
    g5_pool * p = nullptr;
    const int rc = rk2 ? g5_pool_create_rk2(dim, thinc ? 1 : 0, precision, hbm_budget_bytes, &p)
                       : g5_pool_create(dim, thinc ? 1 : 0, precision, hbm_budget_bytes, &p);
    if (rc != G5_OK)
      throw std::runtime_error(std::string("g5_pool_create: ") + g5_pool_last_error(nullptr));
    m_pool.reset(p);
  }

  // admit a solve with its configuration; -1 when the HBM budget has no room now
  int32_t
  admit(g5_config const & cfg, double t_end, int32_t max_steps)
  {
    int32_t   id = -1;
    const int rc = g5_pool_admit(m_pool.get(), &cfg, t_end, max_steps, &id);
    if (rc == G5_ERR_NOMEM)
      return -1;
    pcheck(rc);
    return id;
  }

  // admit a seeded workload solve, its initial condition generated on the device
  int32_t
  admit(g5_solve_spec const & spec)
  {
    int32_t   id = -1;
    const int rc = g5_pool_admit_spec(m_pool.get(), &spec, &id);
    if (rc == G5_ERR_NOMEM)
      return -1;
    pcheck(rc);
    return id;
  }

  void
  upload(int32_t id, std::vector<double> const & U)
  {
    pcheck(g5_pool_upload(m_pool.get(), id, U.data(), U.size()));
  }
  void
  download(int32_t id, std::vector<double> & U)
  {
    pcheck(g5_pool_download(m_pool.get(), id, U.data(), U.size()));
  }
  void
  start(std::vector<int32_t> const & ids)
  {
    pcheck(g5_pool_start(m_pool.get(), ids.data(), static_cast<int32_t>(ids.size())));
  }
  void
  step(int32_t nsteps)
  {
    pcheck(g5_pool_step(m_pool.get(), nsteps));
  }
  std::vector<int32_t>
  finished()
  {
    int32_t n = 0;
    pcheck(g5_pool_finished(m_pool.get(), nullptr, 0, &n));
    std::vector<int32_t> ids(static_cast<size_t>(n));
    pcheck(g5_pool_finished(m_pool.get(), ids.data(), n, &n));
    return ids;
  }
  void
  release(int32_t id)
  {
    pcheck(g5_pool_release(m_pool.get(), id));
  }
  // (t, steps done, active)
  std::tuple<double, int32_t, bool>
  status(int32_t id)
  {
    double  t = 0.0;
    int32_t n = 0, a = 0;
    pcheck(g5_pool_status(m_pool.get(), id, &t, &n, &a));
    return { t, n, a != 0 };
  }
  g5_pool *
  get() const
  {
    return m_pool.get();
  }

private:
  void
  pcheck(int rc)
  {
    if (rc != G5_OK)
      throw std::runtime_error(std::string("g5_pool: ") + g5_pool_last_error(m_pool.get()));
  }
  struct Del
  {
    void
    operator()(g5_pool * p) const
    {
      g5_pool_destroy(p);
    }
  };
  std::unique_ptr<g5_pool, Del> m_pool;
};

} // namespace g5cpp
