// The host C++ of the library on the CPU alone, for the sanitizer builds of the test suite: the
// .ini layer (include/godunov_five_eq.hpp: config_from_ini, initial_state, its refusals), the host
// initial conditions and Rankine-Hugoniot states (g5_init.cpp), the seeded workload specs
// (g5_specs.cpp), the two-level mesh refusals and the HDF5 checkpoint round trip (g5_hdf5.cpp).
// Prints checksums so that a sanitized build can be compared with a plain one.
//   host_cpu <golden dir> <scratch dir>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/godunov_five_eq.hpp"
#include "../../include/godunov_five_eq_io.h"

namespace
{

double
checksum(std::vector<double> const & U)
{
  double s = 0.0;
  for (double x : U)
    s += x;
  return s;
}

int
refused(g5cpp::ConfigMap cm, std::string const & section, std::string const & key, std::string const & value)
{
  cm.setString(section, key, value);
  try
  {
    (void)g5cpp::config_from_ini(cm);
  }
  catch (std::exception const & e)
  {
    std::printf("refused %s.%s=%s: %s\n", section.c_str(), key.c_str(), value.c_str(), e.what());
    return 0;
  }
  return 1;
}

} // namespace

int
main(int argc, char ** argv)
{
  if (argc < 3)
    return 2;
  const std::string gold = argv[1], scratch = argv[2];
  int               bad = 0;
  for (const char * ini : { "shock_bubble_air_helium_2d_512x128.ini", "shock_bubble_air_water_3d_256.ini",
                            "jet_fuel_n2_2d_fixed.ini", "triple_point_2d_fixed.ini" })
  {
    const auto cm = g5cpp::ConfigMap::from_file(gold + "/" + ini);
    const auto cfg = g5cpp::config_from_ini(cm);
    double     sum = 0.0;
    if (cfg.dim == 2) // the 256^3 3-D state is left to the GPU tests
      sum = checksum(g5cpp::initial_state(cm, cfg));
    std::printf("%s dim=%d nb=%d,%d,%d dx=%.17g inlet=%d abort_p=%d init_sum=%.17g\n", ini, cfg.dim, cfg.nb[0],
                cfg.nb[1], cfg.nb[2], cfg.dx, cfg.inlet, cfg.abort_negative_pressure, sum);
  }
  const auto base = g5cpp::ConfigMap::from_file(gold + "/shock_bubble_air_helium_2d_512x128.ini");
  bad += refused(base, "gravity", "enabled", "true");
  bad += refused(base, "amr", "by", "8");
  bad += refused(base, "material1", "mg_eos_name", "MG_JWL");
  bad += refused(base, "amr", "time_integrator", "RK3");
  bad += refused(base, "amr", "level_max", "5");
  // a two-level mesh whose refined block touches the brick's border is refused before any device
  // allocation
  try
  {
    g5_config c2 = g5cpp::config_from_ini(base);
    g5cpp::GodunovImplemHip<2> impl(c2, { 0, 0 }, 1);
    bad += 1;
  }
  catch (std::exception const & e)
  {
    std::printf("refused two-level mesh: %s\n", e.what());
  }
  // workload specs and their configurations / region states
  double spec_sum = 0.0;
  for (int k = 0; k < 256; ++k)
  {
    g5_solve_spec s;
    g5_config     c;
    g5_region     r[3];
    bad += g5_batch_spec(20251203, k, &s) != G5_OK;
    bad += g5_spec_config(&s, k & 1, G5_PREC_MIXED, &c) != G5_OK;
    bad += g5_spec_regions(&s, r) != G5_OK;
    spec_sum += s.mach + s.gamma[0] + s.gamma[1] + r[0].p + r[0].u + c.dx;
  }
  std::vector<g5_solve_spec> serve(400);
  bad += g5_serve_specs(20251203, 400, 160.0, 0.1, serve.data()) != G5_OK;
  for (auto const & s : serve)
  {
    size_t n = 0;
    bad += g5_spec_state_size(&s, &n) != G5_OK;
    spec_sum += s.arrival + static_cast<double>(n);
  }
  std::printf("specs_sum=%.17g\n", spec_sum);
  // host initial conditions (triple point in 3-D) and an HDF5 checkpoint round trip
  g5_config c{};
  c.dim = 3;
  c.bsize = 16;
  c.nb[0] = 3;
  c.nb[1] = 2;
  c.nb[2] = 2;
  c.dx = 1.0 / 32;
  c.cfl = 0.5;
  c.gamma[0] = 1.4;
  c.gamma[1] = 1.5;
  for (int f = 0; f < 6; ++f)
    c.bc[f] = G5_BC_ZERO_GRADIENT;
  g5_region r[3] = {};
  for (int k = 0; k < 3; ++k)
  {
    r[k].alpha0 = k == 2 ? 0.0 : 1.0;
    r[k].alpha_rho[0] = 1.0 + k;
    r[k].alpha_rho[1] = 0.125 * (k + 1);
    r[k].p = 1.0 / (k + 1);
  }
  std::vector<double> U(static_cast<size_t>(12) * 9 * 4096);
  bad += g5_init_triple_point(&c, r, 1.0 / 3.0, 0.5, U.data(), U.size()) != G5_OK;
  bad += g5_h5_write(scratch.c_str(), "ckpt", &c, U.data(), U.size(), 0.25, 7, 0u) != G5_OK;
  std::vector<double> V(U.size());
  double              t = 0.0;
  int64_t             it = 0;
  bad += g5_h5_read((scratch + "/ckpt.h5").c_str(), &c, V.data(), V.size(), &t, &it) != G5_OK;
  const bool same = V == U and t == 0.25 and it == 7;
  std::printf("triple_point_3d_sum=%.17g h5_round_trip=%d\n", checksum(U), same ? 1 : 0);
  bad += same ? 0 : 1;
  double rho2, u2, p2, ws;
  bad += g5_rankine_hugoniot(4.4, 6e8, 1000.0, 1e5, 0.0, 1e9, &rho2, &u2, &p2, &ws) != G5_OK;
  std::printf("rh rho2=%.17g u2=%.17g ws=%.17g\n", rho2, u2, ws);
  return bad;
}
