// Exercises the C++ host layer (include/godunov_five_eq.hpp) above the C ABI.
//   host_layer <file.ini>          parse the .ini and print the derived g5_config
//   host_layer <file.ini> --run N  also build the shock-bubble state and run N steps on the GPU
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/godunov_five_eq.hpp"

template <int dim>
int
run(g5cpp::ConfigMap const & cm, int nsteps)
{
  g5cpp::SolverGodunovFiveEq<dim> solver(cm);
  solver.init();
  for (int n = 0; n < 2; ++n)
    solver.next_iteration_impl();
  const double mcups = solver.run();
  std::printf("steps=%d t=%.9e Mcell-updates/s=%.1f\n", solver.iteration(), solver.time(), mcups);
  (void)nsteps;
  return 0;
}

// checkpoint after n1 iterations, restart a second solver from the file, continue both n2
// iterations: the states, times and iterations must be identical
template <int dim>
int
run_restart(g5cpp::ConfigMap cm, std::string const & dir, int n1, int n2)
{
  cm.setString("output", "outputDir", dir);
  cm.setString("output", "outputPrefix", "ckpt");
  g5cpp::SolverGodunovFiveEq<dim> a(cm);
  a.init();
  for (int n = 0; n < n1; ++n)
    a.next_iteration_impl();
  const std::string h5 = a.save_solution_hdf5(true);
  for (int n = 0; n < n2; ++n)
    a.next_iteration_impl();
  char prefix[32];
  std::snprintf(prefix, sizeof(prefix), "ckpt_%07d", n1);
  cm.setString("restart", "filePrefix", prefix);
  cm.setString("restart", "enabled", "true");
  g5cpp::SolverGodunovFiveEq<dim> b(cm);
  b.init();
  const int it_restart = b.iteration();
  for (int n = 0; n < n2; ++n)
    b.next_iteration_impl();
  const bool same = a.get_state() == b.get_state() and a.time() == b.time() and a.iteration() == b.iteration();
  std::printf("file=%s restart_iteration=%d iteration=%d t=%.17g identical=%d\n", h5.c_str(), it_restart,
              b.iteration(), b.time(), same ? 1 : 0);
  return same ? 0 : 1;
}

// a few solves of the seeded batch in one pool through the C++ PoolHip class
int
run_pool(int nsolves)
{
  g5cpp::PoolHip       pool(2, false, G5_PREC_MIXED, 1ull << 30);
  std::vector<int32_t> ids;
  for (int k = 0; k < nsolves; ++k)
  {
    g5_solve_spec s;
    g5cpp::check(g5_batch_spec(20251203, k, &s), nullptr);
    s.nb[0] = 6 + k % 3;
    s.nb[1] = 4;
    s.steps = 20 + k;
    ids.push_back(pool.admit(s));
  }
  pool.start(ids);
  pool.step(20 + nsolves);
  const auto fin = pool.finished();
  int        ok = 0;
  for (int k = 0; k < nsolves; ++k)
  {
    const auto st = pool.status(ids[k]);
    ok += std::get<1>(st) == 20 + k and !std::get<2>(st);
  }
  std::printf("pool solves=%d finished=%zu complete=%d\n", nsolves, fin.size(), ok);
  return ok == nsolves and static_cast<int>(fin.size()) == nsolves ? 0 : 1;
}

int
main(int argc, char ** argv)
{
  if (argc >= 3 and std::strcmp(argv[1], "--pool") == 0)
    return run_pool(std::atoi(argv[2]));
  if (argc < 2)
    return 2;
  const auto cm = g5cpp::ConfigMap::from_file(argv[1]);
  const auto cfg = g5cpp::config_from_ini(cm);
  const g5cpp::HydroSettings hs(cm);
  std::printf("dim=%d bsize=%d nb=%d,%d,%d periodic=%d,%d,%d bc=%d,%d,%d,%d,%d,%d\n", cfg.dim, cfg.bsize, cfg.nb[0],
              cfg.nb[1], cfg.nb[2], cfg.periodic[0], cfg.periodic[1], cfg.periodic[2], cfg.bc[0], cfg.bc[1], cfg.bc[2],
              cfg.bc[3], cfg.bc[4], cfg.bc[5]);
  std::printf("dx=%.17g cfl=%.17g slope_type=%g smallr=%g smallc=%g smallp=%g\n", cfg.dx, cfg.cfl, cfg.slope_type,
              cfg.smallr, cfg.smallc, hs.smallp);
  std::printf("gamma=%.17g,%.17g pinf=%.17g,%.17g thinc=%d beta=%g eps=%g integrator=%d\n", cfg.gamma[0], cfg.gamma[1],
              cfg.pinf[0], cfg.pinf[1], cfg.thinc_enabled, cfg.thinc_beta, cfg.thinc_epsilon, cfg.integrator);
  std::printf("source_term=%d abort_negative_pressure=%d\n", cfg.source_term, cfg.abort_negative_pressure);
  std::printf("inlet=%d jet=%.17g,%.17g,%.17g inlet_u=", cfg.inlet, cfg.jet[0], cfg.jet[1], cfg.jet[2]);
  for (int v = 0; v < (cfg.dim == 2 ? 8 : 9); ++v)
    std::printf(v ? ",%.17g" : "%.17g", cfg.inlet_u[v]);
  std::printf("\n");
  const auto r0 = g5cpp::region_from_ini(cm, 0);
  std::printf("region0 alpha0=%g alpha_rho=%.17g,%.17g p=%.17g u=%.17g\n", r0.alpha0, r0.alpha_rho[0], r0.alpha_rho[1],
              r0.p, r0.u);
  if (argc >= 3 and std::strcmp(argv[2], "--init") == 0)
  {
    // host only: the initial state of [hydro] problem, written raw to argv[3] when given
    const auto U = g5cpp::initial_state(cm, cfg);
    double     sum = 0.0;
    for (double x : U)
      sum += x;
    std::printf("problem=%s values=%zu sum=%.17g\n", cm.getString("hydro", "problem", "").c_str(), U.size(), sum);
    if (argc >= 4)
    {
      FILE * f = std::fopen(argv[3], "wb");
      if (!f or std::fwrite(U.data(), sizeof(double), U.size(), f) != U.size())
        return 3;
      std::fclose(f);
    }
    return 0;
  }
  if (argc >= 4 and std::strcmp(argv[2], "--outputs") == 0)
  {
    // run() with periodic HDF5 output: 40 steps, nOutput 10, into argv[3]
    g5cpp::ConfigMap m = cm;
    m.setString("run", "nStepmax", "40");
    m.setString("run", "nOutput", "10");
    m.setString("output", "hdf5_enabled", "true");
    m.setString("output", "outputDir", argv[3]);
    m.setString("output", "outputPrefix", "o");
    m.setString("output", "write_variables", "rho_mix,alpha_0,e_tot");
    if (cfg.dim != 2)
      return 2;
    g5cpp::SolverGodunovFiveEq<2> solver(m);
    solver.init();
    solver.run();
    const auto   U = solver.get_state();
    double       sum = 0.0;
    for (double x : U)
      sum += x;
    std::printf("steps=%d t=%.17g sum=%.17g\n", solver.iteration(), solver.time(), sum);
    return 0;
  }
  if (argc >= 4 and std::strcmp(argv[2], "--restart") == 0)
    return cfg.dim == 2 ? run_restart<2>(cm, argv[3], 30, 20) : run_restart<3>(cm, argv[3], 6, 4);
  if (argc >= 4 and std::strcmp(argv[2], "--run") == 0)
    return cfg.dim == 2 ? run<2>(cm, std::atoi(argv[3])) : run<3>(cm, std::atoi(argv[3]));
  return 0;
}
