// Runs the CPU oracle (oracle/g5_oracle.cpp, compiled into this driver) over configurations that
// reach every code path it has -- 3-D THINC with walls, 2-D RK2_SSP with MC slopes, a periodic
// direction, the MILLER_PUCKETT source term and the shock-bubble inlet, 2-D RK1 THINC with the jet
// inlet -- for the sanitizer builds of the test suite (AddressSanitizer + UndefinedBehaviorSanitizer
// with GCC, ThreadSanitizer with LLVM's OpenMP and its archer tool).  Prints one checksum per case.
#include "../../oracle/g5_oracle.cpp"

#include <cstdio>
#include <vector>

namespace
{

// conservative state of one cell: phase volume fraction a0 (material 0), phase densities, p, u
void
cell_state(g5o_params const & P, double a0, double r0, double r1, double p, double u, double v, double w, double * U)
{
  const double g1[2] = { 1.0 / (P.gamma[0] - 1.0), 1.0 / (P.gamma[1] - 1.0) };
  const double gp[2] = { P.gamma[0] * P.pinf[0] / (P.gamma[0] - 1.0), P.gamma[1] * P.pinf[1] / (P.gamma[1] - 1.0) };
  const double a1 = 1.0 - a0, rho = a0 * r0 + a1 * r1;
  const double rhoe = p * (a0 * g1[0] + a1 * g1[1]) + (a0 * gp[0] + a1 * gp[1]);
  const bool   d3 = P.dim == 3;
  const double k = 0.5 * rho * (u * u + v * v + (d3 ? w * w : 0.0));
  int          n = 0;
  U[n++] = rho;
  U[n++] = rhoe + k;
  U[n++] = rho * u;
  U[n++] = rho * v;
  if (d3)
    U[n++] = rho * w;
  U[n++] = a0;
  U[n++] = a0 * r0;
  U[n++] = a1;
  U[n++] = a1 * r1;
}

// shock-bubble state: post-shock material 1 for x < 0.25 (pressure ratio 5), pre-shock material 1,
// a sphere (circle) of material 0 of radius 0.2 at the domain centre with a smoothed edge
std::vector<double>
shock_bubble(g5o_params const & P, double p1, double r1, double r0)
{
  const int           nvar = P.dim == 2 ? 8 : 9, b = P.bsize, ncell = P.dim == 2 ? b * b : b * b * b;
  const int           nz = P.dim == 3 ? P.nb[2] : 1;
  const int           nblk = P.nb[0] * P.nb[1] * nz;
  std::vector<double> U(static_cast<size_t>(nblk) * nvar * ncell);
  const double        ext[3] = { P.nb[0] * b * P.dx, P.nb[1] * b * P.dx, nz * b * P.dx };
  double              s[9];
  for (int blk = 0; blk < nblk; ++blk)
  {
    const int bx = blk % P.nb[0], by = (blk / P.nb[0]) % P.nb[1], bz = blk / (P.nb[0] * P.nb[1]);
    for (int c = 0; c < ncell; ++c)
    {
      const int    i = c % b, j = (c / b) % b, k = c / (b * b);
      const double x = (bx * b + i + 0.5) * P.dx, y = (by * b + j + 0.5) * P.dx, z = (bz * b + k + 0.5) * P.dx;
      double       d2 = (x - 0.5 * ext[0]) * (x - 0.5 * ext[0]) + (y - 0.5 * ext[1]) * (y - 0.5 * ext[1]);
      if (P.dim == 3)
        d2 += (z - 0.5 * ext[2]) * (z - 0.5 * ext[2]);
      const double a0 = 0.5 * (1.0 + std::tanh((0.2 * ext[1] - std::sqrt(d2)) / (2.0 * P.dx)));
      if (x < 0.25 * ext[0])
        cell_state(P, 0.0, r0, 1.8 * r1, 5.0 * p1, 0.5 * std::sqrt(P.gamma[1] * (p1 + P.pinf[1]) / r1), 0.1, 0.05, s);
      else
        cell_state(P, a0, r0, r1, p1, 0.0, 0.0, 0.0, s);
      for (int v = 0; v < nvar; ++v)
        U[(static_cast<size_t>(blk) * nvar + v) * ncell + c] = s[v];
    }
  }
  return U;
}

double
run_case(g5o_params const & P, std::vector<double> const & U0, int steps)
{
  void * h = g5o_create(&P);
  g5o_set_state(h, U0.data());
  double t = 0.0;
  const int n = g5o_run(h, steps, &t, 1e300);
  std::vector<double> U(U0.size());
  g5o_get_state(h, U.data());
  g5o_destroy(h);
  double sum = 0.0;
  for (double x : U)
    sum += x;
  return n == steps and std::isfinite(sum) ? sum : NAN;
}

g5o_params
base(int dim, int nbx, int nby, int nbz)
{
  g5o_params P{};
  P.dim = dim;
  P.bsize = 16;
  P.nb[0] = nbx;
  P.nb[1] = nby;
  P.nb[2] = dim == 3 ? nbz : 1;
  for (int f = 0; f < 6; ++f)
    P.bc[f] = f < 2 * dim ? 2 : 0;
  P.dx = 1.0 / (16.0 * nby);
  P.cfl = 0.5;
  P.slope_type = 1.0;
  P.smallr = 1e-10;
  P.smallc = 1e-10;
  P.gamma[0] = 1.4;
  P.gamma[1] = 1.4;
  P.thinc_beta = 1.6;
  P.thinc_epsilon = 1e-3;
  return P;
}

} // namespace

int
main()
{
  int bad = 0;
  {
    // 3-D air-water THINC, HANCOCK, walls on every face
    g5o_params P = base(3, 3, 2, 2);
    P.gamma[1] = 4.4;
    P.pinf[1] = 6e8;
    P.thinc_enabled = 1;
    for (int f = 0; f < 6; ++f)
      P.bc[f] = 3;
    const double s = run_case(P, shock_bubble(P, 1e5, 1000.0, 1.0), 4);
    std::printf("3d_thinc_walls %.17g\n", s);
    bad += std::isfinite(s) ? 0 : 1;
  }
  {
    // 2-D gas-gas RK2_SSP, MC slopes, periodic x, MILLER_PUCKETT, shock-bubble inlet
    g5o_params P = base(2, 4, 2, 1);
    P.gamma[0] = 1.648;
    P.integrator = 2;
    P.slope_type = 2.0;
    P.periodic[1] = 1;
    P.bc[2] = P.bc[3] = 1;
    P.source_term = 1;
    P.inlet = 1;
    cell_state(P, 0.0, 0.2, 1.8, 5.0, 0.6, 0.0, 0.0, P.inlet_u);
    const double s = run_case(P, shock_bubble(P, 1.0, 1.0, 0.2), 6);
    std::printf("2d_rk2_mc_periodic_mp_inlet %.17g\n", s);
    bad += std::isfinite(s) ? 0 : 1;
  }
  {
    // 2-D RK1 THINC with the jet inlet, walls in y
    g5o_params P = base(2, 3, 3, 1);
    P.integrator = 1;
    P.thinc_enabled = 1;
    P.bc[2] = P.bc[3] = 3;
    P.inlet = 2;
    P.jet[0] = 0.5;
    P.jet[2] = 0.3;
    cell_state(P, 1.0, 3.0, 1.0, 1.0, 1.2, 0.0, 0.0, P.inlet_u);
    const double s = run_case(P, shock_bubble(P, 1.0, 1.0, 3.0), 6);
    std::printf("2d_rk1_thinc_jet %.17g\n", s);
    bad += std::isfinite(s) ? 0 : 1;
  }
  return bad;
}
