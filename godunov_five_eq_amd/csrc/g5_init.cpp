// g5_init.cpp -- host-side initial conditions and the Rankine-Hugoniot helper of the C ABI.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/godunov_five_eq.h"

namespace
{

struct Mix
{
  double g1[2], gp[2];
  explicit Mix(g5_config const & c)
  {
    for (int k = 0; k < 2; ++k)
    {
      g1[k] = 1.0 / (c.gamma[k] - 1.0);
      gp[k] = c.gamma[k] * c.pinf[k] / (c.gamma[k] - 1.0);
    }
  }
  // mixture specific internal energy e(rho, p) of the isobaric stiffened-gas closure
  double
  eint(double rho, double p, double a0) const
  {
    const double a1 = 1.0 - a0;
    return (p * (a0 * g1[0] + a1 * g1[1]) + (a0 * gp[0] + a1 * gp[1])) / rho;
  }
};

// common.h:68-144 get_region_init_state: primitive region -> conservative state
void
region_to_cons(g5_config const & c, Mix const & m, g5_region const & r, double * U)
{
  const int    dim = c.dim;
  const double rho = r.alpha_rho[0] + r.alpha_rho[1];
  const double e = m.eint(rho, r.p, r.alpha0);
  const double ekin = dim == 2 ? 0.5 * (r.u * r.u + r.v * r.v) : 0.5 * (r.u * r.u + r.v * r.v + r.w * r.w);
  // FiveEq ids
  const int IA0 = dim == 2 ? 4 : 5;
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

// fraction of the cell [lo, lo+dx]^d inside the sphere (centre c, radius R): exact 0 or 1 away
// from the surface, midpoint quadrature on 16^d (2-D) or 8^3 (3-D) sub-cells near it
double
sphere_fraction(int dim, double const * lo, double dx, double const * c, double R)
{
  double d2 = 0.0;
  for (int k = 0; k < dim; ++k)
  {
    const double x = lo[k] + 0.5 * dx - c[k];
    d2 += x * x;
  }
  const double dist = std::sqrt(d2);
  const double half_diag = 0.5 * dx * std::sqrt(static_cast<double>(dim));
  if (dist + half_diag <= R)
    return 1.0;
  if (dist - half_diag >= R)
    return 0.0;
  const int n = dim == 2 ? 16 : 8;
  int       inside = 0, total = 0;
  for (int kk = 0; kk < (dim == 3 ? n : 1); ++kk)
    for (int jj = 0; jj < n; ++jj)
      for (int ii = 0; ii < n; ++ii)
      {
        const double p[3] = { lo[0] + (ii + 0.5) * dx / n, lo[1] + (jj + 0.5) * dx / n,
                              dim == 3 ? lo[2] + (kk + 0.5) * dx / n : 0.0 };
        double       s = 0.0;
        for (int k = 0; k < dim; ++k)
          s += (p[k] - c[k]) * (p[k] - c[k]);
        inside += s <= R * R ? 1 : 0;
        ++total;
      }
  return static_cast<double>(inside) / total;
}

} // namespace

extern "C"
{

// common.h:68-144 get_region_init_state
int
g5_region_conservative(const g5_config * cfg, const g5_region * region, double * U, size_t n)
{
  if (not cfg or not region or not U or (cfg->dim != 2 and cfg->dim != 3))
    return G5_ERR_INVALID;
  if (n != static_cast<size_t>(cfg->dim == 2 ? 8 : 9))
    return G5_ERR_INVALID;
  region_to_cons(*cfg, Mix(*cfg), *region, U);
  return G5_OK;
}

// init/InitShockBubble.cpp:50-225 InitShockBubbleDataFunctor (one bubble)
int
g5_init_shock_bubble(const g5_config * cfg, const g5_region * regions3, double x_front,
                     const double * center3, double radius, double * U_host, size_t n)
{
  if (not cfg or not regions3 or not center3 or not U_host)
    return G5_ERR_INVALID;
  const int dim = cfg->dim;
  if ((dim != 2 and dim != 3) or cfg->bsize < 1)
    return G5_ERR_INVALID;
  const int    b = cfg->bsize;
  const int    nvar = dim == 2 ? 8 : 9;
  const int    ncell = dim == 2 ? b * b : b * b * b;
  const int    nx = cfg->nb[0], ny = cfg->nb[1], nz = dim == 3 ? cfg->nb[2] : 1;
  const size_t need = static_cast<size_t>(nx) * ny * nz * nvar * ncell;
  if (n != need)
    return G5_ERR_INVALID;
  const Mix m(*cfg);
  double    S[3][9];
  for (int r = 0; r < 3; ++r)
    region_to_cons(*cfg, m, regions3[r], S[r]);
  const int IA0 = dim == 2 ? 4 : 5, IAD0 = IA0 + 1, IA1 = IA0 + 2, IAD1 = IA0 + 3;
  const int IU = 2, IV = 3, IW = 4, IE = 1, ID = 0;
  // specific internal energies of the pure states (InitShockBubble.cpp:183-207)
  auto eint_of = [&](double const * s) {
    double ekin = s[IU] * s[IU] + s[IV] * s[IV];
    if (dim == 3)
      ekin += s[IW] * s[IW];
    ekin /= (2.0 * s[ID]);
    return s[IE] - ekin;
  };
  const double eint_pre = eint_of(S[1]);
  const double eint_bub = eint_of(S[2]);

  for (long blk = 0; blk < static_cast<long>(nx) * ny * nz; ++blk)
  {
    const int bx = static_cast<int>(blk % nx), by = static_cast<int>((blk / nx) % ny),
              bz = static_cast<int>(blk / (static_cast<long>(nx) * ny));
    double * Ub = U_host + static_cast<size_t>(blk) * nvar * ncell;
    for (int cell = 0; cell < ncell; ++cell)
    {
      const int    i = cell % b, j = (cell / b) % b, k = dim == 3 ? cell / (b * b) : 0;
      const double lo[3] = { cfg->xmin[0] + (bx * b + i) * cfg->dx, cfg->xmin[1] + (by * b + j) * cfg->dx,
                             cfg->xmin[2] + (bz * b + k) * cfg->dx };
      const double xc = lo[0] + 0.5 * cfg->dx;
      double const * st = xc <= x_front ? S[0] : S[1];
      double         u[9];
      for (int v = 0; v < nvar; ++v)
        u[v] = st[v];
      u[ID] = u[IAD0] + u[IAD1];
      u[IA1] = 1.0 - u[IA0];
      const double f = sphere_fraction(dim, lo, cfg->dx, center3, radius);
      if (f > 0.0)
      {
        const double phi0 = 1.0 - f;
        u[IAD0] = phi0 * S[1][IAD0] + (1.0 - phi0) * S[2][IAD0];
        u[IAD1] = phi0 * S[1][IAD1] + (1.0 - phi0) * S[2][IAD1];
        u[ID] = u[IAD0] + u[IAD1];
        u[IA0] = phi0 * S[1][IA0] + (1.0 - phi0) * S[2][IA0];
        u[IA1] = 1.0 - u[IA0];
        u[IU] = phi0 * S[1][IU] + (1.0 - phi0) * S[2][IU];
        u[IV] = phi0 * S[1][IV] + (1.0 - phi0) * S[2][IV];
        if (dim == 3)
          u[IW] = phi0 * S[1][IW] + (1.0 - phi0) * S[2][IW];
        const double rho_mixed = u[IAD0] + u[IAD1];
        double       ekin = u[IU] * u[IU] + u[IV] * u[IV];
        if (dim == 3)
          ekin += u[IW] * u[IW];
        ekin /= (2.0 * rho_mixed);
        const double eint_mixed = phi0 * eint_pre + (1.0 - phi0) * eint_bub;
        u[IE] = eint_mixed + ekin;
      }
      for (int v = 0; v < nvar; ++v)
        Ub[v * ncell + cell] = u[v];
    }
  }
  return G5_OK;
}

// init/InitTriplePoint.cpp:30-90 (InitTriplePointDataFunctor): region 0 for x <= xd, else
// region 1 where y <= yd (sqrt(y^2 + z^2) <= yd in 3-D) and region 2 elsewhere, at the cell
// centre; the same arithmetic as the device generator (kind 2 of k_init_ic)
int
g5_init_triple_point(const g5_config * cfg, const g5_region * regions3, double xd, double yd, double * U_host,
                     size_t n)
{
  if (not cfg or not regions3 or not U_host)
    return G5_ERR_INVALID;
  const int dim = cfg->dim;
  if ((dim != 2 and dim != 3) or cfg->bsize < 1)
    return G5_ERR_INVALID;
  const int    b = cfg->bsize;
  const int    nvar = dim == 2 ? 8 : 9;
  const int    ncell = dim == 2 ? b * b : b * b * b;
  const int    nx = cfg->nb[0], ny = cfg->nb[1], nz = dim == 3 ? cfg->nb[2] : 1;
  const size_t need = static_cast<size_t>(nx) * ny * nz * nvar * ncell;
  if (n != need)
    return G5_ERR_INVALID;
  const Mix m(*cfg);
  double    S[3][9];
  for (int r = 0; r < 3; ++r)
    region_to_cons(*cfg, m, regions3[r], S[r]);
  const int IA0 = dim == 2 ? 4 : 5, IAD0 = IA0 + 1, IA1 = IA0 + 2, IAD1 = IA0 + 3, ID = 0;
  for (long blk = 0; blk < static_cast<long>(nx) * ny * nz; ++blk)
  {
    const int bx = static_cast<int>(blk % nx), by = static_cast<int>((blk / nx) % ny),
              bz = static_cast<int>(blk / (static_cast<long>(nx) * ny));
    double * Ub = U_host + static_cast<size_t>(blk) * nvar * ncell;
    for (int cell = 0; cell < ncell; ++cell)
    {
      const int    i = cell % b, j = (cell / b) % b, k = dim == 3 ? cell / (b * b) : 0;
      const double lo[3] = { cfg->xmin[0] + (bx * b + i) * cfg->dx, cfg->xmin[1] + (by * b + j) * cfg->dx,
                             cfg->xmin[2] + (bz * b + k) * cfg->dx };
      const double xc = lo[0] + 0.5 * cfg->dx;
      int          r = 0;
      if (xc > xd)
      {
        const double yc = lo[1] + 0.5 * cfg->dx;
        const double zc = lo[2] + 0.5 * cfg->dx;
        const bool   in1 = dim == 2 ? (yc <= yd) : (std::sqrt(yc * yc + zc * zc) <= yd);
        r = in1 ? 1 : 2;
      }
      double u[9];
      for (int v = 0; v < nvar; ++v)
        u[v] = S[r][v];
      u[ID] = u[IAD0] + u[IAD1];
      u[IA1] = 1.0 - u[IA0];
      for (int v = 0; v < nvar; ++v)
        Ub[v * ncell + cell] = u[v];
    }
  }
  return G5_OK;
}

// Rankine-Hugoniot jump of a shock moving into a stiffened gas at rest.  With P = p + pinf the
// stiffened gas obeys the ideal-gas relations in P (app/tools/moving_shock.py:79-163 is the
// pinf = 0 case): P2/P1 = 1 + 2 gamma/(gamma+1) (M^2 - 1), rho2/rho1 = ((g+1)P2 + (g-1)P1) /
// ((g-1)P2 + (g+1)P1), W = M c1 with c1^2 = gamma P1/rho1, u2 = W (1 - rho1/rho2).
int
g5_rankine_hugoniot(double gamma, double pinf, double rho1, double p1, double mach, double p2_in,
                    double * rho2, double * u2, double * p2, double * shock_speed)
{
  if (not rho2 or not u2 or not p2 or not shock_speed or not(gamma > 1.0) or not(rho1 > 0.0))
    return G5_ERR_INVALID;
  const double P1 = p1 + pinf;
  if (not(P1 > 0.0))
    return G5_ERR_INVALID;
  const double c1 = std::sqrt(gamma * P1 / rho1);
  double       P2, M;
  if (mach > 0.0)
  {
    M = mach;
    P2 = P1 * (1.0 + 2.0 * gamma / (gamma + 1.0) * (M * M - 1.0));
  }
  else if (p2_in > p1)
  {
    P2 = p2_in + pinf;
    M = std::sqrt(1.0 + (gamma + 1.0) / (2.0 * gamma) * (P2 / P1 - 1.0));
  }
  else
    return G5_ERR_INVALID;
  const double ratio = ((gamma + 1.0) * P2 + (gamma - 1.0) * P1) / ((gamma - 1.0) * P2 + (gamma + 1.0) * P1);
  const double W = M * c1;
  *rho2 = rho1 * ratio;
  *u2 = W * (1.0 - 1.0 / ratio);
  *p2 = P2 - pinf;
  *shock_speed = W;
  return G5_OK;
}

} // extern "C"
