// g5_oracle.cpp -- CPU FP64 restatement of the reference five-equation Godunov step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the library built from it.
// The product (godunov_five_eq_amd/) never links, loads or calls it.
//
// It restates, in plain C++17 + OpenMP, the V0 implementation of the reference
// (scheme/GodunovImplemV0.cpp:83-151) with V0's own structure: a materialised outside-block ring
// filled by FillOutside, a ghosted primitive array Q_ghosted (ghost width 2), slope arrays
// (ghost width 1), a per-direction flux array and a u_star array, a read-modify-write update of
// U2 per direction, then FixVolumeFractionsNormalization and the U <- U2 copy.
// Every function cites the reference file:line it follows.  Operation order is kept literally
// (build with -ffp-contract=off) so that results are reproducible to the bit.
//
// Parity status of each part (details in the oracle's tests):
//  - slopes, MUSCL-Hancock reconstruction, HLLC, update, normalisation, dt, FillOutside:
//    restated from the reference sources; pinned by closed forms, invariants and brute force.
//  - mixture EOS: the reference calls core::eos::MieGruneisenMixture (eos/eos_utils.h:28), which is
//    absent from the reference tree.  This file implements the Allaire et al. (2002) isobaric
//    stiffened-gas mixture; pinned against the textbook pure-phase limits and round trips.
//    Exact operation order versus the absent library: parity unpinned.
//  - Youngs normal: the reference calls vof::youngs_normal_inplace (absent).  Restated from
//    Youngs (1982); pinned by planar-interface invariants only.  Parity unpinned.
//  - MILLER_PUCKETT source factors: the bulk moduli come from the same absent EOS class;
//    restated as the stiffened-gas K_k = gamma_k (p + pinf_k) and Wood's mixture rule, pinned by
//    the sum rule, pure-phase limits and equilibrium tests only.  Parity unpinned.
//  - tanh: a deterministic expm1-based tanh (Cody-Waite reduction + Taylor polynomial) used
//    identically by the HIP path; pinned against libm tanh to a few ulp.
//  - initial conditions (region states, shock-bubble, triple point): restated from
//    common.h:68-144 and init/InitShockBubble.cpp / init/InitTriplePoint.cpp; the bubble's cell
//    volume fraction uses the reference's tangent-plane construction, whose two helpers
//    (get_tangent_to_sphere, vof::compute_volume_fraction_of_rect_below_plane) are absent: restated
//    from their names' definitions and pinned by an exact clipping and the sphere-volume bound;
//    exact formulas of the absent helpers: parity unpinned.
//
// Layout: U[block][var][cell], cells x-fastest, owned blocks in the caller's order followed by
// the outside ring.  Variable ids follow models/FiveEq.h:35-50.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <algorithm>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace g5o
{

// ---------------------------------------------------------------------------------------------
// Parameters (mirrors HydroSettings models/HydroSettings.cpp:18-29, EOS material parameters,
// THINCParams keys [THINC] enabled/beta/epsilon, [mesh] boundary types, time integrator).
// ---------------------------------------------------------------------------------------------
extern "C" struct g5o_params
{
  int32_t dim;         // 2 or 3
  int32_t bsize;       // cells per block edge (bx=by=bz)
  int32_t nb[3];       // owned blocks per direction
  int32_t periodic[3]; // brick periodicity per direction
  int32_t bc[6];       // per face xmin,xmax,ymin,ymax,zmin,zmax: 0 NONE 1 PERIODIC 2 ZERO_GRADIENT 3 WALL
  int32_t integrator;  // 0 HANCOCK, 1 RK1, 2 RK2_SSP
  int32_t thinc_enabled;
  int32_t pad0;
  double  dx;          // cell length (compute_cell_length * scaling_factor)
  double  cfl;
  double  slope_type;
  double  smallr;
  double  smallc;
  double  gamma[2];    // material gamma
  double  pinf[2];     // material p_inf
  double  thinc_beta;
  double  thinc_epsilon;
  // inflow override of SolverGodunovFiveEq.cpp:997-1018: 0 none, 1 shock-bubble inlet
  // (border_conditions/BCShockBubble.h), 2 jet (border_conditions/BCJet.h)
  int32_t inlet;
  int32_t pad1;
  double  inlet_u[9]; // conservative inflow state
  double  xmin[3];    // domain origin ([mesh] xmin, ymin, zmin)
  double  jet[3];     // [jet] y, z, width
  // volume-fraction source term ([volume_fraction_advection] source_term_type):
  // 0 ALLAIRE_KOKH, 1 MILLER_PUCKETT (scheme/ReadFluxesAndConservativeUpdateFunctor.h:412-443)
  int32_t source_term;
  int32_t pad2;
};

enum
{
  BC_NONE = 0,
  BC_PERIODIC = 1,
  BC_ZERO_GRADIENT = 2,
  BC_WALL = 3
};
enum
{
  TI_HANCOCK = 0,
  TI_RK1 = 1,
  TI_RK2_SSP = 2
};

// variable ids, models/FiveEq.h:35-50
struct Ids
{
  int ID, IP, IE, IU, IV, IW, IA0, IAD0, IA1, IAD1, nbvar;
  explicit Ids(int dim)
  {
    ID = 0;
    IP = IE = 1;
    IU = 2;
    IV = 3;
    IW = (dim == 2) ? IV : IV + 1;
    const int hydro = IW + 1;
    IA0 = hydro;
    IAD0 = hydro + 1;
    IA1 = IA0 + 2;
    IAD1 = IAD0 + 2;
    nbvar = IAD1 + 1;
  }
};

struct State
{
  double v[9];
  double &       operator[](int i) { return v[i]; }
  double const & operator[](int i) const { return v[i]; }
};

// ---------------------------------------------------------------------------------------------
// Mixture stiffened-gas EOS (Allaire, Clerc, Kokh 2002 isobaric closure).
//   1/(gamma-1)        = sum_k alpha_k / (gamma_k - 1)                  (xi)
//   gamma p_inf/(g-1)  = sum_k alpha_k gamma_k p_inf_k / (gamma_k - 1)  (Pi)
//   p   = (rho e - Pi) / xi,   e = (p xi + Pi)/rho,   rho c^2 = ((xi + 1) p + Pi) / xi.
// Call signatures follow the reference call sites (models/utils.h:178,220,243;
// models/RiemannSolvers.h:84-95,106-121): (rho, e|p, alpha0, alpha1, alpha_rho0, alpha_rho1).
// ---------------------------------------------------------------------------------------------
struct Eos
{
  double g1[2]; // 1/(gamma_k - 1)
  double gp[2]; // gamma_k p_inf_k / (gamma_k - 1)
  double gamma[2], pinf[2];
  void
  init(double const gamma_[2], double const pinf_[2])
  {
    for (int k = 0; k < 2; ++k)
    {
      g1[k] = 1.0 / (gamma_[k] - 1.0);
      gp[k] = gamma_[k] * pinf_[k] / (gamma_[k] - 1.0);
      gamma[k] = gamma_[k];
      pinf[k] = pinf_[k];
    }
  }
  // Bulk moduli for the MILLER_PUCKETT source term (call sites
  // scheme/ReadFluxesAndConservativeUpdateFunctor.h:427-437; the reference's EOS class is not in
  // its tree -- PARITY UNPINNED, restated from the stiffened-gas closure): material
  // K_k = rho_k c_k^2 = gamma_k (p + pinf_k), mixture by Wood's formula 1/K = sum alpha_k / K_k.
  double
  material_bulk_modulus(int k, double p, double /*alpha_k*/, double /*alpha_rho_k*/) const
  {
    return gamma[k] * (p + pinf[k]);
  }
  double
  mixture_bulk_modulus(double /*rho*/, double p, double a0, double a1, double ar0, double ar1) const
  {
    const double k0 = material_bulk_modulus(0, p, a0, ar0);
    const double k1 = material_bulk_modulus(1, p, a1, ar1);
    return 1.0 / (a0 / k0 + a1 / k1);
  }
  double
  xi(double a0, double a1) const
  {
    return a0 * g1[0] + a1 * g1[1];
  }
  double
  pi(double a0, double a1) const
  {
    return a0 * gp[0] + a1 * gp[1];
  }
  double
  mixture_pressure(double rho, double eint, double a0, double a1, double, double) const
  {
    return (rho * eint - pi(a0, a1)) / xi(a0, a1);
  }
  double
  mixture_specific_eint(double rho, double p, double a0, double a1, double, double) const
  {
    return (p * xi(a0, a1) + pi(a0, a1)) / rho;
  }
  double
  mixture_sound_speed(double rho, double p, double a0, double a1, double, double) const
  {
    const double x = xi(a0, a1);
    const double rc2 = ((x + 1.0) * p + pi(a0, a1)) / x;
    return std::sqrt(rc2 / rho);
  }
};

struct Settings
{
  double cfl, slope_type, smallr, smallc, smallp;
};

// ---------------------------------------------------------------------------------------------
// Deterministic tanh: tanh(a) = -m/(2+m), m = expm1(-2a), a=|x|; expm1 by Cody-Waite reduction
// y = k ln2 + r, |r| <= ln2/2, and the degree-13 Taylor polynomial of expm1(r).
// ---------------------------------------------------------------------------------------------
static double
det_expm1_neg(double y) // y <= 0
{
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double inv_ln2 = 1.44269504088896338700e+00;
  double       k = std::nearbyint(y * inv_ln2);
  if (k < -60.0)
    k = -60.0;
  const double r = (y - k * ln2_hi) - k * ln2_lo;
  // Horner for r*(1 + r/2 + r^2/6 + ... + r^12/13!)
  double p = 1.6059043836821613e-10;     // 1/13!
  p = p * r + 2.08767569878681e-09;    // 1/12!
  p = p * r + 2.505210838544172e-08;    // 1/11!
  p = p * r + 2.755731922398589e-07;    // 1/10!
  p = p * r + 2.7557319223985893e-06;    // 1/9!
  p = p * r + 2.48015873015873e-05;    // 1/8!
  p = p * r + 1.98412698412698413e-04;    // 1/7!
  p = p * r + 1.38888888888888894e-03;    // 1/6!
  p = p * r + 8.33333333333333322e-03;    // 1/5!
  p = p * r + 4.16666666666666644e-02;    // 1/4!
  p = p * r + 1.66666666666666657e-01;    // 1/3!
  p = p * r + 5.00000000000000000e-01;    // 1/2!
  p = p * r + 1.0;
  p = p * r; // expm1(r)
  if (k == 0.0)
    return p;
  const double two_k = std::ldexp(1.0, static_cast<int>(k));
  return two_k * p + (two_k - 1.0);
}

double
det_tanh(double x)
{
#ifdef G5O_LIBM_TANH
  return std::tanh(x); // the C library's tanh (variant build, see oracle/Makefile)
#endif
  const double a = std::fabs(x);
  if (a > 22.0)
    return x > 0 ? 1.0 : -1.0;
  const double m = det_expm1_neg(-2.0 * a);
  const double t = -m / (2.0 + m);
  return x < 0 ? -t : t;
}

// ---------------------------------------------------------------------------------------------
// models/utils.h:36-73 hydro_sanitizer ; :82-87 sanitize_alpha
// ---------------------------------------------------------------------------------------------
static void
hydro_sanitizer(State & q, Ids const & H)
{
  const double eps = 1e-12;
  if (q[H.IAD0] < eps)
    q[H.IAD0] = 0.0;
  if (q[H.IAD1] < eps)
    q[H.IAD1] = 0.0;
  if (q[H.IA0] < eps)
    q[H.IA0] = 0.0;
  if (q[H.IA0] > 1.0 - eps)
    q[H.IA0] = 1.0;
  if (q[H.IA1] < eps)
    q[H.IA1] = 0.0;
  if (q[H.IA1] > 1.0 - eps)
    q[H.IA1] = 1.0;
}

static inline double
sanitize_alpha(double alpha)
{
  return std::fmax(0.0, std::fmin(1.0, alpha));
}

// models/utils.h:124-151 compute_mixture_ekin
static double
compute_mixture_ekin(State const & u, Settings const & s, int dim, Ids const & H)
{
  const double d = std::fmax(u[H.ID], s.smallr);
  if (dim == 2)
  {
    const double ux = u[H.IU] / d;
    const double uy = u[H.IV] / d;
    return 0.5 * (ux * ux + uy * uy);
  }
  const double ux = u[H.IU] / d;
  const double uy = u[H.IV] / d;
  const double uz = u[H.IW] / d;
  return 0.5 * (ux * ux + uy * uy + uz * uz);
}

// models/utils.h:160-182 compute_mixture_pressure: density alpha_rho0 + alpha_rho1, kinetic
// energy with U[ID] (compute_mixture_ekin), alpha1 = 1 - alpha0
static double
compute_mixture_pressure(State const & u, Settings const & s, Eos const & eos, int dim, Ids const & H)
{
  const double d = std::fmax(u[H.IAD0] + u[H.IAD1], s.smallr);
  const double ekin = compute_mixture_ekin(u, s, dim, H);
  const double eint = u[H.IE] / d - ekin;
  return eos.mixture_pressure(d, eint, u[H.IA0], 1.0 - u[H.IA0], u[H.IAD0], u[H.IAD1]);
}

// scheme/ReadFluxesAndConservativeUpdateFunctor.h:412-443 get_source_term_factor
static double
source_term_factor(int source_term, State const & u, Settings const & s, Eos const & eos, int dim, Ids const & H,
                   int mat)
{
  if (source_term != 1)
    return 1.0; // ALLAIRE_KOKH
  const double pressure = compute_mixture_pressure(u, s, eos, dim, H);
  const double kappa0 = mat == 0 ? eos.material_bulk_modulus(0, pressure, u[H.IA0], u[H.IAD0])
                                 : eos.material_bulk_modulus(1, pressure, u[H.IA1], u[H.IAD1]);
  const double kappa_bar = eos.mixture_bulk_modulus(u[H.ID], pressure, u[H.IA0], u[H.IA1], u[H.IAD0], u[H.IAD1]);
  return 1.0 - kappa_bar / kappa0;
}

// models/utils.h:200-250 computePrimitives ; returns c
double
compute_primitives(State const & u, State & q, Settings const & s, Eos const & eos, int dim,
                   Ids const & H)
{
  const double d = std::fmax(u[H.ID], s.smallr);
  const double ux = u[H.IU] / d;
  const double uy = u[H.IV] / d;
  const double ekin = compute_mixture_ekin(u, s, dim, H);
  const double eint = u[H.IE] / d - ekin;
  const double p =
    eos.mixture_pressure(d, eint, u[H.IA0], 1.0 - u[H.IA0], u[H.IAD0], u[H.IAD1]);
  q[H.ID] = d;
  q[H.IP] = p;
  q[H.IU] = ux;
  q[H.IV] = uy;
  if (dim == 3)
  {
    const double uz = u[H.IW] / d;
    q[H.IW] = uz;
  }
  q[H.IA0] = sanitize_alpha(u[H.IA0]);
  q[H.IAD0] = u[H.IAD0];
  q[H.IA1] = 1.0 - q[H.IA0];
  q[H.IAD1] = u[H.IAD1];
  hydro_sanitizer(q, H);
  return eos.mixture_sound_speed(d, p, q[H.IA0], 1.0 - q[H.IA0], u[H.IAD0], u[H.IAD1]);
}

// ---------------------------------------------------------------------------------------------
// models/RiemannSolvers.h:57-262 riemann_hllc (wrapped by riemann_hydro :481-500)
// ---------------------------------------------------------------------------------------------
struct Riemann
{
  State  flux;
  double ustar, phistar0, phistar1;
};

Riemann
riemann_hllc(State const & ql, State const & qr, Settings const & s, Eos const & eos, int dim,
             Ids const & H)
{
  Riemann       rs;
  State &       flux = rs.flux;
  double &      ustar = rs.ustar;
  const double smallr = s.smallr, smallp = s.smallp, smallc = s.smallc;

  const double rl = std::fmax(ql[H.IAD0] + ql[H.IAD1], smallr);
  double       pl = std::fmax(ql[H.IP], rl * smallp);
  const double ul = ql[H.IU];
  const double eintl =
    eos.mixture_specific_eint(rl, pl, ql[H.IA0], 1.0 - ql[H.IA0], ql[H.IAD0], ql[H.IAD1]);
  pl = eos.mixture_pressure(rl, eintl, ql[H.IA0], 1.0 - ql[H.IA0], ql[H.IAD0], ql[H.IAD1]);
  double ecinl = 0.5 * rl * ul * ul;
  ecinl += 0.5 * rl * ql[H.IV] * ql[H.IV];
  if (dim == 3)
    ecinl += 0.5 * rl * ql[H.IW] * ql[H.IW];
  const double cl =
    eos.mixture_sound_speed(rl, pl, ql[H.IA0], 1.0 - ql[H.IA0], ql[H.IAD0], ql[H.IAD1]);
  const double etotl = rl * eintl + ecinl;
  const double ptotl = pl;

  const double rr = std::fmax(qr[H.IAD0] + qr[H.IAD1], smallr);
  double       pr = std::fmax(qr[H.IP], rr * smallp);
  const double ur = qr[H.IU];
  const double eintr =
    eos.mixture_specific_eint(rr, pr, qr[H.IA0], 1.0 - qr[H.IA0], qr[H.IAD0], qr[H.IAD1]);
  pr = eos.mixture_pressure(rr, eintr, qr[H.IA0], 1.0 - qr[H.IA0], qr[H.IAD0], qr[H.IAD1]);
  double ecinr = 0.5 * rr * ur * ur;
  ecinr += 0.5 * rr * qr[H.IV] * qr[H.IV];
  if (dim == 3)
    ecinr += 0.5 * rr * qr[H.IW] * qr[H.IW];
  const double cr =
    eos.mixture_sound_speed(rr, pr, qr[H.IA0], 1.0 - qr[H.IA0], qr[H.IAD0], qr[H.IAD1]);
  const double etotr = rr * eintr + ecinr;
  const double ptotr = pr;

  const double cfastl = std::sqrt(std::fmax(cl * cl, smallc * smallc));
  const double cfastr = std::sqrt(std::fmax(cr * cr, smallc * smallc));
  const double SL = std::fmin(ul, ur) - std::fmax(cfastl, cfastr);
  const double SR = std::fmax(ul, ur) + std::fmax(cfastl, cfastr);

  const double rcl = rl * (ul - SL);
  const double rcr = rr * (SR - ur);

  ustar = (rcr * ur + rcl * ul + (ptotl - ptotr)) / (rcr + rcl);
  const double ptotstar = (rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);

  const double factorL = (SL - ul) / (SL - ustar);
  const double etotstarl = ((SL - ul) * etotl - ptotl * ul + ptotstar * ustar) / (SL - ustar);
  const double factorR = (SR - ur) / (SR - ustar);
  const double etotstarr = ((SR - ur) * etotr - ptotr * ur + ptotstar * ustar) / (SR - ustar);

  double ro, ro_0, ro_1, uo, ptoto, etoto;
  if (SL > 0.0)
  {
    ro_0 = ql[H.IAD0];
    ro_1 = ql[H.IAD1];
    ro = rl;
    uo = ul;
    ptoto = ptotl;
    etoto = etotl;
  }
  else if (ustar > 0.0)
  {
    ro_0 = ql[H.IAD0] * factorL;
    ro_1 = ql[H.IAD1] * factorL;
    ro = rl * factorL;
    uo = ustar;
    ptoto = ptotstar;
    etoto = etotstarl;
  }
  else if (SR > 0.0)
  {
    ro_0 = qr[H.IAD0] * factorR;
    ro_1 = qr[H.IAD1] * factorR;
    ro = rr * factorR;
    uo = ustar;
    ptoto = ptotstar;
    etoto = etotstarr;
  }
  else
  {
    ro_0 = qr[H.IAD0];
    ro_1 = qr[H.IAD1];
    ro = rr;
    uo = ur;
    ptoto = ptotr;
    etoto = etotr;
  }

  flux[H.IAD0] = ro_0 * uo;
  flux[H.IAD1] = ro_1 * uo;
  flux[H.ID] = (ro_0 + ro_1) * uo;
  flux[H.IU] = ro * uo * uo + ptoto;
  flux[H.IP] = (etoto + ptoto) * uo;
  if (flux[H.ID] > 0.0)
    flux[H.IV] = flux[H.ID] * ql[H.IV];
  else
    flux[H.IV] = flux[H.ID] * qr[H.IV];
  if (dim == 3)
  {
    if (flux[H.ID] > 0.0)
      flux[H.IW] = flux[H.ID] * ql[H.IW];
    else
      flux[H.IW] = flux[H.ID] * qr[H.IW];
  }
  if (flux[H.ID] > 0.0)
  {
    rs.phistar0 = ql[H.IA0];
    rs.phistar1 = 1.0 - ql[H.IA0];
  }
  else
  {
    rs.phistar0 = qr[H.IA0];
    rs.phistar1 = 1.0 - qr[H.IA0];
  }
  ustar = uo;
  // the flux slots IA0/IA1 are set by the caller (scheme/ComputeFluxesAndStoreFunctor.cpp:400-401)
  flux[H.IA0] = 0.0;
  flux[H.IA1] = 0.0;
  return rs;
}

// scheme/ComputeLimitedSlopesFunctor.cpp:114-137 slope_unsplit_scalar
double
slope_unsplit_scalar(double q, double qPlus, double qMinus, double slope_type)
{
  double dq = 0;
  if (slope_type == 1 or slope_type == 2)
  {
    const double dlft = slope_type * (q - qMinus);
    const double drgt = slope_type * (qPlus - q);
    const double dcen = 0.5 * (qPlus - qMinus);
    const double dsgn = (dcen >= 0.0) ? 1.0 : -1.0;
    const double slop = std::fmin(std::fabs(dlft), std::fabs(drgt));
    const double dlim = (dlft * drgt) <= 0.0 ? 0.0 : slop;
    dq = dsgn * std::fmin(dlim, std::fabs(dcen));
  }
  return dq;
}

// ---------------------------------------------------------------------------------------------
// Youngs (1982) interface normal on the 3^d alpha stencil, unit cell size: the gradient of alpha
// averaged from the 2^d corners of the cell, i.e. central differences with (1,2,1) weights in
// each transverse direction, normalised to unit length (zero vector if the gradient vanishes).
// a[] holds alpha on the stencil with index (di+1) + 3*(dj+1) [+ 9*(dk+1)].
// ---------------------------------------------------------------------------------------------
void
youngs_normal(int dim, double const * a, double * n)
{
  if (dim == 2)
  {
    auto A = [&](int di, int dj) { return a[(di + 1) + 3 * (dj + 1)]; };
    const double nx = ((A(1, -1) + 2.0 * A(1, 0)) + A(1, 1)) - ((A(-1, -1) + 2.0 * A(-1, 0)) + A(-1, 1));
    const double ny = ((A(-1, 1) + 2.0 * A(0, 1)) + A(1, 1)) - ((A(-1, -1) + 2.0 * A(0, -1)) + A(1, -1));
    const double nrm = std::sqrt(nx * nx + ny * ny);
    if (nrm > 0.0)
    {
      n[0] = nx / nrm;
      n[1] = ny / nrm;
    }
    else
    {
      n[0] = 0.0;
      n[1] = 0.0;
    }
    return;
  }
  auto A = [&](int di, int dj, int dk) { return a[(di + 1) + 3 * (dj + 1) + 9 * (dk + 1)]; };
  // weighted sum over the transverse 3x3 plane at offset s along direction d
  auto plane = [&](int d, int s) {
    double acc = 0.0;
    for (int t2 = -1; t2 <= 1; ++t2)
      for (int t1 = -1; t1 <= 1; ++t1)
      {
        const double w = (t1 == 0 ? 2.0 : 1.0) * (t2 == 0 ? 2.0 : 1.0);
        int          o[3];
        o[d] = s;
        o[(d + 1) % 3] = t1;
        o[(d + 2) % 3] = t2;
        acc = acc + w * A(o[0], o[1], o[2]);
      }
    return acc;
  };
  const double nx = plane(0, 1) - plane(0, -1);
  const double ny = plane(1, 1) - plane(1, -1);
  const double nz = plane(2, 1) - plane(2, -1);
  const double nrm = std::sqrt(nx * nx + ny * ny + nz * nz);
  if (nrm > 0.0)
  {
    n[0] = nx / nrm;
    n[1] = ny / nrm;
    n[2] = nz / nrm;
  }
  else
  {
    n[0] = n[1] = n[2] = 0.0;
  }
}

// =============================================================================================
// Solver: mesh, arrays, V0 step
// =============================================================================================
// THINC face value of alpha (ComputeFluxesAndStoreTHINCFunctor.cpp:176-196, Zhang JCP 2025
// eq. 43): am, ac, ap = alpha at -1, 0, +1 along the face axis, nd = |normal component|,
// bb = +1 for the right face of the cell, -1 for the left one.
static double
thinc_face_alpha(double am, double ac, double ap, double nd, double thinc_beta, int bb)
{
  const double beta = thinc_beta * nd + 0.01;
  const double A = (ap + am) / 2;
  const double B = (ap - am) / 2;
  const double C = det_tanh(0.5 * beta);
  const double phi = ac;
  const double D = det_tanh(0.5 * beta * (phi - A) / B);
  return A + bb * B * (C + bb * D / C) / (1.0 + bb * D);
}

struct Solver
{
  g5o_params P;
  Ids        H;
  Eos        eos;
  Settings   s;
  int        dim, b, nvar;
  int        nowned, noutside, ntotal;
  int        ncell;  // b^dim
  int        ncellg; // (b+4)^dim
  int        ncells; // (b+2)^dim
  int        ncellf; // (b+1) b^(dim-1)
  std::vector<int32_t> bcoord;  // 3 per block (owned then outside), outside coords may be -1 or nb
  std::vector<int32_t> lookup;  // extended grid (nb+2)^dim -> block index or -1
  std::vector<double>  U, U2, Q, Sx, Sy, Sz, F, Us, Urk;
  // Two-level static mesh (level 0: the brick's blocks; level 1: coarse blocks refined into 2^dim
  // blocks of half the cell length, at least one block away from the brick's border).  bcoord
  // holds a block's coordinates at its own level; refined coarse positions are -1 in lookup.
  bool                 amr = false;
  int                  prolong = 0;  // 0 SIMPLE_COPY, 1 EXTRAPOLATE_LINEAR_MINMOD, 2/3 conservative order 2/4
  std::vector<int32_t> blev;         // level of every block (outside blocks: 0)
  std::vector<int32_t> lookup1;      // fine grid (2 nb)^dim -> owned block index or -1
  std::vector<char>    refined0;     // coarse grid nb^dim: refined

  explicit Solver(g5o_params const & p, std::vector<int32_t> const & refined = {}, int prolongation = 0)
    : P(p)
    , H(p.dim)
  {
    amr = not refined.empty();
    prolong = prolongation;
    refined0.assign(static_cast<size_t>(nbd_of(p, 0)) * nbd_of(p, 1) * nbd_of(p, 2), 0);
    for (size_t k = 0; k + 2 < refined.size(); k += 3)
      refined0[refined[k] + nbd_of(p, 0) * (refined[k + 1] + nbd_of(p, 1) * (p.dim == 3 ? refined[k + 2] : 0))] = 1;
    dim = p.dim;
    b = p.bsize;
    nvar = H.nbvar;
    eos.init(p.gamma, p.pinf);
    s.cfl = p.cfl;
    s.slope_type = p.slope_type;
    s.smallr = p.smallr;
    s.smallc = p.smallc;
    s.smallp = p.smallc * p.smallc; // models/HydroSettings.cpp:24
    ncell = ipow(b, dim);
    ncellg = ipow(b + 4, dim);
    ncells = ipow(b + 2, dim);
    ncellf = (b + 1) * ipow(b, dim - 1);
    build_mesh();
  }

  static int
  nbd_of(g5o_params const & p, int d)
  {
    return d < p.dim ? p.nb[d] : 1;
  }
  static int
  ipow(int a, int e)
  {
    int r = 1;
    for (int i = 0; i < e; ++i)
      r *= a;
    return r;
  }
  int
  nbd(int d) const
  {
    return d < dim ? P.nb[d] : 1;
  }
  int
  ext_index(int x, int y, int z) const
  {
    const int ex = P.nb[0] + 2, ey = P.nb[1] + 2;
    return (x + 1) + ex * ((y + 1) + ey * (dim == 3 ? (z + 1) : 0));
  }

  // Owned blocks in lexicographic order (x fastest), then the outside ring (blocks whose
  // coordinates leave the brick along at least one non-periodic direction).
  void
  build_mesh()
  {
    const int nx = nbd(0), ny = nbd(1), nz = nbd(2);
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x)
          if (not refined0[x + nx * (y + ny * z)])
          {
            bcoord.push_back(x);
            bcoord.push_back(y);
            bcoord.push_back(z);
            blev.push_back(0);
          }
    // fine blocks: lexicographic over the fine grid, those whose coarse parent is refined
    const int fz = dim == 3 ? 2 * nz : 1;
    lookup1.assign(static_cast<size_t>(2 * nx) * (2 * ny) * fz, -1);
    for (int z = 0; z < fz; ++z)
      for (int y = 0; y < 2 * ny; ++y)
        for (int x = 0; x < 2 * nx; ++x)
          if (refined0[x / 2 + nx * (y / 2 + ny * (dim == 3 ? z / 2 : 0))])
          {
            lookup1[x + 2 * nx * (y + 2 * ny * z)] = static_cast<int32_t>(bcoord.size() / 3);
            bcoord.push_back(x);
            bcoord.push_back(y);
            bcoord.push_back(z);
            blev.push_back(1);
          }
    nowned = static_cast<int>(bcoord.size() / 3);
    const int zlo = dim == 3 ? -1 : 0, zhi = dim == 3 ? nz : 0;
    for (int z = zlo; z <= zhi; ++z)
      for (int y = -1; y <= ny; ++y)
        for (int x = -1; x <= nx; ++x)
        {
          const int c[3] = { x, y, z };
          bool      outside = false, invalid = false;
          for (int d = 0; d < dim; ++d)
          {
            if (c[d] < 0 or c[d] >= nbd(d))
            {
              if (P.periodic[d])
                invalid = true; // periodic directions have no outside blocks
              else
                outside = true;
            }
          }
          if (outside and not invalid)
          {
            bcoord.push_back(x);
            bcoord.push_back(y);
            bcoord.push_back(z);
            blev.push_back(0);
          }
        }
    ntotal = static_cast<int>(bcoord.size() / 3);
    noutside = ntotal - nowned;
    const int ez = dim == 3 ? nz + 2 : 1;
    lookup.assign(static_cast<size_t>(nx + 2) * (ny + 2) * ez, -1);
    for (int ib = 0; ib < ntotal; ++ib)
      if (blev[ib] == 0)
        lookup[ext_index(bcoord[3 * ib], bcoord[3 * ib + 1], bcoord[3 * ib + 2])] = ib;
    U.assign(static_cast<size_t>(ntotal) * nvar * ncell, 0.0);
    U2.assign(static_cast<size_t>(nowned) * nvar * ncell, 0.0);
    Q.assign(static_cast<size_t>(nowned) * nvar * ncellg, 0.0);
    Sx.assign(static_cast<size_t>(nowned) * nvar * ncells, 0.0);
    Sy.assign(Sx.size(), 0.0);
    if (dim == 3)
      Sz.assign(Sx.size(), 0.0);
    F.assign(static_cast<size_t>(nowned) * nvar * ncellf, 0.0);
    Us.assign(static_cast<size_t>(nowned) * ncellf, 0.0);
  }

  // neighbour block of block ib shifted by o (periodic wrap), or -1
  int
  neighbor(int ib, int const o[3]) const
  {
    int c[3] = { bcoord[3 * ib], bcoord[3 * ib + 1], bcoord[3 * ib + 2] };
    for (int d = 0; d < dim; ++d)
    {
      c[d] += o[d];
      if (P.periodic[d])
      {
        const int n = nbd(d);
        c[d] = ((c[d] % n) + n) % n;
      }
      if (c[d] < -1 or c[d] > nbd(d))
        return -1;
    }
    if (dim == 2)
      c[2] = 0;
    return lookup[ext_index(c[0], c[1], c[2])];
  }

  // accessors
  double &
  u(std::vector<double> & A, int ib, int var, int cell)
  {
    return A[(static_cast<size_t>(ib) * nvar + var) * ncell + cell];
  }
  int
  cidx(int i, int j, int k) const
  {
    return i + b * (j + (dim == 3 ? b * k : 0));
  }
  // ghosted index with shift -g, extent b+2g
  int
  gidx(int i, int j, int k, int g) const
  {
    const int e = b + 2 * g;
    return (i + g) + e * ((j + g) + (dim == 3 ? e * (k + g) : 0));
  }
  double &
  q(int ib, int var, int i, int j, int k)
  {
    return Q[(static_cast<size_t>(ib) * nvar + var) * ncellg + gidx(i, j, k, 2)];
  }
  double &
  sl(std::vector<double> & S, int ib, int var, int i, int j, int k)
  {
    return S[(static_cast<size_t>(ib) * nvar + var) * ncells + gidx(i, j, k, 1)];
  }
  // flux array of direction dir: extent (b+1) along dir, b otherwise, x fastest
  int
  fidx(int dir, int i, int j, int k) const
  {
    const int ex = dir == 0 ? b + 1 : b;
    const int ey = dir == 1 ? b + 1 : b;
    return i + ex * (j + (dim == 3 ? ey * k : 0));
  }
  double &
  fl(int ib, int var, int dir, int i, int j, int k)
  {
    return F[(static_cast<size_t>(ib) * nvar + var) * ncellf + fidx(dir, i, j, k)];
  }
  double &
  us(int ib, int dir, int i, int j, int k)
  {
    return Us[static_cast<size_t>(ib) * ncellf + fidx(dir, i, j, k)];
  }

  State
  load_cons(std::vector<double> & A, int ib, int cell)
  {
    // common.h:191-213 get_conservative_variables
    State st{};
    for (int v = 0; v < nvar; ++v)
      st[v] = u(A, ib, v, cell);
    return st;
  }

  // ---- two-level mesh ------------------------------------------------------------------------
  // cell length of a block's level (compute_cell_length(level, bx) * scaling_factor: halved per level)
  double
  dx_of(int ib) const
  {
    return blev[ib] ? 0.5 * P.dx : P.dx;
  }
  static int
  floor_div(int a, int n)
  {
    return a >= 0 ? a / n : -((-a + n - 1) / n);
  }

  // Conservative state at global cell g of level l (the getNeighLoc lookup of StencilHelper,
  // ConvertToPrimitivesVariablesFunctor.cpp:876-918, for a two-level mesh):
  //  - level 0, a coarse leaf (or the outside ring, periodic wrap): the cell itself;
  //  - level 0 inside a refined block: restriction, get_conservative_vars_restriction (:214-250),
  //    the average of the 2^dim fine siblings (compute_siblings_average of the absent core:
  //    summed in lexicographic order, x fastest, then divided by 2^dim) for IE, IU, IV, IW, IAD0,
  //    IAD1, IA0, with IA1 = 1 - IA0 and ID = IAD0 + IAD1;
  //  - level 1 outside the fine blocks: prolongation of the coarse parent (:889-910), SIMPLE_COPY
  //    or EXTRAPOLATE_LINEAR_MINMOD (linear_extrapolate_using_limited_slopes, :553-770).
  State
  cons_level(std::vector<double> & A, int l, int const g[3])
  {
    if (l == 0)
    {
      int c[3] = { 0, 0, 0 }, cell[3] = { 0, 0, 0 };
      for (int d = 0; d < dim; ++d)
      {
        c[d] = floor_div(g[d], b);
        cell[d] = g[d] - c[d] * b;
        if (P.periodic[d])
        {
          const int n = nbd(d);
          c[d] = ((c[d] % n) + n) % n;
        }
      }
      const int ib = lookup[ext_index(c[0], c[1], c[2])];
      if (ib >= 0)
        return load_cons(A, ib, cidx(cell[0], cell[1], cell[2]));
      return restriction(A, g);
    }
    int c[3] = { 0, 0, 0 }, cell[3] = { 0, 0, 0 };
    for (int d = 0; d < dim; ++d)
    {
      c[d] = floor_div(g[d], b);
      cell[d] = g[d] - c[d] * b;
    }
    const int ib = lookup1[c[0] + 2 * nbd(0) * (c[1] + 2 * nbd(1) * c[2])];
    if (ib >= 0)
      return load_cons(A, ib, cidx(cell[0], cell[1], cell[2]));
    return prolongation(A, g);
  }

  // coarse cell g covered by fine blocks: average of its 2^dim fine cells
  State
  restriction(std::vector<double> & A, int const g[3])
  {
    // the averaged variables (IW only in 3-D: in 2-D its id is IV's)
    const int vars[7] = { H.IE, H.IU, H.IV, H.IAD0, H.IAD1, H.IA0, H.IW };
    const int nv = dim == 3 ? 7 : 6;
    State     sum{};
    bool      first = true;
    for (int kz = 0; kz < (dim == 3 ? 2 : 1); ++kz)
      for (int ky = 0; ky < 2; ++ky)
        for (int kx = 0; kx < 2; ++kx)
        {
          const int   f[3] = { 2 * g[0] + kx, 2 * g[1] + ky, dim == 3 ? 2 * g[2] + kz : 0 };
          const State u1 = cons_level(A, 1, f);
          for (int t = 0; t < nv; ++t)
            sum[vars[t]] = first ? u1[vars[t]] : sum[vars[t]] + u1[vars[t]];
          first = false;
        }
    const double vr = dim == 3 ? 8.0 : 4.0;
    State        u{};
    for (int t = 0; t < nv; ++t)
      u[vars[t]] = sum[vars[t]] / vr;
    u[H.IA1] = 1.0 - u[H.IA0];
    u[H.ID] = u[H.IAD0] + u[H.IAD1];
    return u;
  }

  // fine cell g outside the fine blocks: its coarse parent, copied or extrapolated with minmod
  // slopes of the conservative variables (slope_type 1; the neighbours of the parent at the
  // coarse level, restricted where they are refined) at the fine cell's side -1/+1 of the parent
  State
  prolongation(std::vector<double> & A, int const g[3])
  {
    int gc[3] = { 0, 0, 0 }, ix[3] = { 0, 0, 0 };
    for (int d = 0; d < dim; ++d)
    {
      gc[d] = floor_div(g[d], 2);
      ix[d] = (g[d] - 2 * gc[d]) == 1 ? 1 : -1;
    }
    const State center = cons_level(A, 0, gc);
    if (prolong == 0)
      return center;
    if (prolong >= 2)
      return conservative_interpolation(A, g, gc);
    State slope[3];
    for (int d = 0; d < dim; ++d)
    {
      int gl[3] = { gc[0], gc[1], gc[2] }, gr[3] = { gc[0], gc[1], gc[2] };
      gl[d] -= 1;
      gr[d] += 1;
      const State left = cons_level(A, 0, gl), right = cons_level(A, 0, gr);
      for (int v = 0; v < nvar; ++v)
        slope[d][v] = slope_unsplit_scalar(center[v], right[v], left[v], 1.0);
    }
    // the extrapolated variables (IW only in 3-D: in 2-D its id is IV's)
    const int vars[7] = { H.IAD0, H.IAD1, H.IA0, H.IE, H.IU, H.IV, H.IW };
    State     u{};
    for (int t = 0; t < (dim == 3 ? 7 : 6); ++t)
    {
      const int v = vars[t];
      double val = center[v] + 0.25 * static_cast<double>(ix[0]) * slope[0][v] +
                   0.25 * static_cast<double>(ix[1]) * slope[1][v];
      if (dim == 3)
        val = center[v] + 0.25 * static_cast<double>(ix[0]) * slope[0][v] +
              0.25 * static_cast<double>(ix[1]) * slope[1][v] + 0.25 * static_cast<double>(ix[2]) * slope[2][v];
      u[v] = val;
    }
    u[H.ID] = u[H.IAD0] + u[H.IAD1];
    u[H.IA1] = 1.0 - u[H.IA0];
    return u;
  }

  // CONSERVATIVE_INTERPOLATION_ORDER_2 / _4 (prolong 2 / 3; conservative_interpolation_order{2,4}
  // and interpol_order{2,4}, ConvertToPrimitivesVariablesFunctor.cpp:339-490,751-827): the fine
  // cell's half of its coarse parent along each direction (position 0 left, 1 right), by the
  // tensor product of 1-D conservative interpolations over the parent's 3^dim (5^dim) coarse
  // neighbourhood, x first, then y (then z).  The weights are the averages over the left half of
  // the polynomial of degree 2 (4) with the given cell averages: (1/8, 1, -1/8) and
  // (-3/128, 22/128, 1, -22/128, 3/128), mirrored for the right half (the absent core's
  // coefficient tables restated from that definition); each combination sums in stencil order.
  State
  conservative_interpolation(std::vector<double> & A, int const g[3], int const gc[3])
  {
    static const double w2[3] = { 0.125, 1.0, -0.125 };
    static const double w4[5] = { -3.0 / 128.0, 22.0 / 128.0, 1.0, -22.0 / 128.0, 3.0 / 128.0 };
    const int      m = prolong == 2 ? 1 : 2, n = 2 * m + 1;
    const double * wl = prolong == 2 ? w2 : w4;
    int            pos[3] = { 0, 0, 0 };
    for (int d = 0; d < dim; ++d)
      pos[d] = g[d] - 2 * gc[d];
    auto w = [&](int d, int k) { return pos[d] == 0 ? wl[k] : wl[n - 1 - k]; };
    // the coarse states of the stencil, once
    std::vector<State> cs(static_cast<size_t>(n) * n * (dim == 3 ? n : 1));
    for (int kz = 0; kz < (dim == 3 ? n : 1); ++kz)
      for (int ky = 0; ky < n; ++ky)
        for (int kx = 0; kx < n; ++kx)
        {
          const int c[3] = { gc[0] + kx - m, gc[1] + ky - m, dim == 3 ? gc[2] + kz - m : 0 };
          cs[kx + n * (ky + n * kz)] = cons_level(A, 0, c);
        }
    const int vars[7] = { H.IAD0, H.IAD1, H.IA0, H.IE, H.IU, H.IV, H.IW };
    State     u{};
    for (int t = 0; t < (dim == 3 ? 7 : 6); ++t)
    {
      const int v = vars[t];
      double    valz = 0.0;
      for (int kz = 0; kz < (dim == 3 ? n : 1); ++kz)
      {
        double valy = 0.0;
        for (int ky = 0; ky < n; ++ky)
        {
          double valx = 0.0;
          for (int kx = 0; kx < n; ++kx)
          {
            const double x = w(0, kx) * cs[kx + n * (ky + n * kz)][v];
            valx = kx == 0 ? x : valx + x;
          }
          valy = ky == 0 ? w(1, ky) * valx : valy + w(1, ky) * valx;
        }
        valz = dim == 3 ? (kz == 0 ? w(2, kz) * valy : valz + w(2, kz) * valy) : valy;
      }
      u[v] = valz;
    }
    u[H.ID] = u[H.IAD0] + u[H.IAD1];
    u[H.IA1] = 1.0 - u[H.IA0];
    return u;
  }

  // the coarse-level neighbour of owned coarse block ib across face (dir, side) is refined
  // (conformal status NEIGHBOR_IS_FINER)
  bool
  finer_neighbor(int ib, int dir, int side) const
  {
    if (not amr or blev[ib] != 0)
      return false;
    int c[3] = { bcoord[3 * ib], bcoord[3 * ib + 1], bcoord[3 * ib + 2] };
    c[dir] += side;
    for (int d = 0; d < dim; ++d)
      if (c[d] < 0 or c[d] >= nbd(d))
        return false;
    return refined0[c[0] + nbd(0) * (c[1] + nbd(1) * (dim == 3 ? c[2] : 0))] != 0;
  }

  // get_flux_from_fine_neighbor / get_ustar_flux_from_fine_neighbor
  // (ReadFluxesAndConservativeUpdateFunctor.h:246-307): the sum of the fluxes (or of u*) of the
  // 2^(dim-1) fine faces covering the coarse face of cell (i, j, k) of block ib on side `side` of
  // direction dir (compute_face_siblings_sum of the absent core: summed with the first transverse
  // index fastest); var < 0: u*
  double
  fine_face_sum(int ib, int dir, int side, int i, int j, int k, int var)
  {
    const int ci[3] = { i, j, k };
    int       gc[3];
    for (int d = 0; d < 3; ++d)
      gc[d] = bcoord[3 * ib + d] * b + ci[d];
    const int t1 = (dir + 1) % dim, t2 = dim == 3 ? (dir + 2) % 3 : -1;
    double    sum = 0.0;
    bool      first = true;
    for (int k2 = 0; k2 < (dim == 3 ? 2 : 1); ++k2)
      for (int k1 = 0; k1 < 2; ++k1)
      {
        int f[3] = { 2 * gc[0], 2 * gc[1], dim == 3 ? 2 * gc[2] : 0 };
        // the fine cells across the face: left of it (side -1) or right of it (side +1)
        f[dir] = side < 0 ? 2 * gc[dir] - 1 : 2 * gc[dir] + 2;
        f[t1] += k1;
        if (t2 >= 0)
          f[t2] += k2;
        int fb[3] = { 0, 0, 0 }, fl_[3] = { 0, 0, 0 };
        for (int d = 0; d < dim; ++d)
        {
          fb[d] = floor_div(f[d], b);
          fl_[d] = f[d] - fb[d] * b;
        }
        const int jb = lookup1[fb[0] + 2 * nbd(0) * (fb[1] + 2 * nbd(1) * fb[2])];
        // the fine cell's right face (use_right_flux) for side -1, its left face for side +1
        int fi[3] = { fl_[0], fl_[1], fl_[2] };
        if (side < 0)
          fi[dir] += 1;
        const double val = var < 0 ? us(jb, dir, fi[0], fi[1], fi[2]) : fl(jb, var, dir, fi[0], fi[1], fi[2]);
        sum = first ? val : sum + val;
        first = false;
      }
    return sum;
  }

  // -------------------------------------------------------------------------------------------
  // border_conditions/FillOutside.cpp:69-170 FillOutsideCellFunctor::operator()
  // -------------------------------------------------------------------------------------------
  void
  fill_outside(std::vector<double> & A)
  {
#pragma omp parallel for schedule(static)
    for (int io = nowned; io < ntotal; ++io)
    {
      const int co[3] = { bcoord[3 * io], bcoord[3 * io + 1], bcoord[3 * io + 2] };
      int       nrm[3] = { 0, 0, 0 }, ci[3] = { co[0], co[1], co[2] };
      for (int d = 0; d < dim; ++d)
      {
        if (co[d] < 0)
          nrm[d] = -1;
        else if (co[d] >= nbd(d))
          nrm[d] = 1;
        ci[d] = co[d] - nrm[d];
      }
      if (dim == 2)
        ci[2] = 0;
      const int ii = lookup[ext_index(ci[0], ci[1], ci[2])];
      for (int cell = 0; cell < ncell; ++cell)
      {
        const int cout[3] = { cell % b, (cell / b) % b, dim == 3 ? cell / (b * b) : 0 };
        for (int face = 0; face < 2 * dim; ++face)
        {
          const int  dir = face / 2;
          const bool left = (face % 2) == 0;
          const bool at_border = left ? (ci[dir] == 0) : (ci[dir] == nbd(dir) - 1);
          if (at_border and nrm[dir] != 0)
          {
            int cin[3] = { cout[0], cout[1], cout[2] };
            if (P.bc[face] == BC_ZERO_GRADIENT)
            {
              cin[dir] = left ? 0 : b - 1;
              const int ic = cidx(cin[0], cin[1], cin[2]);
              for (int v = 0; v < nvar; ++v)
                u(A, io, v, cell) = u(A, ii, v, ic);
            }
            else if (P.bc[face] == BC_WALL)
            {
              cin[dir] = b - 1 - cout[dir];
              const int ic = cidx(cin[0], cin[1], cin[2]);
              for (int v = 0; v < nvar; ++v)
                u(A, io, v, cell) = u(A, ii, v, ic);
              u(A, io, H.IU + dir, cell) *= -1;
            }
          }
        }
      }
    }
    if (P.inlet != 0)
      override_inlet(A);
  }

  // FillOutsideShockBubble / FillOutsideJet (border_conditions/BCShockBubble.cpp:72-172,
  // BCJet.cpp), applied after the border fill (SolverGodunovFiveEq.cpp:997-1018): outside cells
  // whose centre has x < xmin (for the jet: and lies in the jet, BCJet.h:71-88) take the inflow
  // state.  The cell centre is xmin + (global cell index + 1/2) dx (the reference's coordinate
  // helpers are in the absent core).
  void
  override_inlet(std::vector<double> & A)
  {
#pragma omp parallel for schedule(static)
    for (int io = nowned; io < ntotal; ++io)
    {
      const int co[3] = { bcoord[3 * io], bcoord[3 * io + 1], bcoord[3 * io + 2] };
      for (int cell = 0; cell < ncell; ++cell)
      {
        const int    cout[3] = { cell % b, (cell / b) % b, dim == 3 ? cell / (b * b) : 0 };
        double       xc[3];
        for (int d = 0; d < 3; ++d)
          xc[d] = P.xmin[d] + (static_cast<double>(co[d] * b + cout[d]) + 0.5) * P.dx;
        bool in = xc[0] < P.xmin[0];
        if (in and P.inlet == 2)
        {
          if (dim == 2)
            in = xc[1] > P.jet[0] - P.jet[2] / 2 and xc[1] < P.jet[0] + P.jet[2] / 2;
          else
          {
            const double r2 = (xc[1] - P.jet[0]) * (xc[1] - P.jet[0]) + (xc[2] - P.jet[1]) * (xc[2] - P.jet[1]);
            in = r2 < P.jet[2] * P.jet[2] / 4;
          }
        }
        if (in)
          for (int v = 0; v < nvar; ++v)
            u(A, io, v, cell) = P.inlet_u[v];
      }
    }
  }

  // -------------------------------------------------------------------------------------------
  // scheme/ConvertToPrimitivesVariablesFunctor.cpp:831-951 (same-level path :884-888),
  // fill_inner :492-510, fill_ghost_copy :514-537, set_primitive_vars :252-270
  // -------------------------------------------------------------------------------------------
  void
  convert_to_primitives(std::vector<double> & A)
  {
    const int  e = b + 4;
    const long total = static_cast<long>(nowned) * ncellg;
#pragma omp parallel for schedule(static)
    for (long gi = 0; gi < total; ++gi)
    {
      const int ib = static_cast<int>(gi / ncellg);
      const int c = static_cast<int>(gi % ncellg);
      int       co[3] = { c % e - 2, (c / e) % e - 2, dim == 3 ? c / (e * e) - 2 : 0 };
      int       dir[3] = { 0, 0, 0 }, cin[3] = { co[0], co[1], co[2] };
      for (int d = 0; d < dim; ++d)
      {
        if (co[d] < 0)
          dir[d] = -1;
        else if (co[d] >= b)
          dir[d] = 1;
        cin[d] = co[d] - dir[d] * b;
      }
      State uLoc;
      if (not amr or (dir[0] == 0 and dir[1] == 0 and dir[2] == 0))
      {
        const int src = (dir[0] == 0 and dir[1] == 0 and dir[2] == 0) ? ib : neighbor(ib, dir);
        uLoc = load_cons(A, src, cidx(cin[0], cin[1], cin[2]));
      }
      else
      {
        // fill_ghosts (ConvertToPrimitivesVariablesFunctor.cpp:884-918): same level, finer
        // (restriction) or coarser (prolongation) neighbour, by the level of the cell's leaf
        const int g[3] = { bcoord[3 * ib] * b + co[0], bcoord[3 * ib + 1] * b + co[1],
                           dim == 3 ? bcoord[3 * ib + 2] * b + co[2] : 0 };
        uLoc = cons_level(A, blev[ib], g);
      }
      State       qLoc{};
      compute_primitives(uLoc, qLoc, s, eos, dim, H);
      q(ib, H.ID, co[0], co[1], co[2]) = qLoc[H.ID];
      q(ib, H.IAD0, co[0], co[1], co[2]) = qLoc[H.IAD0];
      q(ib, H.IAD1, co[0], co[1], co[2]) = qLoc[H.IAD1];
      q(ib, H.IA0, co[0], co[1], co[2]) = qLoc[H.IA0];
      q(ib, H.IA1, co[0], co[1], co[2]) = 1.0 - qLoc[H.IA0];
      q(ib, H.IP, co[0], co[1], co[2]) = qLoc[H.IP];
      q(ib, H.IU, co[0], co[1], co[2]) = qLoc[H.IU];
      q(ib, H.IV, co[0], co[1], co[2]) = qLoc[H.IV];
      if (dim == 3)
        q(ib, H.IW, co[0], co[1], co[2]) = qLoc[H.IW];
    }
  }

  // scheme/ComputeLimitedSlopesFunctor.cpp:141-193 compute_limited_slopes
  void
  compute_slopes()
  {
    const int  e = b + 2;
    const long total = static_cast<long>(nowned) * ncells;
#pragma omp parallel for schedule(static)
    for (long gi = 0; gi < total; ++gi)
    {
      const int ib = static_cast<int>(gi / ncells);
      const int c = static_cast<int>(gi % ncells);
      const int i = c % e - 1, j = (c / e) % e - 1, k = dim == 3 ? c / (e * e) - 1 : 0;
      for (int v = 0; v < nvar; ++v)
      {
        sl(Sx, ib, v, i, j, k) =
          slope_unsplit_scalar(q(ib, v, i, j, k), q(ib, v, i + 1, j, k), q(ib, v, i - 1, j, k), s.slope_type);
        sl(Sy, ib, v, i, j, k) =
          slope_unsplit_scalar(q(ib, v, i, j, k), q(ib, v, i, j + 1, k), q(ib, v, i, j - 1, k), s.slope_type);
        if (dim == 3)
          sl(Sz, ib, v, i, j, k) = slope_unsplit_scalar(
            q(ib, v, i, j, k), q(ib, v, i, j, k + 1), q(ib, v, i, j, k - 1), s.slope_type);
      }
    }
  }

  // scheme/ComputeFluxesAndStoreFunctor.h:196-241 get_prim_variables
  State
  get_prim(int ib, int i, int j, int k)
  {
    State st{};
    st[H.IAD0] = q(ib, H.IAD0, i, j, k);
    st[H.IAD1] = q(ib, H.IAD1, i, j, k);
    st[H.IA0] = q(ib, H.IA0, i, j, k);
    st[H.IA1] = q(ib, H.IA1, i, j, k);
    st[H.IP] = q(ib, H.IP, i, j, k);
    st[H.IU] = q(ib, H.IU, i, j, k);
    st[H.IV] = q(ib, H.IV, i, j, k);
    if (dim == 3)
      st[H.IW] = q(ib, H.IW, i, j, k);
    return st;
  }

  // scheme/ComputeFluxesAndStoreFunctor.cpp:123-327 reconstruct_state_{2d,3d} (MUSCL-Hancock)
  State
  reconstruct_muscl(State const & qc, int ib, int is, int js, int ks, double const off[3], double dt_dx)
  {
    const double r0 = qc[H.IAD0], r1 = qc[H.IAD1];
    const double r = r0 + r1;
    const double p = qc[H.IP], uu = qc[H.IU], vv = qc[H.IV];
    const double c = eos.mixture_sound_speed(r, p, qc[H.IA0], 1.0 - qc[H.IA0], r0, r1);
    const double dtdx = dt_dx, dtdy = dt_dx, dtdz = dt_dx;
    State        qr{};
    const bool   hancock = P.integrator == TI_HANCOCK;
    if (dim == 2)
    {
      const double drx0 = sl(Sx, ib, H.IAD0, is, js, ks), drx1 = sl(Sx, ib, H.IAD1, is, js, ks);
      const double dpx = sl(Sx, ib, H.IP, is, js, ks), dux = sl(Sx, ib, H.IU, is, js, ks);
      const double dvx = sl(Sx, ib, H.IV, is, js, ks);
      const double dry0 = sl(Sy, ib, H.IAD0, is, js, ks), dry1 = sl(Sy, ib, H.IAD1, is, js, ks);
      const double dpy = sl(Sy, ib, H.IP, is, js, ks), duy = sl(Sy, ib, H.IU, is, js, ks);
      const double dvy = sl(Sy, ib, H.IV, is, js, ks);
      qr[H.IAD0] = r0 + off[0] * drx0 + off[1] * dry0;
      qr[H.IAD1] = r1 + off[0] * drx1 + off[1] * dry1;
      qr[H.IP] = p + off[0] * dpx + off[1] * dpy;
      qr[H.IU] = uu + off[0] * dux + off[1] * duy;
      qr[H.IV] = vv + off[0] * dvx + off[1] * dvy;
      if (hancock)
      {
        const double sr0 = (-uu * (drx0 + drx1) - dux * r) * dtdx + (-vv * (dry0 + dry1) - dvy * r) * dtdy;
        const double sr0_0 = sr0 * qc[H.IA0];
        const double sr0_1 = sr0 * (1.0 - qc[H.IA0]);
        const double su0 = (-uu * dux - dpx / r) * dtdx + (-vv * duy) * dtdy;
        const double sv0 = (-uu * dvx) * dtdx + (-vv * dvy - dpy / r) * dtdy;
        const double sp0 = (-uu * dpx - dux * r * c * c) * dtdx + (-vv * dpy - dvy * r * c * c) * dtdy;
        qr[H.IAD0] += 0.5 * sr0_0;
        qr[H.IAD1] += 0.5 * sr0_1;
        qr[H.IP] += 0.5 * sp0;
        qr[H.IU] += 0.5 * su0;
        qr[H.IV] += 0.5 * sv0;
      }
    }
    else
    {
      const double ww = qc[H.IW];
      const double drx0 = sl(Sx, ib, H.IAD0, is, js, ks), drx1 = sl(Sx, ib, H.IAD1, is, js, ks);
      const double dpx = sl(Sx, ib, H.IP, is, js, ks), dux = sl(Sx, ib, H.IU, is, js, ks);
      const double dvx = sl(Sx, ib, H.IV, is, js, ks), dwx = sl(Sx, ib, H.IW, is, js, ks);
      const double dry0 = sl(Sy, ib, H.IAD0, is, js, ks), dry1 = sl(Sy, ib, H.IAD1, is, js, ks);
      const double dpy = sl(Sy, ib, H.IP, is, js, ks), duy = sl(Sy, ib, H.IU, is, js, ks);
      const double dvy = sl(Sy, ib, H.IV, is, js, ks), dwy = sl(Sy, ib, H.IW, is, js, ks);
      const double drz0 = sl(Sz, ib, H.IAD0, is, js, ks), drz1 = sl(Sz, ib, H.IAD1, is, js, ks);
      const double dpz = sl(Sz, ib, H.IP, is, js, ks), duz = sl(Sz, ib, H.IU, is, js, ks);
      const double dvz = sl(Sz, ib, H.IV, is, js, ks), dwz = sl(Sz, ib, H.IW, is, js, ks);
      qr[H.IAD0] = r0 + off[0] * drx0 + off[1] * dry0 + off[2] * drz0;
      qr[H.IAD1] = r1 + off[0] * drx1 + off[1] * dry1 + off[2] * drz1;
      qr[H.IP] = p + off[0] * dpx + off[1] * dpy + off[2] * dpz;
      qr[H.IU] = uu + off[0] * dux + off[1] * duy + off[2] * duz;
      qr[H.IV] = vv + off[0] * dvx + off[1] * dvy + off[2] * dvz;
      qr[H.IW] = ww + off[0] * dwx + off[1] * dwy + off[2] * dwz;
      if (hancock)
      {
        const double sr0 = (-uu * (drx0 + drx1) - dux * r) * dtdx + (-vv * (dry0 + dry1) - dvy * r) * dtdy +
                           (-ww * (drz0 + drz1) - dwz * r) * dtdz;
        const double sr0_0 = sr0 * qc[H.IA0];
        const double sr0_1 = sr0 * (1.0 - qc[H.IA0]);
        const double su0 = (-uu * dux - dpx / r) * dtdx + (-vv * duy) * dtdy + (-ww * duz) * dtdz;
        const double sv0 = (-uu * dvx) * dtdx + (-vv * dvy - dpy / r) * dtdy + (-ww * dvz) * dtdz;
        const double sw0 = (-uu * dwx) * dtdx + (-vv * dwy) * dtdy + (-ww * dwz - dpz / r) * dtdz;
        const double sp0 = (-uu * dpx - dux * r * c * c) * dtdx + (-vv * dpy - dvy * r * c * c) * dtdy +
                           (-ww * dpz - dwz * r * c * c) * dtdz;
        qr[H.IAD0] += 0.5 * sr0_0;
        qr[H.IAD1] += 0.5 * sr0_1;
        qr[H.IP] += 0.5 * sp0;
        qr[H.IU] += 0.5 * su0;
        qr[H.IV] += 0.5 * sv0;
        qr[H.IW] += 0.5 * sw0;
      }
    }
    qr[H.IAD0] = std::fmax(s.smallr, qr[H.IAD0]);
    qr[H.IAD1] = std::fmax(s.smallr, qr[H.IAD1]);
    qr[H.IA0] = qc[H.IA0];
    return qr;
  }

  // scheme/ComputeFluxesAndStoreTHINCFunctor.cpp:119-473 reconstruct_state_{2d,3d} (THINC)
  // ioff[] holds the integer offsets {+-1, 0, 0} of the THINC functor.
  State
  reconstruct_thinc(State const & qc, int ib, int is, int js, int ks, int const ioff[3], double dt_dx)
  {
    State        qr{};
    const int    dir = ioff[0] != 0 ? 0 : (ioff[1] != 0 ? 1 : 2);
    int          uv[3] = { 0, 0, 0 };
    uv[dir] = 1;
    const double am = q(ib, H.IA0, is - uv[0], js - uv[1], ks - uv[2]);
    const double a0c = q(ib, H.IA0, is, js, ks);
    const double ap = q(ib, H.IA0, is + uv[0], js + uv[1], ks + uv[2]);
    const double monotonicity = (ap - a0c) * (a0c - am);
    const double r0 = qc[H.IAD0], r1 = qc[H.IAD1];
    const bool   thinc_requested =
      (qc[H.IA0] > P.thinc_epsilon) and (qc[H.IA0] < 1.0 - P.thinc_epsilon) and monotonicity > 0;
    const double drx0 = sl(Sx, ib, H.IAD0, is, js, ks), drx1 = sl(Sx, ib, H.IAD1, is, js, ks);
    const double dry0 = sl(Sy, ib, H.IAD0, is, js, ks), dry1 = sl(Sy, ib, H.IAD1, is, js, ks);
    double       drz0 = 0.0, drz1 = 0.0;
    if (dim == 3)
    {
      drz0 = sl(Sz, ib, H.IAD0, is, js, ks);
      drz1 = sl(Sz, ib, H.IAD1, is, js, ks);
    }
    const double ox = static_cast<double>(ioff[0]), oy = static_cast<double>(ioff[1]),
                 oz = static_cast<double>(ioff[2]);
    if (thinc_requested)
    {
      double st[27], n[3];
      if (dim == 2)
      {
        for (int dj = -1; dj <= 1; ++dj)
          for (int di = -1; di <= 1; ++di)
            st[(di + 1) + 3 * (dj + 1)] = q(ib, H.IA0, is + di, js + dj, ks);
      }
      else
      {
        for (int dk = -1; dk <= 1; ++dk)
          for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di)
              st[(di + 1) + 3 * (dj + 1) + 9 * (dk + 1)] = q(ib, H.IA0, is + di, js + dj, ks + dk);
      }
      youngs_normal(dim, st, n);
      const int bb = ioff[dir] < 0 ? -1 : 1;
      qr[H.IA0] = thinc_face_alpha(am, a0c, ap, std::fabs(n[dir]), P.thinc_beta, bb);
      qr[H.IAD0] = r0 + r0 / qc[H.IA0] * (qr[H.IA0] - qc[H.IA0]);
      qr[H.IAD1] = r1 + r1 / (1.0 - qc[H.IA0]) * (qc[H.IA0] - qr[H.IA0]);
    }
    else
    {
      qr[H.IA0] = qc[H.IA0];
      if (dim == 2)
      {
        qr[H.IAD0] = r0 + 0.5 * ox * drx0 + 0.5 * oy * dry0;
        qr[H.IAD1] = r1 + 0.5 * ox * drx1 + 0.5 * oy * dry1;
      }
      else
      {
        qr[H.IAD0] = r0 + 0.5 * ox * drx0 + 0.5 * oy * dry0 + 0.5 * oz * drz0;
        qr[H.IAD1] = r1 + 0.5 * ox * drx1 + 0.5 * oy * dry1 + 0.5 * oz * drz1;
      }
    }
    const double p = qc[H.IP], uu = qc[H.IU], vv = qc[H.IV];
    const double r = r0 + r1;
    const double c = eos.mixture_sound_speed(r, p, qc[H.IA0], 1.0 - qc[H.IA0], r0, r1);
    const double dtdx = dt_dx, dtdy = dt_dx, dtdz = dt_dx;
    const bool   hancock = P.integrator == TI_HANCOCK;
    if (dim == 2)
    {
      const double dpx = sl(Sx, ib, H.IP, is, js, ks), dux = sl(Sx, ib, H.IU, is, js, ks);
      const double dvx = sl(Sx, ib, H.IV, is, js, ks);
      const double dpy = sl(Sy, ib, H.IP, is, js, ks), duy = sl(Sy, ib, H.IU, is, js, ks);
      const double dvy = sl(Sy, ib, H.IV, is, js, ks);
      qr[H.IP] = p + 0.5 * ox * dpx + 0.5 * oy * dpy;
      qr[H.IU] = uu + 0.5 * ox * dux + 0.5 * oy * duy;
      qr[H.IV] = vv + 0.5 * ox * dvx + 0.5 * oy * dvy;
      if (hancock)
      {
        const double sr0 = (-uu * (drx0 + drx1) - dux * r) * dtdx + (-vv * (dry0 + dry1) - dvy * r) * dtdy;
        const double sr0_0 = sr0 * qc[H.IA0];
        const double sr0_1 = sr0 * (1.0 - qc[H.IA0]);
        const double su0 = (-uu * dux - dpx / r) * dtdx + (-vv * duy) * dtdy;
        const double sv0 = (-uu * dvx) * dtdx + (-vv * dvy - dpy / r) * dtdy;
        const double sp0 = (-uu * dpx - dux * r * c * c) * dtdx + (-vv * dpy - dvy * r * c * c) * dtdy;
        qr[H.IAD0] += 0.5 * sr0_0;
        qr[H.IAD1] += 0.5 * sr0_1;
        qr[H.IP] += 0.5 * sp0;
        qr[H.IU] += 0.5 * su0;
        qr[H.IV] += 0.5 * sv0;
      }
    }
    else
    {
      const double ww = qc[H.IW];
      const double dpx = sl(Sx, ib, H.IP, is, js, ks), dux = sl(Sx, ib, H.IU, is, js, ks);
      const double dvx = sl(Sx, ib, H.IV, is, js, ks), dwx = sl(Sx, ib, H.IW, is, js, ks);
      const double dpy = sl(Sy, ib, H.IP, is, js, ks), duy = sl(Sy, ib, H.IU, is, js, ks);
      const double dvy = sl(Sy, ib, H.IV, is, js, ks), dwy = sl(Sy, ib, H.IW, is, js, ks);
      const double dpz = sl(Sz, ib, H.IP, is, js, ks), duz = sl(Sz, ib, H.IU, is, js, ks);
      const double dvz = sl(Sz, ib, H.IV, is, js, ks), dwz = sl(Sz, ib, H.IW, is, js, ks);
      qr[H.IP] = p + 0.5 * ox * dpx + 0.5 * oy * dpy + 0.5 * oz * dpz;
      qr[H.IU] = uu + 0.5 * ox * dux + 0.5 * oy * duy + 0.5 * oz * duz;
      qr[H.IV] = vv + 0.5 * ox * dvx + 0.5 * oy * dvy + 0.5 * oz * dvz;
      qr[H.IW] = ww + 0.5 * ox * dwx + 0.5 * oy * dwy + 0.5 * oz * dwz;
      if (hancock)
      {
        const double sr0 = (-uu * (drx0 + drx1) - dux * r) * dtdx + (-vv * (dry0 + dry1) - dvy * r) * dtdy +
                           (-ww * (drz0 + drz1) - dwz * r) * dtdz;
        const double sr0_0 = sr0 * qc[H.IA0];
        const double sr0_1 = sr0 * (1.0 - qc[H.IA0]);
        const double su0 = (-uu * dux - dpx / r) * dtdx + (-vv * duy) * dtdy + (-ww * duz) * dtdz;
        const double sv0 = (-uu * dvx) * dtdx + (-vv * dvy - dpy / r) * dtdy + (-ww * dvz) * dtdz;
        const double sw0 = (-uu * dwx) * dtdx + (-vv * dwy) * dtdy + (-ww * dwz - dpz / r) * dtdz;
        const double sp0 = (-uu * dpx - dux * r * c * c) * dtdx + (-vv * dpy - dvy * r * c * c) * dtdy +
                           (-ww * dpz - dwz * r * c * c) * dtdz;
        qr[H.IAD0] += 0.5 * sr0_0;
        qr[H.IAD1] += 0.5 * sr0_1;
        qr[H.IP] += 0.5 * sp0;
        qr[H.IU] += 0.5 * su0;
        qr[H.IV] += 0.5 * sv0;
        qr[H.IW] += 0.5 * sw0;
      }
    }
    return qr;
  }

  // scheme/ComputeFluxesAndStoreFunctor.cpp:331-642 compute_fluxes_and_store_{2d,3d} and the THINC
  // variant scheme/ComputeFluxesAndStoreTHINCFunctor.cpp:477-788; set_flux (.h:253-307).
  void
  compute_fluxes(double dt, int dir)
  {
    const long   total = static_cast<long>(nowned) * ncellf;
    const bool   thinc = P.thinc_enabled != 0;
    const int    ex = dir == 0 ? b + 1 : b, ey = dir == 1 ? b + 1 : b;
#pragma omp parallel for schedule(static)
    for (long gi = 0; gi < total; ++gi)
    {
      const int    ib = static_cast<int>(gi / ncellf);
      const double dtdS = dt / dx_of(ib); // dt / the block's cell length
      const int c = static_cast<int>(gi % ncellf);
      const int i = c % ex, j = (c / ex) % ey, k = dim == 3 ? c / (ex * ey) : 0;
      int       sh[3] = { 0, 0, 0 };
      sh[dir] = 1;
      const State qprim = get_prim(ib, i, j, k);
      const State qprim_n = get_prim(ib, i - sh[0], j - sh[1], k - sh[2]);
      State       qL, qR;
      if (thinc)
      {
        int offL[3] = { 0, 0, 0 }, offR[3] = { 0, 0, 0 };
        offL[dir] = 1;
        offR[dir] = -1;
        qL = reconstruct_thinc(qprim_n, ib, i - sh[0], j - sh[1], k - sh[2], offL, dtdS);
        qR = reconstruct_thinc(qprim, ib, i, j, k, offR, dtdS);
      }
      else
      {
        double offL[3] = { 0.0, 0.0, 0.0 }, offR[3] = { 0.0, 0.0, 0.0 };
        offL[dir] = 0.5;
        offR[dir] = -0.5;
        qL = reconstruct_muscl(qprim_n, ib, i - sh[0], j - sh[1], k - sh[2], offL, dtdS);
        qR = reconstruct_muscl(qprim, ib, i, j, k, offR, dtdS);
      }
      const int swp = dir == 0 ? -1 : (dir == 1 ? H.IV : H.IW);
      if (swp >= 0)
      {
        std::swap(qL[H.IU], qL[swp]);
        std::swap(qR[H.IU], qR[swp]);
      }
      Riemann rs = riemann_hllc(qL, qR, s, eos, dim, H);
      if (swp >= 0)
        std::swap(rs.flux[H.IU], rs.flux[swp]);
      rs.flux[H.IA0] = rs.ustar * rs.phistar0;
      rs.flux[H.IA1] = rs.ustar * rs.phistar1;
      State fc;
      for (int v = 0; v < nvar; ++v)
        fc[v] = rs.flux[v] * dtdS;
      fl(ib, H.IAD0, dir, i, j, k) = fc[H.IAD0];
      fl(ib, H.IAD1, dir, i, j, k) = fc[H.IAD1];
      fl(ib, H.IA0, dir, i, j, k) = fc[H.IA0];
      fl(ib, H.IA1, dir, i, j, k) = fc[H.IA1];
      fl(ib, H.IP, dir, i, j, k) = fc[H.IP];
      fl(ib, H.IU, dir, i, j, k) = fc[H.IU];
      fl(ib, H.IV, dir, i, j, k) = fc[H.IV];
      if (dim == 3)
        fl(ib, H.IW, dir, i, j, k) = fc[H.IW];
      us(ib, dir, i, j, k) = rs.ustar * dtdS;
    }
  }

  // scheme/ReadFluxesAndConservativeUpdateFunctor.cpp:112-592 read_fluxes_and_update_{2d,3d}
  // (same-level branch) and update_U (.h:319-393); source term factors of .h:412-443 from the
  // cell's time-n state (1 for ALLAIRE_KOKH).
  void
  update(std::vector<double> & Uin, std::vector<double> & Uout, int dir)
  {
    const long total = static_cast<long>(nowned) * ncell;
    const double eps = 1e-12;
#pragma omp parallel for schedule(static)
    for (long gi = 0; gi < total; ++gi)
    {
      const int ib = static_cast<int>(gi / ncell);
      const int c = static_cast<int>(gi % ncell);
      double    st0 = 1.0, st1 = 1.0;
      if (P.source_term == 1)
      {
        State uloc{};
        for (int v = 0; v < nvar; ++v)
          uloc[v] = u(Uin, ib, v, c);
        st0 = source_term_factor(1, uloc, s, eos, dim, H, 0);
        st1 = source_term_factor(1, uloc, s, eos, dim, H, 1);
      }
      const int i = c % b, j = (c / b) % b, k = dim == 3 ? c / (b * b) : 0;
      int       sh[3] = { 0, 0, 0 };
      sh[dir] = 1;
      const int vars[8] = { H.IAD0, H.IAD1, H.IA0, H.IA1, H.IP, H.IU, H.IV, H.IW };
      const int nv = dim == 3 ? 8 : 7;
      double    acc[9];
      for (int v = 0; v < 9; ++v)
        acc[v] = 0.0;
      const double ua0 = u(Uin, ib, H.IA0, c), ua1 = u(Uin, ib, H.IA1, c);
      const int    ci[3] = { i, j, k };
      // a coarse cell on a face whose neighbour is finer takes the fine faces' sum over the
      // volume ratio 2^dim (read_fluxes_and_update_{2d,3d} NEIGHBOR_IS_FINER branches,
      // ReadFluxesAndConservativeUpdateFunctor.cpp:137-199)
      const double vr = dim == 3 ? 8.0 : 4.0;
      if (ci[dir] == 0 and finer_neighbor(ib, dir, -1))
      {
        for (int t = 0; t < nv; ++t)
          acc[vars[t]] = acc[vars[t]] + fine_face_sum(ib, dir, -1, i, j, k, vars[t]) / vr;
        const double ustar_flx = fine_face_sum(ib, dir, -1, i, j, k, -1);
        acc[H.IA0] -= st0 * ustar_flx * ua0 / vr;
        acc[H.IA1] -= st1 * ustar_flx * ua1 / vr;
      }
      else
      {
        for (int t = 0; t < nv; ++t)
          acc[vars[t]] = acc[vars[t]] + fl(ib, vars[t], dir, i, j, k);
        acc[H.IA0] -= st0 * us(ib, dir, i, j, k) * ua0;
        acc[H.IA1] -= st1 * us(ib, dir, i, j, k) * ua1;
      }
      if (ci[dir] == b - 1 and finer_neighbor(ib, dir, 1))
      {
        for (int t = 0; t < nv; ++t)
          acc[vars[t]] = acc[vars[t]] - fine_face_sum(ib, dir, 1, i, j, k, vars[t]) / vr;
        const double ustar_flx = fine_face_sum(ib, dir, 1, i, j, k, -1);
        acc[H.IA0] += st0 * ustar_flx * ua0 / vr;
        acc[H.IA1] += st1 * ustar_flx * ua1 / vr;
      }
      else
      {
        for (int t = 0; t < nv; ++t)
          acc[vars[t]] = acc[vars[t]] - fl(ib, vars[t], dir, i + sh[0], j + sh[1], k + sh[2]);
        acc[H.IA0] += st0 * us(ib, dir, i + sh[0], j + sh[1], k + sh[2]) * ua0;
        acc[H.IA1] += st1 * us(ib, dir, i + sh[0], j + sh[1], k + sh[2]) * ua1;
      }
      double & vad0 = u(Uout, ib, H.IAD0, c);
      vad0 += acc[H.IAD0];
      if (vad0 < eps)
        vad0 = 0.0;
      double & vad1 = u(Uout, ib, H.IAD1, c);
      vad1 += acc[H.IAD1];
      if (vad1 < eps)
        vad1 = 0.0;
      double & va0 = u(Uout, ib, H.IA0, c);
      va0 += acc[H.IA0];
      if (va0 < eps)
        va0 = 0.0;
      if (va0 > 1.0 - eps)
        va0 = 1.0;
      double & va1 = u(Uout, ib, H.IA1, c);
      va1 += acc[H.IA1];
      if (va1 < eps)
        va1 = 0.0;
      if (va1 > 1.0 - eps)
        va1 = 1.0;
      u(Uout, ib, H.IP, c) += acc[H.IP];
      u(Uout, ib, H.IU, c) += acc[H.IU];
      u(Uout, ib, H.IV, c) += acc[H.IV];
      if (dim == 3)
        u(Uout, ib, H.IW, c) += acc[H.IW];
    }
  }

  // scheme/FixVolumeFractionsNormalization.cpp:47-72
  void
  fix_volume_fractions(std::vector<double> & A)
  {
    const long total = static_cast<long>(nowned) * ncell;
#pragma omp parallel for schedule(static)
    for (long gi = 0; gi < total; ++gi)
    {
      const int    ib = static_cast<int>(gi / ncell);
      const int    c = static_cast<int>(gi % ncell);
      const double sum_alpha = u(A, ib, H.IA0, c) + u(A, ib, H.IA1, c);
      const double sum_alpha_rho = u(A, ib, H.IAD0, c) + u(A, ib, H.IAD1, c);
      const double inv = 1.0 / sum_alpha;
      u(A, ib, H.IA0, c) *= inv;
      u(A, ib, H.IA1, c) *= inv;
      u(A, ib, H.ID, c) = sum_alpha_rho;
    }
  }

  // scheme/GodunovImplemV0.cpp:83-151 do_time_step(U, U2, dt)
  void
  do_time_step(std::vector<double> & Uin, std::vector<double> & Uout, double dt)
  {
    convert_to_primitives(Uin);
    compute_slopes();
    for (int dir = 0; dir < dim; ++dir)
    {
      compute_fluxes(dt, dir);
      update(Uin, Uout, dir);
    }
    fix_volume_fractions(Uout);
  }

  // scheme/ComputeDtFiveEqFunctor.cpp:85-134 compute_cfl + Max reduce ; returns invDt
  double
  compute_inv_dt(std::vector<double> & A)
  {
    double     invDt = 0.0;
    const long total = static_cast<long>(nowned) * ncell;
#pragma omp parallel for schedule(static) reduction(max : invDt)
    for (long gi = 0; gi < total; ++gi)
    {
      const int   ib = static_cast<int>(gi / ncell);
      const int   c = static_cast<int>(gi % ncell);
      const State uLoc = load_cons(A, ib, c);
      State       qLoc{};
      const double cs = compute_primitives(uLoc, qLoc, s, eos, dim, H);
      const double vx = cs + std::fabs(qLoc[H.IU]);
      const double vy = cs + std::fabs(qLoc[H.IV]);
      double       val;
      const double dx = dx_of(ib);
      if (dim == 2)
        val = vx / dx + vy / dx;
      else
      {
        const double vz = cs + std::fabs(qLoc[H.IW]);
        val = vx / dx + vy / dx + vz / dx;
      }
      invDt = std::fmax(invDt, val);
    }
    return invDt;
  }

  // ---- refine flags (scheme/ComputeRefineFlags.cpp) ------------------------------------------
  // Variable v of the cell one step (side -1/+1 along dir) from cell ci of owned block ib, as
  // getNeighborDataSameLevel (:67-107) reads it: outside a non-periodic border the current cell's
  // own value; a same-level or coarser neighbour's cell value (a fine cell outside the fine
  // blocks reads its coarse parent); a finer neighbour's average of the 2^dim fine cells
  // (compute_siblings_average of the absent core: summed x fastest, divided by 2^dim).
  double
  neighbor_value(int ib, int const ci[3], int dir, int side, int v)
  {
    int g[3] = { 0, 0, 0 };
    for (int d = 0; d < dim; ++d)
      g[d] = bcoord[3 * ib + d] * b + ci[d];
    g[dir] += side;
    if (amr and blev[ib] == 1)
    {
      int c[3] = { 0, 0, 0 }, f[3] = { 0, 0, 0 };
      for (int d = 0; d < dim; ++d)
      {
        c[d] = floor_div(g[d], b);
        f[d] = g[d] - c[d] * b;
      }
      const int jb = lookup1[c[0] + 2 * nbd(0) * (c[1] + 2 * nbd(1) * c[2])];
      if (jb >= 0)
        return u(U, jb, v, cidx(f[0], f[1], f[2]));
      int gc[3] = { 0, 0, 0 };
      for (int d = 0; d < dim; ++d)
        gc[d] = floor_div(g[d], 2);
      return coarse_value(gc, v);
    }
    for (int d = 0; d < dim; ++d)
    {
      const int n = nbd(d) * b;
      if (g[d] < 0 or g[d] >= n)
      {
        if (not P.periodic[d])
          return u(U, ib, v, cidx(ci[0], ci[1], ci[2]));
        g[d] = ((g[d] % n) + n) % n;
      }
    }
    return coarse_value(g, v);
  }

  // variable v of level-0 cell g inside the brick: a coarse leaf's value or the average of the
  // fine cells of a refined block
  double
  coarse_value(int const g[3], int v)
  {
    int c[3] = { 0, 0, 0 }, f[3] = { 0, 0, 0 };
    for (int d = 0; d < dim; ++d)
    {
      c[d] = floor_div(g[d], b);
      f[d] = g[d] - c[d] * b;
    }
    const int ib = lookup[ext_index(c[0], c[1], c[2])];
    if (ib >= 0)
      return u(U, ib, v, cidx(f[0], f[1], f[2]));
    double sum = 0.0;
    bool   first = true;
    for (int kz = 0; kz < (dim == 3 ? 2 : 1); ++kz)
      for (int ky = 0; ky < 2; ++ky)
        for (int kx = 0; kx < 2; ++kx)
        {
          const int fg[3] = { 2 * g[0] + kx, 2 * g[1] + ky, dim == 3 ? 2 * g[2] + kz : 0 };
          int       fc[3] = { 0, 0, 0 }, ff[3] = { 0, 0, 0 };
          for (int d = 0; d < dim; ++d)
          {
            fc[d] = floor_div(fg[d], b);
            ff[d] = fg[d] - fc[d] * b;
          }
          const int    jb = lookup1[fc[0] + 2 * nbd(0) * (fc[1] + 2 * nbd(1) * fc[2])];
          const double val = u(U, jb, v, cidx(ff[0], ff[1], ff[2]));
          sum = first ? val : sum + val;
          first = false;
        }
    return sum / (dim == 3 ? 8.0 : 4.0);
  }

  // compute_lohner_split (:117-135): |L - 2C + R| / (|R - C| + |C - L| + eps (|L| + 2|C| + |R|))
  static double
  lohner_split(double L, double C, double R, double eps)
  {
    return std::fabs(L - 2 * C + R) / (std::fabs(R - C) + std::fabs(C - L) +
                                       eps * (std::fabs(L) + 2 * std::fabs(C) + std::fabs(R)));
  }
  // normalized_gradient (:100-113)
  static double
  normalized_gradient(double v1, double v2)
  {
    double vmax = std::fmax(std::fabs(v1), std::fabs(v2));
    if (vmax < 0.001)
      return 0.0;
    vmax = std::fabs(v1 - v2) / vmax;
    return std::fmax(std::fmin(vmax, 1.0), 0.0);
  }

  // Refine flags of the owned blocks (ComputeRefineFlags::operator(), :253-372) from the
  // indicator on the partial densities IAD0 and IAD1: per cell, update_flag (:236-249) for each
  // indicator value (LOHNER_SPLIT: per variable, per direction; LOHNER_UNSPLIT and SIMPLE_GRADIENT:
  // per variable), then the maximum over the block's cells into flags[ib] (the caller's flags are
  // max-accumulated, as the reference's atomic_max into the reset flag array).  Flag values in the
  // order the maximum needs: 0 initial, 1 coarsen, 2 nothing, 3 refine.  indmax[ib] (optional):
  // the largest indicator value of the block (a NaN value -- 0/0 where a partial density and its
  // neighbours vanish -- takes part in no comparison and no maximum).
  // indicator: 0 LOHNER_SPLIT, 1 LOHNER_UNSPLIT, 2 SIMPLE_GRADIENT; base_level: level of level-0
  void
  refine_flags(int indicator, int level_min, int level_max, int base_level, double refine_th,
               double coarsen_th, double eps, int32_t * flags, double * indmax)
  {
#pragma omp parallel for schedule(static)
    for (int ib = 0; ib < nowned; ++ib)
    {
      const int level = base_level + (amr ? blev[ib] : 0);
      int32_t   flag = 0;
      double    imax = 0.0;
      auto      update = [&](double ind) {
        if (level < level_max and ind > refine_th)
          flag = std::max(flag, 3);
        else if (level > level_min and ind < coarsen_th)
          flag = std::max(flag, 1);
        else
          flag = std::max(flag, 2);
        imax = std::fmax(imax, ind);
      };
      for (int c = 0; c < ncell; ++c)
      {
        const int ci[3] = { c % b, (c / b) % b, dim == 3 ? c / (b * b) : 0 };
        for (int v : { H.IAD0, H.IAD1 })
        {
          const double C = u(U, ib, v, c);
          double       L[3], R[3];
          for (int d = 0; d < dim; ++d)
          {
            L[d] = neighbor_value(ib, ci, d, -1, v);
            R[d] = neighbor_value(ib, ci, d, 1, v);
          }
          if (indicator == 0)
          {
            for (int d = 0; d < dim; ++d)
              update(lohner_split(L[d], C, R[d], eps));
          }
          else if (indicator == 1)
          {
            // compute_lohner_unsplit (:140-191)
            double num = 0.0, den = 0.0;
            for (int d = 0; d < dim; ++d)
            {
              const double s2 = L[d] - 2 * C + R[d];
              const double t = std::fabs(R[d] - C) + std::fabs(C - L[d]) +
                               eps * (std::fabs(L[d]) + 2 * std::fabs(C) + std::fabs(R[d]));
              num = d == 0 ? s2 * s2 : num + s2 * s2;
              den = d == 0 ? t * t : den + t * t;
            }
            update(std::sqrt(num / den));
          }
          else
          {
            // compute_simple_gradient (:196-231), from the identity of a max reduction
            double res = -std::numeric_limits<double>::max();
            for (int d = 0; d < dim; ++d)
            {
              res = std::fmax(res, normalized_gradient(C, L[d]));
              res = std::fmax(res, normalized_gradient(C, R[d]));
            }
            update(res);
          }
        }
      }
      flags[ib] = std::max(flags[ib], flag);
      if (indmax)
        indmax[ib] = imax;
    }
  }

  // SolverGodunovFiveEq.cpp:368-390 compute_dt_local
  double
  compute_dt_local()
  {
    const double invDt = compute_inv_dt(U);
    return s.cfl / invDt;
  }

  // SolverGodunovFiveEq.cpp:481-526 godunov_unsplit (HANCOCK/RK1 and RK2_SSP branches)
  void
  godunov_unsplit(double dt)
  {
    const size_t nown = static_cast<size_t>(nowned) * nvar * ncell;
    if (P.integrator == TI_RK2_SSP)
    {
      fill_outside(U);
      Urk.assign(U.begin(), U.end());
      do_time_step(U, Urk, dt);
      // stage 1 combination is the identity on U_RK
      fill_outside(Urk);
      std::copy(Urk.begin(), Urk.begin() + nown, U2.begin());
      do_time_step(Urk, U2, dt);
      // stage 2: U2 = 1/2 U + 1/2 U2
      for (size_t idx = 0; idx < nown; ++idx)
        U2[idx] = 0.5 * U[idx] + 0.5 * U2[idx];
      std::copy(U2.begin(), U2.begin() + nown, U.begin());
      return;
    }
    fill_outside(U);
    do_time_step(U, U2, dt);
    std::copy(U2.begin(), U2.begin() + nown, U.begin());
  }
};

// ---------------------------------------------------------------------------------------------
// Initial conditions: region states (common.h:68-144 get_region_init_state), the shock-bubble
// (init/InitShockBubble.cpp:50-225) and the triple point (init/InitTriplePoint.cpp:46-133).
// The bubble's volume fraction follows the reference's construction: the plane tangent to the
// sphere at the point nearest the cell centre (get_tangent_to_sphere) and the fraction of the cell
// below it (vof::compute_volume_fraction_of_rect_below_plane).  Both are in the absent core and
// are restated here: outward unit normal n = (x_c - c)/|x_c - c|, plane n.x = n.c + R, and the
// exact volume of a box below a plane by inclusion-exclusion over its corners.
// A region is given as {alpha0, alpha_rho0, alpha_rho1, p, u, v, w}.
// ---------------------------------------------------------------------------------------------
State
region_state(int dim, Eos const & e, double const * r)
{
  Ids          H(dim);
  const double alpha0 = r[0], ar0 = r[1], ar1 = r[2], p = r[3];
  State        q{}, U{};
  q[H.IU] = r[4];
  q[H.IV] = r[5];
  if (dim == 3)
    q[H.IW] = r[6];
  const double rho = ar0 + ar1;
  const double eint_specific = e.mixture_specific_eint(rho, p, alpha0, 1 - alpha0, ar0, ar1);
  const double ekin_specific = dim == 2 ? 0.5 * (q[H.IU] * q[H.IU] + q[H.IV] * q[H.IV])
                                        : 0.5 * (q[H.IU] * q[H.IU] + q[H.IV] * q[H.IV] + q[H.IW] * q[H.IW]);
  U[H.ID] = rho;
  U[H.IA0] = alpha0;
  U[H.IA1] = 1.0 - alpha0;
  U[H.IAD0] = ar0;
  U[H.IAD1] = ar1;
  U[H.IE] = (eint_specific + ekin_specific) * rho;
  U[H.IU] = q[H.IU] * rho;
  U[H.IV] = q[H.IV] * rho;
  if (dim == 3)
    U[H.IW] = q[H.IW] * rho;
  return U;
}

// fraction of the box [lo, lo + dx]^dim where n.x <= alpha: with m_i = |n_i| dx (axes with
// n_i < 0 mirrored) and beta the plane's offset from the box's lowest corner, the volume of
// {xi in [0,1]^k : m.xi <= beta} is sum_v (-1)^|v| max(0, beta - m.v)^k / (k! prod m) over the
// corners v of the k axes with m_i > 0 (axes along which the plane does not vary drop out)
double
box_fraction_below_plane(int dim, double const * n, double alpha, double const * lo, double dx)
{
  double m[3], beta = alpha, mmax = 0.0;
  for (int d = 0; d < dim; ++d)
  {
    m[d] = n[d] * dx;
    beta -= n[d] * lo[d];
    if (m[d] < 0.0)
    {
      beta -= m[d];
      m[d] = -m[d];
    }
    mmax = std::max(mmax, m[d]);
  }
  if (mmax == 0.0)
    return beta >= 0.0 ? 1.0 : 0.0;
  double mk[3], msum = 0.0;
  int    k = 0;
  for (int d = 0; d < dim; ++d)
    if (m[d] > 1e-10 * mmax)
    {
      mk[k++] = m[d];
      msum += m[d];
    }
  if (beta <= 0.0)
    return 0.0;
  if (beta >= msum)
    return 1.0;
  double sum = 0.0;
  for (int v = 0; v < (1 << k); ++v)
  {
    double s = beta;
    int    bits = 0;
    for (int d = 0; d < k; ++d)
      if ((v >> d) & 1)
      {
        s -= mk[d];
        ++bits;
      }
    if (s > 0.0)
      sum += ((bits & 1) ? -1.0 : 1.0) * std::pow(s, k);
  }
  double den = 1.0;
  for (int d = 0; d < k; ++d)
    den *= mk[d] * (d + 1);
  return std::min(1.0, std::max(0.0, sum / den));
}

// volume fraction of the cell [lo, lo + dx]^dim inside the sphere (centre c, radius R) by the
// tangent plane at the point of the sphere nearest the cell centre
double
tangent_sphere_fraction(int dim, double const * lo, double dx, double const * c, double R)
{
  double xc[3], n[3], d2 = 0.0;
  for (int d = 0; d < dim; ++d)
  {
    xc[d] = lo[d] + 0.5 * dx;
    n[d] = xc[d] - c[d];
    d2 += n[d] * n[d];
  }
  if (d2 == 0.0)
    return 1.0;
  const double dist = std::sqrt(d2);
  double       alpha = R;
  for (int d = 0; d < dim; ++d)
  {
    n[d] /= dist;
    alpha += n[d] * c[d];
  }
  return box_fraction_below_plane(dim, n, alpha, lo, dx);
}

// InitShockBubbleDataFunctor::operator() (one bubble): regions 0 (x_c <= x_front), 1 (beyond) and
// the bubble 2, owned blocks in brick order
void
init_shock_bubble(g5o_params const & P, double const * regions, double x_front, double const * center, double radius,
                  double * Uown)
{
  const int dim = P.dim, b = P.bsize;
  Ids       H(dim);
  Eos       e;
  e.init(P.gamma, P.pinf);
  State S[3];
  for (int r = 0; r < 3; ++r)
    S[r] = region_state(dim, e, regions + 7 * r);
  // pure-state specific internal energies (InitShockBubble.cpp:183-207)
  auto eint_of = [&](State const & s) {
    double ekin = s[H.IU] * s[H.IU] + s[H.IV] * s[H.IV];
    if (dim == 3)
      ekin += s[H.IW] * s[H.IW];
    ekin /= (2.0 * s[H.ID]);
    return s[H.IE] - ekin;
  };
  const double eint_pre = eint_of(S[1]), eint_bub = eint_of(S[2]);
  const int    ncell = dim == 2 ? b * b : b * b * b;
  const long   nblk = static_cast<long>(P.nb[0]) * P.nb[1] * (dim == 3 ? P.nb[2] : 1);
  for (long blk = 0; blk < nblk; ++blk)
  {
    const int co[3] = { static_cast<int>(blk % P.nb[0]), static_cast<int>((blk / P.nb[0]) % P.nb[1]),
                        static_cast<int>(blk / (static_cast<long>(P.nb[0]) * P.nb[1])) };
    for (int cell = 0; cell < ncell; ++cell)
    {
      const int cc[3] = { cell % b, (cell / b) % b, dim == 3 ? cell / (b * b) : 0 };
      double    lo[3];
      for (int d = 0; d < 3; ++d)
        lo[d] = P.xmin[d] + static_cast<double>(co[d] * b + cc[d]) * P.dx;
      const double xc = lo[0] + 0.5 * P.dx;
      State        u = xc <= x_front ? S[0] : S[1];
      u[H.ID] = u[H.IAD0] + u[H.IAD1];
      u[H.IA1] = 1.0 - u[H.IA0];
      const double frac = tangent_sphere_fraction(dim, lo, P.dx, center, radius);
      if (frac > 0.0)
      {
        const double phi0 = 1.0 - frac;
        u[H.IAD0] = phi0 * S[1][H.IAD0] + (1.0 - phi0) * S[2][H.IAD0];
        u[H.IAD1] = phi0 * S[1][H.IAD1] + (1.0 - phi0) * S[2][H.IAD1];
        u[H.ID] = u[H.IAD0] + u[H.IAD1];
        u[H.IA0] = phi0 * S[1][H.IA0] + (1.0 - phi0) * S[2][H.IA0];
        u[H.IA1] = 1.0 - u[H.IA0];
        u[H.IU] = phi0 * S[1][H.IU] + (1.0 - phi0) * S[2][H.IU];
        u[H.IV] = phi0 * S[1][H.IV] + (1.0 - phi0) * S[2][H.IV];
        if (dim == 3)
          u[H.IW] = phi0 * S[1][H.IW] + (1.0 - phi0) * S[2][H.IW];
        const double rho_mixed = u[H.IAD0] + u[H.IAD1];
        double       ekin_mixed = u[H.IU] * u[H.IU] + u[H.IV] * u[H.IV];
        if (dim == 3)
          ekin_mixed += u[H.IW] * u[H.IW];
        ekin_mixed /= (2.0 * rho_mixed);
        const double eint_mixed = phi0 * eint_pre + (1.0 - phi0) * eint_bub;
        u[H.IE] = eint_mixed + ekin_mixed;
      }
      for (int v = 0; v < H.nbvar; ++v)
        Uown[(blk * H.nbvar + v) * ncell + cell] = u[v];
    }
  }
}

// InitTriplePointDataFunctor::operator(): region 0 for x_c <= xd, else region 1 where y_c <= yd
// (3-D: sqrt(y_c^2 + z_c^2) <= yd) and region 2 elsewhere
void
init_triple_point(g5o_params const & P, double const * regions, double xd, double yd, double * Uown)
{
  const int dim = P.dim, b = P.bsize;
  Ids       H(dim);
  Eos       e;
  e.init(P.gamma, P.pinf);
  State S[3];
  for (int r = 0; r < 3; ++r)
    S[r] = region_state(dim, e, regions + 7 * r);
  const int  ncell = dim == 2 ? b * b : b * b * b;
  const long nblk = static_cast<long>(P.nb[0]) * P.nb[1] * (dim == 3 ? P.nb[2] : 1);
  for (long blk = 0; blk < nblk; ++blk)
  {
    const int co[3] = { static_cast<int>(blk % P.nb[0]), static_cast<int>((blk / P.nb[0]) % P.nb[1]),
                        static_cast<int>(blk / (static_cast<long>(P.nb[0]) * P.nb[1])) };
    for (int cell = 0; cell < ncell; ++cell)
    {
      const int cc[3] = { cell % b, (cell / b) % b, dim == 3 ? cell / (b * b) : 0 };
      double    xyz[3];
      for (int d = 0; d < 3; ++d)
        xyz[d] = P.xmin[d] + static_cast<double>(co[d] * b + cc[d]) * P.dx + 0.5 * P.dx;
      int r = 0;
      if (!(xyz[0] <= xd))
      {
        bool in1 = xyz[1] <= yd;
        if (dim == 3)
          in1 = std::sqrt(xyz[1] * xyz[1] + xyz[2] * xyz[2]) <= yd;
        r = in1 ? 1 : 2;
      }
      State u = S[r];
      u[H.ID] = u[H.IAD0] + u[H.IAD1];
      for (int v = 0; v < H.nbvar; ++v)
        Uown[(blk * H.nbvar + v) * ncell + cell] = u[v];
    }
  }
}

} // namespace g5o

// =============================================================================================
// C ABI for the tests (ctypes)
// =============================================================================================
using namespace g5o;

extern "C"
{

void *
g5o_create(const g5o_params * p)
{
  return new Solver(*p);
}

// a two-level mesh: cfg's brick at level 0 with the n coarse blocks at coords[3k..3k+2] refined into
// 2^dim blocks of half the cell length; prolongation 0 SIMPLE_COPY, 1 EXTRAPOLATE_LINEAR_MINMOD,
// 2 CONSERVATIVE_INTERPOLATION_ORDER_2, 3 CONSERVATIVE_INTERPOLATION_ORDER_4
void *
g5o_create_amr(const g5o_params * p, const int32_t * coords, int32_t n, int32_t prolongation)
{
  std::vector<int32_t> r(coords, coords + 3 * static_cast<size_t>(n));
  return new Solver(*p, r, prolongation);
}

// level of each owned block (0 coarse, 1 fine); g5o_block_coords gives its coordinates at that level
void
g5o_block_levels(void * h, int32_t * levels)
{
  Solver * S = static_cast<Solver *>(h);
  for (int ib = 0; ib < S->nowned; ++ib)
    levels[ib] = S->blev[ib];
}

// conservative state at global cell g (3 ints) of level l of a two-level mesh, as the ghost fill
// reads it (a leaf cell, a restriction or a prolongation)
void
g5o_cons_level(void * h, int32_t level, const int32_t * g, double * out)
{
  Solver *  S = static_cast<Solver *>(h);
  const int gg[3] = { g[0], g[1], g[2] };
  const State u = S->cons_level(S->U, level, gg);
  for (int v = 0; v < S->nvar; ++v)
    out[v] = u[v];
}

// refine flags of the owned blocks (Solver::refine_flags), flags max-accumulated
void
g5o_refine_flags(void * h, int32_t indicator, int32_t level_min, int32_t level_max, int32_t base_level,
                 double refine_th, double coarsen_th, double eps, int32_t * flags, double * indmax)
{
  static_cast<Solver *>(h)->refine_flags(indicator, level_min, level_max, base_level, refine_th, coarsen_th, eps,
                                         flags, indmax);
}

void
g5o_destroy(void * h)
{
  delete static_cast<Solver *>(h);
}

int
g5o_num_owned(void * h)
{
  return static_cast<Solver *>(h)->nowned;
}

int
g5o_num_outside(void * h)
{
  return static_cast<Solver *>(h)->noutside;
}

// owned-block state in, [block][var][cell]; U2 <- U as the reference does after init
// (SolverGodunovFiveEq.cpp:168)
void
g5o_set_state(void * h, const double * Uown)
{
  Solver *     S = static_cast<Solver *>(h);
  const size_t n = static_cast<size_t>(S->nowned) * S->nvar * S->ncell;
  std::memcpy(S->U.data(), Uown, n * sizeof(double));
  std::memcpy(S->U2.data(), Uown, n * sizeof(double));
}

void
g5o_get_state(void * h, double * Uown)
{
  Solver *     S = static_cast<Solver *>(h);
  const size_t n = static_cast<size_t>(S->nowned) * S->nvar * S->ncell;
  std::memcpy(Uown, S->U.data(), n * sizeof(double));
}

// full U including the outside ring (for FillOutside tests)
void
g5o_get_state_all(void * h, double * Uall)
{
  Solver * S = static_cast<Solver *>(h);
  std::memcpy(Uall, S->U.data(), S->U.size() * sizeof(double));
}

void
g5o_block_coords(void * h, int32_t * coords)
{
  Solver * S = static_cast<Solver *>(h);
  std::memcpy(coords, S->bcoord.data(), S->bcoord.size() * sizeof(int32_t));
}

void
g5o_fill_outside(void * h)
{
  Solver * S = static_cast<Solver *>(h);
  S->fill_outside(S->U);
}

double
g5o_compute_inv_dt(void * h)
{
  Solver * S = static_cast<Solver *>(h);
  return S->compute_inv_dt(S->U);
}

double
g5o_compute_dt(void * h)
{
  return static_cast<Solver *>(h)->compute_dt_local();
}

void
g5o_godunov_unsplit(void * h, double dt)
{
  static_cast<Solver *>(h)->godunov_unsplit(dt);
}

// runs nsteps of next_iteration_impl (dt from compute_dt_local, clamped so that t never exceeds
// tEnd); returns the number of steps done, updates *t.
int
g5o_run(void * h, int nsteps, double * t, double tEnd)
{
  Solver * S = static_cast<Solver *>(h);
  int      n = 0;
  for (; n < nsteps; ++n)
  {
    if (*t >= tEnd)
      break;
    double dt = S->compute_dt_local();
    if (*t + dt > tEnd)
      dt = tEnd - *t;
    S->godunov_unsplit(dt);
    *t += dt;
  }
  return n;
}

void
g5o_set_num_threads(int n)
{
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int
g5o_get_max_threads()
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// ---- unit-level entry points ----
static Eos
make_eos(const double * gamma, const double * pinf)
{
  Eos e;
  e.init(gamma, pinf);
  return e;
}

double
g5o_eos_pressure(const double * gamma, const double * pinf, double rho, double eint, double a0)
{
  return make_eos(gamma, pinf).mixture_pressure(rho, eint, a0, 1.0 - a0, 0, 0);
}

double
g5o_eos_eint(const double * gamma, const double * pinf, double rho, double p, double a0)
{
  return make_eos(gamma, pinf).mixture_specific_eint(rho, p, a0, 1.0 - a0, 0, 0);
}

double
g5o_eos_sound_speed(const double * gamma, const double * pinf, double rho, double p, double a0)
{
  return make_eos(gamma, pinf).mixture_sound_speed(rho, p, a0, 1.0 - a0, 0, 0);
}

double
g5o_compute_primitives(const g5o_params * p, const double * uin, double * qout)
{
  Ids      H(p->dim);
  Eos      e = make_eos(p->gamma, p->pinf);
  Settings s{ p->cfl, p->slope_type, p->smallr, p->smallc, p->smallc * p->smallc };
  State    u{}, q{};
  for (int v = 0; v < H.nbvar; ++v)
    u[v] = uin[v];
  const double c = compute_primitives(u, q, s, e, p->dim, H);
  for (int v = 0; v < H.nbvar; ++v)
    qout[v] = q[v];
  return c;
}

// source term factors {st0, st1} of one cell's conservative state (p->source_term selects the model)
void
g5o_source_factors(const g5o_params * p, const double * uin, double * st2)
{
  Ids      H(p->dim);
  Eos      e = make_eos(p->gamma, p->pinf);
  Settings s{ p->cfl, p->slope_type, p->smallr, p->smallc, p->smallc * p->smallc };
  State    u{};
  for (int v = 0; v < H.nbvar; ++v)
    u[v] = uin[v];
  st2[0] = source_term_factor(p->source_term, u, s, e, p->dim, H, 0);
  st2[1] = source_term_factor(p->source_term, u, s, e, p->dim, H, 1);
}

// flux out: nbvar values (IA0/IA1 = ustar*phistar), plus out3 = {ustar, phistar0, phistar1}
void
g5o_riemann_hllc(const g5o_params * p, const double * ql, const double * qr, double * flux, double * out3)
{
  Ids      H(p->dim);
  Eos      e = make_eos(p->gamma, p->pinf);
  Settings s{ p->cfl, p->slope_type, p->smallr, p->smallc, p->smallc * p->smallc };
  State    a{}, c{};
  for (int v = 0; v < H.nbvar; ++v)
  {
    a[v] = ql[v];
    c[v] = qr[v];
  }
  Riemann rs = riemann_hllc(a, c, s, e, p->dim, H);
  rs.flux[H.IA0] = rs.ustar * rs.phistar0;
  rs.flux[H.IA1] = rs.ustar * rs.phistar1;
  for (int v = 0; v < H.nbvar; ++v)
    flux[v] = rs.flux[v];
  out3[0] = rs.ustar;
  out3[1] = rs.phistar0;
  out3[2] = rs.phistar1;
}

double
g5o_slope(double q, double qp, double qm, double slope_type)
{
  return slope_unsplit_scalar(q, qp, qm, slope_type);
}

double
g5o_thinc_alpha(double am, double ac, double ap, double nd, double thinc_beta, int bb)
{
  return thinc_face_alpha(am, ac, ap, nd, thinc_beta, bb);
}

double
g5o_tanh(double x)
{
  return det_tanh(x);
}

void
g5o_youngs(int dim, const double * a, double * n)
{
  youngs_normal(dim, a, n);
}

// ---- initial conditions (regions: 7 values each, {alpha0, alpha_rho0, alpha_rho1, p, u, v, w}) ----
void
g5o_region_state(const g5o_params * p, const double * region, double * U)
{
  Eos e = make_eos(p->gamma, p->pinf);
  const State u = region_state(p->dim, e, region);
  for (int v = 0; v < Ids(p->dim).nbvar; ++v)
    U[v] = u[v];
}

double
g5o_box_fraction_below_plane(int dim, const double * n, double alpha, const double * lo, double dx)
{
  return box_fraction_below_plane(dim, n, alpha, lo, dx);
}

double
g5o_tangent_sphere_fraction(int dim, const double * lo, double dx, const double * c, double R)
{
  return tangent_sphere_fraction(dim, lo, dx, c, R);
}

void
g5o_init_shock_bubble(const g5o_params * p, const double * regions, double x_front, const double * center,
                      double radius, double * Uown)
{
  init_shock_bubble(*p, regions, x_front, center, radius, Uown);
}

void
g5o_init_triple_point(const g5o_params * p, const double * regions, double xd, double yd, double * Uown)
{
  init_triple_point(*p, regions, xd, yd, Uown);
}

} // extern "C"
