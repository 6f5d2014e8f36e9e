// g5_device.h -- device-side physics of the five-equation step (FP64 and mixed paths).
//
// Each function restates one reference function; the FP64 operation order is the reference's
// (the file is compiled with -ffp-contract=off, and FP64 division and sqrt are IEEE correctly
// rounded on gfx950), so results are reproducible to the bit on the host.
//
// The reconstruction functions are templates on the arithmetic T of the primitive/face path.
// T = double is the FP64 path.  T = float is the mixed path: slopes, Hancock terms, THINC (with a
// hardware-exponential tanh), Youngs normals and the HLLC wave speeds run in FP32 (FP32 division
// and sqrt are the fast approximate forms), while the pressure recovery from the total energy,
// the HLLC energy flux, the conservative update and the dt reduction stay FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "g5_types.h"

#define G5D __device__ __forceinline__

namespace g5
{

// ---------------------------------------------------------------------------------------------
// state containers (named fields instead of indexed arrays so everything stays in registers)
// ---------------------------------------------------------------------------------------------
struct Cons // conservative state, models/FiveEq.h:35-50 order
{
  double d, e, mu, mv, mw, a0, ar0, a1, ar1;
};

// T is the arithmetic of the primitive/face path: double for the FP64 path, float for the mixed
// path (the conservative state and the fluxes added to it stay double in both)
template <typename T>
struct PrimT // the primitive values the flux path reads (get_prim_variables .h:196-241)
{
  T r0, r1, a, p, u, v, w;
};

template <typename T>
struct SlopeT // limited slopes of the reconstructed variables
{
  T r0, r1, p, u, v, w;
};

template <typename T>
struct FaceT // reconstructed face state (IA0 in a)
{
  T r0, r1, a, p, u, v, w;
};

using Prim = PrimT<double>;
using Slope = SlopeT<double>;
using Face = FaceT<double>;

// What a face exchanges between the two cells that share it.  FP64 path: the face flux scaled by
// dt/dx (set_flux, ComputeFluxesAndStoreFunctor.h:253-307).  Mixed path: the upwinded HLLC star
// state in FP32 (partial and total densities, normal velocity, pressure, transverse velocities,
// volume fraction) with its total energy in FP64; each cell assembles the FP64 flux from it
// (assemble_flux), so both neighbours apply bit-identical fluxes and the mass, momentum, energy
// and volume-fraction fluxes are mutually consistent in FP64.
template <typename T>
struct FluxT;
template <>
struct FluxT<double>
{
  double ar0, ar1, a0, a1, e, mu, mv, mw, us;
};
template <>
struct FluxT<float>
{
  float  ro0, ro1, ro, uo, pt, t1, t2, phi0;
  double etot;
};
using Flux = FluxT<double>;

// ---------------------------------------------------------------------------------------------
// mixture stiffened-gas EOS (Allaire et al. 2002 isobaric closure):
//   xi = sum alpha_k/(gamma_k-1), Pi = sum alpha_k gamma_k pinf_k/(gamma_k-1)
//   p = (rho e - Pi)/xi ; e = (p xi + Pi)/rho ; rho c^2 = ((xi+1) p + Pi)/xi
// ---------------------------------------------------------------------------------------------
template <typename T>
struct Par; // solve parameters in the arithmetic T
template <>
struct Par<double>
{
  G5D static double g1(SolveDesc const & S, int k) { return S.g1[k]; }
  G5D static double gp(SolveDesc const & S, int k) { return S.gp[k]; }
  G5D static double smallr(SolveDesc const & S) { return S.smallr; }
  G5D static double smallp(SolveDesc const & S) { return S.smallp; }
  G5D static double smallc(SolveDesc const & S) { return S.smallc; }
  G5D static double beta(SolveDesc const & S) { return S.thinc_beta; }
  G5D static double eps(SolveDesc const & S) { return S.thinc_eps; }
};
template <>
struct Par<float>
{
  G5D static float g1(SolveDesc const & S, int k) { return S.g1f[k]; }
  G5D static float gp(SolveDesc const & S, int k) { return S.gpf[k]; }
  G5D static float smallr(SolveDesc const & S) { return S.smallrf; }
  G5D static float smallp(SolveDesc const & S) { return S.smallpf; }
  G5D static float smallc(SolveDesc const & S) { return S.smallcf; }
  G5D static float beta(SolveDesc const & S) { return S.thinc_betaf; }
  G5D static float eps(SolveDesc const & S) { return S.thinc_epsf; }
};

// ---------------------------------------------------------------------------------------------
// FP64 arithmetic of the reassociated path (G5_PREC_FP64_FAST).  The hardware reciprocal and
// reciprocal square root estimates (v_rcp_f64, v_rsq_f64; relative error below 2^-24) are refined
// by a third-order step (the reciprocal) or Goldschmidt / Newton steps (the roots) to within about
// one ulp -- instead of the IEEE division and square root sequences (scale, estimate, refinement,
// fused correction and fixup), whose dependent chains are about twice as long.  Arguments are
// finite and > 0.
// ---------------------------------------------------------------------------------------------
G5D double
rcp64(double b)
{
  // one third-order step r (1 + e + e^2), e = 1 - b r: the estimate's relative error (below
  // 2^-24) cubed, then the rounding of the last fused operations
  const double r = __builtin_amdgcn_rcp(b);
  const double e = fma(-b, r, 1.0);
  return fma(r, fma(e, e, e), r);
}

// sqrt(x) for x >= 0: g = x y, h = y/2 from y ~ 1/sqrt(x), one Goldschmidt step, then one
// Newton correction of g with the residual x - g^2 (0 for x = 0)
G5D double
sqrt64(double x)
{
  const double y = __builtin_amdgcn_rsq(x);
  double       g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  const double d = fma(-g, g, x);
  g = fma(d, h, g);
  return x == 0.0 ? 0.0 : g;
}

// 1/sqrt(x) for x > 0: one Goldschmidt step from the estimate
G5D double
rsqrt64(double x)
{
  const double y = __builtin_amdgcn_rsq(x);
  const double g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  return 2.0 * fma(h, r, h);
}

// the reciprocal and reciprocal square root of the reassociated forms in the arithmetic of the
// path: FP64 refined as above; FP32 (mixed path) the hardware estimates themselves (about 1 ulp;
// the mixed translation units flush FP32 denormals, so 1/x is the bare v_rcp_f32)
G5D double
rcp_of(double x)
{
  return rcp64(x);
}
G5D float
rcp_of(float x)
{
  return 1.0f / x;
}
G5D double
rsqrt_of(double x)
{
  return rsqrt64(x);
}
G5D float
rsqrt_of(float x)
{
  return __builtin_amdgcn_rsqf(x);
}

template <typename T>
G5D T
eos_xi(SolveDesc const & S, T a0, T a1)
{
  return a0 * Par<T>::g1(S, 0) + a1 * Par<T>::g1(S, 1);
}
template <typename T>
G5D T
eos_pi(SolveDesc const & S, T a0, T a1)
{
  return a0 * Par<T>::gp(S, 0) + a1 * Par<T>::gp(S, 1);
}
G5D double
eos_pressure(SolveDesc const & S, double rho, double eint, double a0, double a1)
{
  return (rho * eint - eos_pi(S, a0, a1)) / eos_xi(S, a0, a1);
}
G5D double
eos_eint(SolveDesc const & S, double rho, double p, double a0, double a1)
{
  return (p * eos_xi(S, a0, a1) + eos_pi(S, a0, a1)) / rho;
}
template <typename T>
G5D T
eos_sound_speed(SolveDesc const & S, T rho, T p, T a0, T a1)
{
  const T x = eos_xi(S, a0, a1);
  const T rc2 = ((x + T(1.0)) * p + eos_pi(S, a0, a1)) / x;
  return sqrt(rc2 / rho);
}

// ---------------------------------------------------------------------------------------------
// models/utils.h:200-250 computePrimitives (+ hydro_sanitizer :36-73, sanitize_alpha :82-87).
// Returns the mixture sound speed; prim.a is the sanitised IA0.
// ---------------------------------------------------------------------------------------------
template <int DIM>
G5D double
cons_to_prim(SolveDesc const & S, Cons const & U, Prim & q)
{
  const double d = fmax(U.d, S.smallr);
  const double ux = U.mu / d;
  const double uy = U.mv / d;
  double       uz = 0.0;
  double       ekin;
  if (DIM == 2)
  {
    ekin = 0.5 * (ux * ux + uy * uy);
  }
  else
  {
    uz = U.mw / d;
    ekin = 0.5 * (ux * ux + uy * uy + uz * uz);
  }
  const double eint = U.e / d - ekin;
  const double p = eos_pressure(S, d, eint, U.a0, 1.0 - U.a0);
  double       a = fmax(0.0, fmin(1.0, U.a0));
  double       r0 = U.ar0, r1 = U.ar1;
  const double eps = 1e-12;
  if (r0 < eps)
    r0 = 0.0;
  if (r1 < eps)
    r1 = 0.0;
  if (a < eps)
    a = 0.0;
  if (a > 1.0 - eps)
    a = 1.0;
  q.r0 = r0;
  q.r1 = r1;
  q.a = a;
  q.p = p;
  q.u = ux;
  q.v = uy;
  q.w = uz;
  return eos_sound_speed(S, d, p, a, 1.0 - a);
}

// Mixed-precision conversion: the pressure is recovered in FP64 from the total energy
// (p xi = rho e_int - Pi cancels Pi ~ 1e9 against p xi ~ 1e4 when pinf ~ 6e8), then every
// primitive is rounded to FP32.  Same sanitising as cons_to_prim; returns the FP32 sound speed.
// The FP64 reciprocal of the density is the refined hardware one (within about an ulp).
template <int DIM>
G5D float
cons_to_prim_mixed(SolveDesc const & S, Cons const & U, PrimT<float> & q)
{
  const double d = fmax(U.d, S.smallr);
  const double inv_d = rcp64(d);
  const double ux = U.mu * inv_d;
  const double uy = U.mv * inv_d;
  double       m2 = U.mu * ux + U.mv * uy;
  double       uz = 0.0;
  if (DIM == 3)
  {
    uz = U.mw * inv_d;
    m2 += U.mw * uz;
  }
  const double pxi = (U.e - 0.5 * m2) - eos_pi(S, U.a0, 1.0 - U.a0);
  const float  p = static_cast<float>(pxi) / static_cast<float>(eos_xi(S, U.a0, 1.0 - U.a0));
  double       a = fmax(0.0, fmin(1.0, U.a0));
  double       r0 = U.ar0, r1 = U.ar1;
  const double eps = 1e-12;
  if (r0 < eps)
    r0 = 0.0;
  if (r1 < eps)
    r1 = 0.0;
  if (a < eps)
    a = 0.0;
  if (a > 1.0 - eps)
    a = 1.0;
  q.r0 = static_cast<float>(r0);
  q.r1 = static_cast<float>(r1);
  q.a = static_cast<float>(a);
  q.p = p;
  q.u = static_cast<float>(ux);
  q.v = static_cast<float>(uy);
  q.w = static_cast<float>(uz);
  return eos_sound_speed(S, static_cast<float>(d), p, q.a, 1.0f - q.a);
}

// Reassociated FP64 conversion (G5_PREC_FP64_FAST): the same quantities as cons_to_prim with one
// reciprocal of the density for the velocities, rho e = E - m.u/2 (equal to d (E/d - |u|^2/2) in
// exact arithmetic), one reciprocal of xi for the pressure and c^2 = ((xi+1) p + Pi)/(xi d) with
// one reciprocal and the refined square root.  Same sanitising; returns the sound speed.
template <int DIM>
G5D double
cons_to_prim_fast(SolveDesc const & S, Cons const & U, Prim & q)
{
  const double d = fmax(U.d, S.smallr);
  const double inv_d = rcp64(d);
  const double ux = U.mu * inv_d;
  const double uy = U.mv * inv_d;
  double       m2 = U.mu * ux + U.mv * uy;
  double       uz = 0.0;
  if (DIM == 3)
  {
    uz = U.mw * inv_d;
    m2 += U.mw * uz;
  }
  const double pxi = (U.e - 0.5 * m2) - eos_pi(S, U.a0, 1.0 - U.a0);
  const double p = pxi * rcp64(eos_xi(S, U.a0, 1.0 - U.a0));
  double       a = fmax(0.0, fmin(1.0, U.a0));
  double       r0 = U.ar0, r1 = U.ar1;
  const double eps = 1e-12;
  if (r0 < eps)
    r0 = 0.0;
  if (r1 < eps)
    r1 = 0.0;
  if (a < eps)
    a = 0.0;
  if (a > 1.0 - eps)
    a = 1.0;
  q.r0 = r0;
  q.r1 = r1;
  q.a = a;
  q.p = p;
  q.u = ux;
  q.v = uy;
  q.w = uz;
  const double x = eos_xi(S, a, 1.0 - a);
  return sqrt64(((x + 1.0) * p + eos_pi(S, a, 1.0 - a)) * rcp64(x * d));
}

// one name for the three paths (FAST: the reassociated FP64 path; ignored by the mixed one)
template <int DIM, bool FAST = false>
G5D double
to_prim(SolveDesc const & S, Cons const & U, PrimT<double> & q)
{
  if constexpr (FAST)
    return cons_to_prim_fast<DIM>(S, U, q);
  else
    return cons_to_prim<DIM>(S, U, q);
}
template <int DIM, bool FAST = false>
G5D float
to_prim(SolveDesc const & S, Cons const & U, PrimT<float> & q)
{
  return cons_to_prim_mixed<DIM>(S, U, q);
}

// scheme/ComputeLimitedSlopesFunctor.cpp:114-137 slope_unsplit_scalar, for slope_type 1 or 2 (the
// caller tests it; other values give 0)
template <typename T>
G5D T
slope1_limited(T q, T qp, T qm, T slope_type)
{
  const T dlft = slope_type * (q - qm);
  const T drgt = slope_type * (qp - q);
  const T dcen = T(0.5) * (qp - qm);
  const T dsgn = (dcen >= T(0.0)) ? T(1.0) : T(-1.0);
  const T slop = fmin(fabs(dlft), fabs(drgt));
  const T dlim = (dlft * drgt) <= T(0.0) ? T(0.0) : slop;
  return dsgn * fmin(dlim, fabs(dcen));
}

// the median of a, b and 0: minmod(a, b) -- 0 when the signs differ, else the one nearer 0
G5D double
med0(double a, double b)
{
  return fmax(fmin(a, b), fmin(fmax(a, b), 0.0));
}
G5D float
med0(float a, float b)
{
  return __builtin_amdgcn_fmed3f(a, b, 0.0f);
}

// Reassociated forms (FP64 and mixed): the limited slope as medians with 0.  minmod (slope_type 1):
// the median of dl, dr and 0, which is sign(dc) min(|dl|, |dr|) when dl dr > 0 and 0 otherwise,
// without the bound by |dc| = |dl + dr|/2 >= min(|dl|, |dr|) that holds in exact arithmetic (and
// without the product dl dr, so a product that underflows no longer zeroes the slope).  MC
// (slope_type 2): the median of that minmod of 2 dl and 2 dr, dc and 0 -- the reference's
// sign(dc) min(dlim, |dc|).
template <typename T>
G5D T
minmod_fast(T q, T qp, T qm)
{
  return med0(q - qm, qp - q);
}
template <typename T>
G5D T
mc_fast(T q, T qp, T qm)
{
  return med0(med0(T(2.0) * (q - qm), T(2.0) * (qp - q)), T(0.5) * (qp - qm));
}

// the limited slopes of the six reconstructed variables along one axis, for a known limiter
template <typename T>
G5D SlopeT<T>
slopes_minmod_fast(PrimT<T> const & c, PrimT<T> const & p, PrimT<T> const & m)
{
  return SlopeT<T>{ minmod_fast(c.r0, p.r0, m.r0), minmod_fast(c.r1, p.r1, m.r1), minmod_fast(c.p, p.p, m.p),
                    minmod_fast(c.u, p.u, m.u),    minmod_fast(c.v, p.v, m.v),    minmod_fast(c.w, p.w, m.w) };
}
template <typename T>
G5D SlopeT<T>
slopes_mc_fast(PrimT<T> const & c, PrimT<T> const & p, PrimT<T> const & m)
{
  return SlopeT<T>{ mc_fast(c.r0, p.r0, m.r0), mc_fast(c.r1, p.r1, m.r1), mc_fast(c.p, p.p, m.p),
                    mc_fast(c.u, p.u, m.u),    mc_fast(c.v, p.v, m.v),    mc_fast(c.w, p.w, m.w) };
}
template <typename T>
G5D SlopeT<T>
slopes_limited(PrimT<T> const & c, PrimT<T> const & p, PrimT<T> const & m, T st)
{
  return SlopeT<T>{ slope1_limited(c.r0, p.r0, m.r0, st), slope1_limited(c.r1, p.r1, m.r1, st),
                    slope1_limited(c.p, p.p, m.p, st),    slope1_limited(c.u, p.u, m.u, st),
                    slope1_limited(c.v, p.v, m.v, st),    slope1_limited(c.w, p.w, m.w, st) };
}

// the limited slopes of a cell along NAX axes (neighbours p[a] and m[a] on axis a), with one test
// of the limiter for all of them: the reference's slope_unsplit_scalar per variable (slope_type 1
// or 2, otherwise 0), or with FAST the median forms of minmod_fast / mc_fast
template <typename T, bool FAST, int NAX>
G5D void
slopes_axes(PrimT<T> const & c, PrimT<T> const (&p)[NAX], PrimT<T> const (&m)[NAX], T st, SlopeT<T> (&s)[NAX])
{
  if (st == T(1.0))
  {
#pragma unroll
    for (int a = 0; a < NAX; ++a)
      s[a] = FAST ? slopes_minmod_fast(c, p[a], m[a]) : slopes_limited(c, p[a], m[a], T(1.0));
  }
  else if (st == T(2.0))
  {
#pragma unroll
    for (int a = 0; a < NAX; ++a)
      s[a] = FAST ? slopes_mc_fast(c, p[a], m[a]) : slopes_limited(c, p[a], m[a], T(2.0));
  }
  else
  {
#pragma unroll
    for (int a = 0; a < NAX; ++a)
      s[a] = SlopeT<T>{ T(0.0), T(0.0), T(0.0), T(0.0), T(0.0), T(0.0) };
  }
}

// ---------------------------------------------------------------------------------------------
// MUSCL-Hancock half step: the cell-centred part of reconstruct_state_{2d,3d}
// (scheme/ComputeFluxesAndStoreFunctor.cpp:134-202, :230-316).  The sound speed and the Hancock
// source terms depend on the cell only, so they are computed once per cell and reused by every
// face of the cell (the reference recomputes the same values per face).
// ---------------------------------------------------------------------------------------------
template <typename T>
struct HancockT // 0.5 * (sr0_0, sr0_1, sp0, su0, sv0, sw0)
{
  T r0, r1, p, u, v, w;
};
using Hancock = HancockT<double>;

// FAST (the reassociated forms, FP64 and mixed): the pressure-gradient terms multiply by one
// reciprocal of r, and r c^2 is the mixture's rho c^2 = ((xi+1) p + Pi)/xi itself (no square root).
template <int DIM, typename T, bool FAST = false>
G5D HancockT<T>
hancock_terms(SolveDesc const & S, PrimT<T> const & q, SlopeT<T> const & sx, SlopeT<T> const & sy,
              SlopeT<T> const & sz, T dtdx)
{
  HancockT<T> h;
  const T     r = q.r0 + q.r1;
  const T     dtdy = dtdx, dtdz = dtdx;
  const T     uu = q.u, vv = q.v, ww = q.w;
  if constexpr (FAST)
  {
    const T x = eos_xi(S, q.a, T(1.0) - q.a);
    const T rc2 = ((x + T(1.0)) * q.p + eos_pi(S, q.a, T(1.0) - q.a)) * rcp_of(x);
    const T ri = rcp_of(r);
    T       sr0 = (-uu * (sx.r0 + sx.r1) - sx.u * r) * dtdx + (-vv * (sy.r0 + sy.r1) - sy.v * r) * dtdy;
    T       su0 = (-uu * sx.u - sx.p * ri) * dtdx + (-vv * sy.u) * dtdy;
    T       sv0 = (-uu * sx.v) * dtdx + (-vv * sy.v - sy.p * ri) * dtdy;
    T       sp0 = (-uu * sx.p - sx.u * rc2) * dtdx + (-vv * sy.p - sy.v * rc2) * dtdy;
    T       sw0 = T(0.0);
    if (DIM == 3)
    {
      sr0 += (-ww * (sz.r0 + sz.r1) - sz.w * r) * dtdz;
      su0 += (-ww * sz.u) * dtdz;
      sv0 += (-ww * sz.v) * dtdz;
      sw0 = (-uu * sx.w) * dtdx + (-vv * sy.w) * dtdy + (-ww * sz.w - sz.p * ri) * dtdz;
      sp0 += (-ww * sz.p - sz.w * rc2) * dtdz;
    }
    h.r0 = T(0.5) * (sr0 * q.a);
    h.r1 = T(0.5) * (sr0 * (T(1.0) - q.a));
    h.p = T(0.5) * sp0;
    h.u = T(0.5) * su0;
    h.v = T(0.5) * sv0;
    h.w = T(0.5) * sw0;
    return h;
  }
  const T c = eos_sound_speed(S, r, q.p, q.a, T(1.0) - q.a);
  if (DIM == 2)
  {
    const T sr0 = (-uu * (sx.r0 + sx.r1) - sx.u * r) * dtdx + (-vv * (sy.r0 + sy.r1) - sy.v * r) * dtdy;
    const T sr0_0 = sr0 * q.a;
    const T sr0_1 = sr0 * (T(1.0) - q.a);
    const T su0 = (-uu * sx.u - sx.p / r) * dtdx + (-vv * sy.u) * dtdy;
    const T sv0 = (-uu * sx.v) * dtdx + (-vv * sy.v - sy.p / r) * dtdy;
    const T sp0 = (-uu * sx.p - sx.u * r * c * c) * dtdx + (-vv * sy.p - sy.v * r * c * c) * dtdy;
    h.r0 = T(0.5) * sr0_0;
    h.r1 = T(0.5) * sr0_1;
    h.p = T(0.5) * sp0;
    h.u = T(0.5) * su0;
    h.v = T(0.5) * sv0;
    h.w = T(0.0);
  }
  else
  {
    const T sr0 = (-uu * (sx.r0 + sx.r1) - sx.u * r) * dtdx + (-vv * (sy.r0 + sy.r1) - sy.v * r) * dtdy +
                  (-ww * (sz.r0 + sz.r1) - sz.w * r) * dtdz;
    const T sr0_0 = sr0 * q.a;
    const T sr0_1 = sr0 * (T(1.0) - q.a);
    const T su0 = (-uu * sx.u - sx.p / r) * dtdx + (-vv * sy.u) * dtdy + (-ww * sz.u) * dtdz;
    const T sv0 = (-uu * sx.v) * dtdx + (-vv * sy.v - sy.p / r) * dtdy + (-ww * sz.v) * dtdz;
    const T sw0 = (-uu * sx.w) * dtdx + (-vv * sy.w) * dtdy + (-ww * sz.w - sz.p / r) * dtdz;
    const T sp0 = (-uu * sx.p - sx.u * r * c * c) * dtdx + (-vv * sy.p - sy.v * r * c * c) * dtdy +
                  (-ww * sz.p - sz.w * r * c * c) * dtdz;
    h.r0 = T(0.5) * sr0_0;
    h.r1 = T(0.5) * sr0_1;
    h.p = T(0.5) * sp0;
    h.u = T(0.5) * su0;
    h.v = T(0.5) * sv0;
    h.w = T(0.5) * sw0;
  }
  return h;
}

// q + off_x*sx + off_y*sy (+ off_z*sz), one offset +-0.5 on axis AX, the others 0.0 (kept so that
// the operation sequence is the reference's).  AX = -1: the axis is the run-time value ax (one
// code path for lanes that reconstruct faces on different axes; the same operations per lane).
// FAST (the reassociated forms, FP64 and mixed): q + o * s of the face axis alone.
template <int DIM, int AX, typename T, bool FAST = false>
G5D T
muscl_extrap(T qv, T sxv, T syv, T szv, T o, int ax = AX)
{
  const int a = AX >= 0 ? AX : ax;
  if constexpr (FAST)
    return qv + o * (a == 0 ? sxv : (a == 1 ? syv : szv));
  const T   ox = a == 0 ? o : T(0.0);
  const T   oy = a == 1 ? o : T(0.0);
  const T   oz = a == 2 ? o : T(0.0);
  if (DIM == 2)
    return qv + ox * sxv + oy * syv;
  return qv + ox * sxv + oy * syv + oz * szv;
}

// Reassociated forms: (q + h) + o s of the face axis -- the Hancock-advanced cell value q + h is
// the same expression for every face of the cell, so the compiler computes it once per cell
template <bool HANCOCK, int AX, typename T>
G5D T
hancock_extrap(T qv, T hv, T sxv, T syv, T szv, T o, int ax)
{
  const int a = AX >= 0 ? AX : ax;
  return (HANCOCK ? qv + hv : qv) + o * (a == 0 ? sxv : (a == 1 ? syv : szv));
}

// reconstruct_state_{2d,3d} face part (scheme/ComputeFluxesAndStoreFunctor.cpp:164-211)
template <int DIM, int AX, bool HANCOCK, typename T, bool FAST = false>
G5D FaceT<T>
face_muscl(SolveDesc const & S, PrimT<T> const & q, SlopeT<T> const & sx, SlopeT<T> const & sy,
           SlopeT<T> const & sz, HancockT<T> const & h, T o, int ax = AX)
{
  FaceT<T> f;
  if constexpr (FAST)
  {
    f.r0 = hancock_extrap<HANCOCK, AX>(q.r0, h.r0, sx.r0, sy.r0, sz.r0, o, ax);
    f.r1 = hancock_extrap<HANCOCK, AX>(q.r1, h.r1, sx.r1, sy.r1, sz.r1, o, ax);
    f.p = hancock_extrap<HANCOCK, AX>(q.p, h.p, sx.p, sy.p, sz.p, o, ax);
    f.u = hancock_extrap<HANCOCK, AX>(q.u, h.u, sx.u, sy.u, sz.u, o, ax);
    f.v = hancock_extrap<HANCOCK, AX>(q.v, h.v, sx.v, sy.v, sz.v, o, ax);
    f.w = DIM == 3 ? hancock_extrap<HANCOCK, AX>(q.w, h.w, sx.w, sy.w, sz.w, o, ax) : T(0.0);
    f.r0 = fmax(Par<T>::smallr(S), f.r0);
    f.r1 = fmax(Par<T>::smallr(S), f.r1);
    f.a = q.a;
    return f;
  }
  f.r0 = muscl_extrap<DIM, AX, T, FAST>(q.r0, sx.r0, sy.r0, sz.r0, o, ax);
  f.r1 = muscl_extrap<DIM, AX, T, FAST>(q.r1, sx.r1, sy.r1, sz.r1, o, ax);
  f.p = muscl_extrap<DIM, AX, T, FAST>(q.p, sx.p, sy.p, sz.p, o, ax);
  f.u = muscl_extrap<DIM, AX, T, FAST>(q.u, sx.u, sy.u, sz.u, o, ax);
  f.v = muscl_extrap<DIM, AX, T, FAST>(q.v, sx.v, sy.v, sz.v, o, ax);
  f.w = DIM == 3 ? muscl_extrap<DIM, AX, T, FAST>(q.w, sx.w, sy.w, sz.w, o, ax) : T(0.0);
  if (HANCOCK)
  {
    f.r0 += h.r0;
    f.r1 += h.r1;
    f.p += h.p;
    f.u += h.u;
    f.v += h.v;
    if (DIM == 3)
      f.w += h.w;
  }
  f.r0 = fmax(Par<T>::smallr(S), f.r0);
  f.r1 = fmax(Par<T>::smallr(S), f.r1);
  f.a = q.a;
  return f;
}

// ---------------------------------------------------------------------------------------------
// Deterministic tanh (same algorithm as the host oracle): tanh(a) = -m/(2+m), m = expm1(-2a),
// expm1 by Cody-Waite reduction and the degree-13 Taylor polynomial.
// ---------------------------------------------------------------------------------------------
G5D double
det_expm1_neg(double y)
{
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double inv_ln2 = 1.44269504088896338700e+00;
  double       k = rint(y * inv_ln2);
  if (k < -60.0)
    k = -60.0;
  const double r = (y - k * ln2_hi) - k * ln2_lo;
  double       p = 1.6059043836821613e-10;
  p = p * r + 2.08767569878681e-09;
  p = p * r + 2.505210838544172e-08;
  p = p * r + 2.755731922398589e-07;
  p = p * r + 2.7557319223985893e-06;
  p = p * r + 2.48015873015873e-05;
  p = p * r + 1.98412698412698413e-04;
  p = p * r + 1.38888888888888894e-03;
  p = p * r + 8.33333333333333322e-03;
  p = p * r + 4.16666666666666644e-02;
  p = p * r + 1.66666666666666657e-01;
  p = p * r + 5.00000000000000000e-01;
  p = p * r + 1.0;
  p = p * r;
  if (k == 0.0)
    return p;
  const double two_k = ldexp(1.0, static_cast<int>(k));
  return two_k * p + (two_k - 1.0);
}

G5D double
det_tanh(double x)
{
  const double a = fabs(x);
  if (a > 22.0)
    return x > 0 ? 1.0 : -1.0;
  const double m = det_expm1_neg(-2.0 * a);
  const double t = -m / (2.0 + m);
  return x < 0 ? -t : t;
}

// FP32 tanh of the mixed path: 1 - 2/(exp(2|x|)+1) with the hardware exponential
G5D float
fast_tanh(float x)
{
  const float e = __expf(-2.0f * fabsf(x));
  const float t = (1.0f - e) / (1.0f + e);
  return x < 0.0f ? -t : t;
}
// reassociated FP64 path: the same expm1 with the final quotient as a product by rcp64
G5D double
det_tanh_fast(double x)
{
  const double a = fabs(x);
  if (a > 22.0)
    return x > 0 ? 1.0 : -1.0;
  const double m = det_expm1_neg(-2.0 * a);
  const double t = -m * rcp64(2.0 + m);
  return x < 0 ? -t : t;
}
template <bool FAST = false>
G5D double
tanh_of(double x)
{
  if constexpr (FAST)
    return det_tanh_fast(x);
  else
    return det_tanh(x);
}
template <bool FAST = false>
G5D float
tanh_of(float x)
{
  return fast_tanh(x);
}

// ---------------------------------------------------------------------------------------------
// Youngs (1982) normal from the 3^d alpha stencil (same weights and order as the host oracle).
// ---------------------------------------------------------------------------------------------
// FAST (the reassociated forms, FP64 and mixed): the components multiply by one reciprocal square
// root.
template <typename T, bool FAST = false>
G5D void
youngs2(T const (&a)[9], T & nx, T & ny)
{
  // a[(di+1) + 3*(dj+1)]
  const T gx = ((a[2] + T(2.0) * a[5]) + a[8]) - ((a[0] + T(2.0) * a[3]) + a[6]);
  const T gy = ((a[6] + T(2.0) * a[7]) + a[8]) - ((a[0] + T(2.0) * a[1]) + a[2]);
  if constexpr (FAST)
  {
    const T n2 = gx * gx + gy * gy;
    const T rn = n2 > T(0.0) ? rsqrt_of(n2) : T(0.0);
    nx = gx * rn;
    ny = gy * rn;
    return;
  }
  const T nrm = sqrt(gx * gx + gy * gy);
  if (nrm > T(0.0))
  {
    nx = gx / nrm;
    ny = gy / nrm;
  }
  else
  {
    nx = T(0.0);
    ny = T(0.0);
  }
}

// 3-D: a[(di+1) + 3*(dj+1) + 9*(dk+1)]; plane sums in the order t2 outer, t1 inner with
// o[d]=s, o[(d+1)%3]=t1, o[(d+2)%3]=t2 and weights (1,2,1)x(1,2,1)
template <int D, typename T>
G5D T
youngs_plane(T const (&a)[27], int s)
{
  T acc = T(0.0);
#pragma unroll
  for (int t2 = -1; t2 <= 1; ++t2)
#pragma unroll
    for (int t1 = -1; t1 <= 1; ++t1)
    {
      const T w = (t1 == 0 ? T(2.0) : T(1.0)) * (t2 == 0 ? T(2.0) : T(1.0));
      int          o[3];
      o[D] = s;
      o[(D + 1) % 3] = t1;
      o[(D + 2) % 3] = t2;
      acc = acc + w * a[(o[0] + 1) + 3 * (o[1] + 1) + 9 * (o[2] + 1)];
    }
  return acc;
}

template <typename T, bool FAST = false>
G5D void
youngs3(T const (&a)[27], T & nx, T & ny, T & nz)
{
  const T gx = youngs_plane<0>(a, 1) - youngs_plane<0>(a, -1);
  const T gy = youngs_plane<1>(a, 1) - youngs_plane<1>(a, -1);
  const T gz = youngs_plane<2>(a, 1) - youngs_plane<2>(a, -1);
  if constexpr (FAST)
  {
    const T n2 = gx * gx + gy * gy + gz * gz;
    const T rn = n2 > T(0.0) ? rsqrt_of(n2) : T(0.0);
    nx = gx * rn;
    ny = gy * rn;
    nz = gz * rn;
    return;
  }
  const T nrm = sqrt(gx * gx + gy * gy + gz * gz);
  if (nrm > T(0.0))
  {
    nx = gx / nrm;
    ny = gy / nrm;
    nz = gz / nrm;
  }
  else
  {
    nx = ny = nz = T(0.0);
  }
}

// THINC alpha and partial densities of one face side (ComputeFluxesAndStoreTHINCFunctor.cpp
// :140-222 / :310-395).  am, ac, ap: alpha at -1, 0, +1 along the face axis; nd: |normal[axis]|
// (only used when THINC is requested); ib: +1 for the right face of the cell, -1 for the left;
// o: the MUSCL offset (0.5*ib) on the face axis (MUSCL fallback).
template <typename T>
struct ThincAxisT // per-cell, per-axis THINC quantities shared by the two faces
{
  bool req;
  T    A, B, C, D;
};
using ThincAxis = ThincAxisT<double>;

// FAST (the reassociated forms, FP64 and mixed): (phi - A)/B as a product by the reciprocal of B,
// the FAST tanh
template <typename T, bool FAST = false>
G5D ThincAxisT<T>
thinc_axis(SolveDesc const & S, T qa, T am, T ac, T ap, T nd)
{
  ThincAxisT<T> t;
  const T       monotonicity = (ap - ac) * (ac - am);
  t.req = (qa > Par<T>::eps(S)) and (qa < T(1.0) - Par<T>::eps(S)) and monotonicity > 0;
  t.A = t.B = t.C = t.D = T(0.0);
  if (t.req)
  {
    const T beta = Par<T>::beta(S) * nd + T(0.01);
    t.A = (ap + am) / 2;
    t.B = (ap - am) / 2;
    t.C = tanh_of<FAST>(T(0.5) * beta);
    const T phi = ac;
    if constexpr (FAST)
      t.D = tanh_of<FAST>(T(0.5) * beta * (phi - t.A) * rcp_of(t.B));
    else
      t.D = tanh_of(T(0.5) * beta * (phi - t.A) / t.B);
  }
  return t;
}

// reconstruct_state_{2d,3d} of the THINC functor: alpha-related part and MUSCL part for p, u, v, w.
// FAST (the reassociated forms, FP64 and mixed): the quotients by C, 1 + b D, alpha and 1 - alpha as
// products by reciprocals (those of C, alpha and 1 - alpha are shared by the faces of the cell).
template <int DIM, int AX, bool HANCOCK, typename T, bool FAST = false>
G5D FaceT<T>
face_thinc(PrimT<T> const & q, SlopeT<T> const & sx, SlopeT<T> const & sy, SlopeT<T> const & sz,
           HancockT<T> const & h, ThincAxisT<T> const & th, int ib, int ax = AX)
{
  FaceT<T> f;
  const T  o = T(0.5) * static_cast<T>(ib); // HALF_F * offsets[AX]
  if constexpr (FAST)
  {
    // (q + h) shared by the faces of the cell, as in face_muscl
    if (th.req)
    {
      const T bb = static_cast<T>(ib);
      f.a = th.A + bb * th.B * ((th.C + bb * th.D * rcp_of(th.C)) * rcp_of(T(1.0) + bb * th.D));
      f.r0 = (HANCOCK ? q.r0 + h.r0 : q.r0) + q.r0 * rcp_of(q.a) * (f.a - q.a);
      f.r1 = (HANCOCK ? q.r1 + h.r1 : q.r1) + q.r1 * rcp_of(T(1.0) - q.a) * (q.a - f.a);
    }
    else
    {
      f.a = q.a;
      f.r0 = hancock_extrap<HANCOCK, AX>(q.r0, h.r0, sx.r0, sy.r0, sz.r0, o, ax);
      f.r1 = hancock_extrap<HANCOCK, AX>(q.r1, h.r1, sx.r1, sy.r1, sz.r1, o, ax);
    }
    f.p = hancock_extrap<HANCOCK, AX>(q.p, h.p, sx.p, sy.p, sz.p, o, ax);
    f.u = hancock_extrap<HANCOCK, AX>(q.u, h.u, sx.u, sy.u, sz.u, o, ax);
    f.v = hancock_extrap<HANCOCK, AX>(q.v, h.v, sx.v, sy.v, sz.v, o, ax);
    f.w = DIM == 3 ? hancock_extrap<HANCOCK, AX>(q.w, h.w, sx.w, sy.w, sz.w, o, ax) : T(0.0);
    return f;
  }
  if (th.req)
  {
    const T bb = static_cast<T>(ib);
    f.a = th.A + bb * th.B * (th.C + bb * th.D / th.C) / (T(1.0) + bb * th.D);
    f.r0 = q.r0 + q.r0 / q.a * (f.a - q.a);
    f.r1 = q.r1 + q.r1 / (T(1.0) - q.a) * (q.a - f.a);
  }
  else
  {
    f.a = q.a;
    f.r0 = muscl_extrap<DIM, AX, T, FAST>(q.r0, sx.r0, sy.r0, sz.r0, o, ax);
    f.r1 = muscl_extrap<DIM, AX, T, FAST>(q.r1, sx.r1, sy.r1, sz.r1, o, ax);
  }
  f.p = muscl_extrap<DIM, AX, T, FAST>(q.p, sx.p, sy.p, sz.p, o, ax);
  f.u = muscl_extrap<DIM, AX, T, FAST>(q.u, sx.u, sy.u, sz.u, o, ax);
  f.v = muscl_extrap<DIM, AX, T, FAST>(q.v, sx.v, sy.v, sz.v, o, ax);
  f.w = DIM == 3 ? muscl_extrap<DIM, AX, T, FAST>(q.w, sx.w, sy.w, sz.w, o, ax) : T(0.0);
  if (HANCOCK)
  {
    f.r0 += h.r0;
    f.r1 += h.r1;
    f.p += h.p;
    f.u += h.u;
    f.v += h.v;
    if (DIM == 3)
      f.w += h.w;
  }
  return f;
}

// ---------------------------------------------------------------------------------------------
// models/RiemannSolvers.h:57-262 riemann_hllc, for a face normal to axis AX, followed by the flux
// scaling and the u<->v (u<->w) rotation of compute_fluxes_and_store_{2d,3d}
// (ComputeFluxesAndStoreFunctor.cpp:374-451, :502-640).  un/t1/t2 are the normal and the two
// transverse velocities in the reference's swapped order.
// ---------------------------------------------------------------------------------------------
template <int DIM>
G5D Flux
hllc(SolveDesc const & S, Face const & L, Face const & R, double dtdx, int AX)
{
  // swapped velocity slots (IU, IV, IW) after my_swap
  const double ulN = AX == 0 ? L.u : (AX == 1 ? L.v : L.w);
  const double ulT1 = AX == 1 ? L.u : L.v;
  const double ulT2 = AX == 2 ? L.u : L.w;
  const double urN = AX == 0 ? R.u : (AX == 1 ? R.v : R.w);
  const double urT1 = AX == 1 ? R.u : R.v;
  const double urT2 = AX == 2 ? R.u : R.w;

  const double smallr = S.smallr, smallp = S.smallp, smallc = S.smallc;

  const double rl = fmax(L.r0 + L.r1, smallr);
  double       pl = fmax(L.p, rl * smallp);
  const double ul = ulN;
  const double eintl = eos_eint(S, rl, pl, L.a, 1.0 - L.a);
  pl = eos_pressure(S, rl, eintl, L.a, 1.0 - L.a);
  double ecinl = 0.5 * rl * ul * ul;
  ecinl += 0.5 * rl * ulT1 * ulT1;
  if (DIM == 3)
    ecinl += 0.5 * rl * ulT2 * ulT2;
  const double xl = eos_xi(S, L.a, 1.0 - L.a);
  const double ql = (((xl + 1.0) * pl + eos_pi(S, L.a, 1.0 - L.a)) / xl) / rl; // cl^2 before the root
  const double etotl = rl * eintl + ecinl;
  const double ptotl = pl;

  const double rr = fmax(R.r0 + R.r1, smallr);
  double       pr = fmax(R.p, rr * smallp);
  const double ur = urN;
  const double eintr = eos_eint(S, rr, pr, R.a, 1.0 - R.a);
  pr = eos_pressure(S, rr, eintr, R.a, 1.0 - R.a);
  double ecinr = 0.5 * rr * ur * ur;
  ecinr += 0.5 * rr * urT1 * urT1;
  if (DIM == 3)
    ecinr += 0.5 * rr * urT2 * urT2;
  const double xr = eos_xi(S, R.a, 1.0 - R.a);
  const double qr = (((xr + 1.0) * pr + eos_pi(S, R.a, 1.0 - R.a)) / xr) / rr;
  const double etotr = rr * eintr + ecinr;
  const double ptotr = pr;

  // The wave speeds use max(cfastl, cfastr), cfast = sqrt(max(c*c, smallc*smallc)), c = sqrt(q).
  // With correctly rounded multiplication and square root in radix 2, sqrt(RN(x*x)) == x for every
  // x whose square neither overflows nor underflows, and RN(x*x) is monotone in x, so cfast equals
  // max(c, smallc) bit for bit when smallc^2 is normal and c < 1e150; and since sqrt is correctly
  // rounded, hence monotone, max(sqrt(ql), sqrt(qr)) == sqrt(max(ql, qr)) bit for bit (fmax skips
  // a NaN operand on both sides of the identity).  So one square root gives max(cfastl, cfastr).
  // The reference form is kept for the other cases.
  const double qm = fmax(ql, qr);
  double       cfast;
  if (smallc >= 1.5e-154 and qm < 1e300)
    cfast = fmax(sqrt(qm), smallc);
  else
  {
    const double cl = sqrt(ql), cr = sqrt(qr);
    cfast = fmax(sqrt(fmax(cl * cl, smallc * smallc)), sqrt(fmax(cr * cr, smallc * smallc)));
  }
  const double SL = fmin(ul, ur) - cfast;
  const double SR = fmax(ul, ur) + cfast;

  const double rcl = rl * (ul - SL);
  const double rcr = rr * (SR - ur);

  const double ustar0 = (rcr * ur + rcl * ul + (ptotl - ptotr)) / (rcr + rcl);
  const double ptotstar = (rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);

  // Sampling at x/t = 0.  The reference computes the star factor and energy of both sides and then
  // picks one of four states; here the side K of the fan (left when SL > 0 or u* > 0) is selected
  // first and only its star quantities are computed -- the same operations on the same operands,
  // so the same bits -- and the star state or, for a supersonic side, the side's own state taken.
  const bool   left = SL > 0.0 or ustar0 > 0.0;
  const bool   star = left ? !(SL > 0.0) : SR > 0.0;
  const double SK = left ? SL : SR, uK = left ? ul : ur, rK = left ? rl : rr;
  const double pK = left ? ptotl : ptotr, eK = left ? etotl : etotr;
  const double r0K = left ? L.r0 : R.r0, r1K = left ? L.r1 : R.r1;
  const double factorK = (SK - uK) / (SK - ustar0);
  const double etotstarK = ((SK - uK) * eK - pK * uK + ptotstar * ustar0) / (SK - ustar0);
  const double ro_0 = star ? r0K * factorK : r0K;
  const double ro_1 = star ? r1K * factorK : r1K;
  const double ro = star ? rK * factorK : rK;
  const double uo = star ? ustar0 : uK;
  const double ptoto = star ? ptotstar : pK;
  const double etoto = star ? etotstarK : eK;

  const double fr0 = ro_0 * uo;
  const double fr1 = ro_1 * uo;
  const double fd = (ro_0 + ro_1) * uo;
  const double fn = ro * uo * uo + ptoto;
  const double fe = (etoto + ptoto) * uo;
  const bool   up = fd > 0.0;
  const double ft1 = fd * (up ? ulT1 : urT1);
  const double ft2 = DIM == 3 ? fd * (up ? ulT2 : urT2) : 0.0;
  const double phi0 = up ? L.a : R.a;
  const double phi1 = up ? 1.0 - L.a : 1.0 - R.a;
  const double ustar = uo;

  Flux F;
  F.ar0 = fr0 * dtdx;
  F.ar1 = fr1 * dtdx;
  F.a0 = (ustar * phi0) * dtdx;
  F.a1 = (ustar * phi1) * dtdx;
  F.e = fe * dtdx;
  // un-swap: x: (IU,IV,IW) = (n,t1,t2); y: (t1,n,t2); z: (t2,t1,n)
  const double fu = AX == 0 ? fn : (AX == 1 ? ft1 : ft2);
  const double fv = AX == 1 ? fn : ft1;
  const double fw = AX == 2 ? fn : ft2;
  F.mu = fu * dtdx;
  F.mv = fv * dtdx;
  F.mw = DIM == 3 ? fw * dtdx : 0.0;
  F.us = ustar * dtdx;
  return F;
}

// Reassociated FP64 HLLC (G5_PREC_FP64_FAST): the same wave speeds, star state, sampling, flux,
// upwinding and scaling as hllc, with
//   - no e(p) -> p(e) round trip on each side (the identity in exact arithmetic): rho e = p xi + Pi;
//   - c^2 = ((xi+1) p + Pi)/(xi rho), one reciprocal, and one square root for max(cl, cr);
//   - u* and p* multiplied by one reciprocal of rcl + rcr, and the star factor and star energy of
//     the sampled side by one reciprocal of S - u*;
//   - the kinetic energy as rho/2 |u|^2, and the fluxes times dt/dx through u* dt/dx.
template <int DIM>
G5D Flux
hllc_fast(SolveDesc const & S, Face const & L, Face const & R, double dtdx, int AX)
{
  const double ulN = AX == 0 ? L.u : (AX == 1 ? L.v : L.w);
  const double ulT1 = AX == 1 ? L.u : L.v;
  const double ulT2 = AX == 2 ? L.u : L.w;
  const double urN = AX == 0 ? R.u : (AX == 1 ? R.v : R.w);
  const double urT1 = AX == 1 ? R.u : R.v;
  const double urT2 = AX == 2 ? R.u : R.w;
  const double smallr = S.smallr, smallp = S.smallp, smallc = S.smallc;

  const double rl = fmax(L.r0 + L.r1, smallr);
  const double pl = fmax(L.p, rl * smallp);
  const double ul = ulN;
  const double xl = eos_xi(S, L.a, 1.0 - L.a), pil = eos_pi(S, L.a, 1.0 - L.a);
  double       kl = ul * ul + ulT1 * ulT1; // |u|^2, one product by rl/2 for the kinetic energy
  if (DIM == 3)
    kl += ulT2 * ulT2;
  const double ecinl = 0.5 * rl * kl;
  const double ql = ((xl + 1.0) * pl + pil) * rcp64(xl * rl);
  const double etotl = (pl * xl + pil) + ecinl;

  const double rr = fmax(R.r0 + R.r1, smallr);
  const double pr = fmax(R.p, rr * smallp);
  const double ur = urN;
  const double xr = eos_xi(S, R.a, 1.0 - R.a), pir = eos_pi(S, R.a, 1.0 - R.a);
  double       kr = ur * ur + urT1 * urT1; // |u|^2, one product by rr/2 for the kinetic energy
  if (DIM == 3)
    kr += urT2 * urT2;
  const double ecinr = 0.5 * rr * kr;
  const double qr = ((xr + 1.0) * pr + pir) * rcp64(xr * rr);
  const double etotr = (pr * xr + pir) + ecinr;

  const double cfast = fmax(sqrt64(fmax(fmax(ql, qr), 0.0)), smallc);
  const double SL = fmin(ul, ur) - cfast;
  const double SR = fmax(ul, ur) + cfast;

  const double rcl = rl * (ul - SL);
  const double rcr = rr * (SR - ur);
  const double inv_rc = rcp64(rcr + rcl);
  const double ustar0 = (rcr * ur + rcl * ul + (pl - pr)) * inv_rc;
  const double ptotstar = (rcr * pl + rcl * pr + rcl * rcr * (ul - ur)) * inv_rc;

  double ro, ro_0, ro_1, uo, ptoto, etoto;
  if constexpr (DIM == 3)
  {
    // sampling at x/t = 0 without branches: the fan side K (left when SL > 0 or u* > 0), then its
    // star state or, for a supersonic side, its own state -- the same operations on the selected
    // operands as the four branches of hllc
    const bool   left = SL > 0.0 or ustar0 > 0.0;
    const bool   star = left ? !(SL > 0.0) : SR > 0.0;
    const double SK = left ? SL : SR, uK = left ? ul : ur, rK = left ? rl : rr, pK = left ? pl : pr;
    const double r0K = left ? L.r0 : R.r0, r1K = left ? L.r1 : R.r1, eK = left ? etotl : etotr;
    const double inv = rcp64(SK - ustar0);
    const double factor = (SK - uK) * inv;
    ro_0 = star ? r0K * factor : r0K;
    ro_1 = star ? r1K * factor : r1K;
    ro = star ? rK * factor : rK;
    uo = star ? ustar0 : uK;
    ptoto = star ? ptotstar : pK;
    etoto = star ? ((SK - uK) * eK - pK * uK + ptotstar * ustar0) * inv : eK;
  }
  else
  {
    // 2-D: the four branches (the selects of the branch-free form raise the 2-D step from 130 to
    // 185 registers: two wavefronts per SIMD instead of three)
    if (SL > 0.0)
    {
      ro_0 = L.r0;
      ro_1 = L.r1;
      ro = rl;
      uo = ul;
      ptoto = pl;
      etoto = etotl;
    }
    else if (ustar0 > 0.0)
    {
      const double inv = rcp64(SL - ustar0);
      const double factor = (SL - ul) * inv;
      ro_0 = L.r0 * factor;
      ro_1 = L.r1 * factor;
      ro = rl * factor;
      uo = ustar0;
      ptoto = ptotstar;
      etoto = ((SL - ul) * etotl - pl * ul + ptotstar * ustar0) * inv;
    }
    else if (SR > 0.0)
    {
      const double inv = rcp64(SR - ustar0);
      const double factor = (SR - ur) * inv;
      ro_0 = R.r0 * factor;
      ro_1 = R.r1 * factor;
      ro = rr * factor;
      uo = ustar0;
      ptoto = ptotstar;
      etoto = ((SR - ur) * etotr - pr * ur + ptotstar * ustar0) * inv;
    }
    else
    {
      ro_0 = R.r0;
      ro_1 = R.r1;
      ro = rr;
      uo = ur;
      ptoto = pr;
      etoto = etotr;
    }
  }

  // the fluxes times dt/dx, with u* dt/dx formed once
  const double fd = (ro_0 + ro_1) * uo;
  const bool   up = fd > 0.0;
  const double uod = uo * dtdx;
  const double fdd = fd * dtdx;
  const double fn = (ro * uo) * uod + ptoto * dtdx;
  const double ft1 = fdd * (up ? ulT1 : urT1);
  const double ft2 = DIM == 3 ? fdd * (up ? ulT2 : urT2) : 0.0;
  const double phi0 = up ? L.a : R.a;
  const double phi1 = up ? 1.0 - L.a : 1.0 - R.a;

  Flux F;
  F.ar0 = ro_0 * uod;
  F.ar1 = ro_1 * uod;
  F.a0 = phi0 * uod;
  F.a1 = phi1 * uod;
  F.e = (etoto + ptoto) * uod;
  F.mu = AX == 0 ? fn : (AX == 1 ? ft1 : ft2);
  F.mv = AX == 1 ? fn : ft1;
  F.mw = DIM == 3 ? (AX == 2 ? fn : ft2) : 0.0;
  F.us = uod;
  return F;
}

// Mixed-precision HLLC on FP32 face states: wave speeds, star velocity/pressure, mass, momentum
// and volume-fraction fluxes in FP32; the total energies, their star values and the energy flux in
// FP64 (rho e = p xi + Pi is dominated by Pi when pinf is large, and the update takes differences
// of energy fluxes).  The reference's e(p) -> p(e) round trip on each side is the identity in exact
// arithmetic and is not repeated in FP32, where it would lose the pressure to the cancellation.
template <int DIM>
G5D FluxT<float>
hllc_mixed(SolveDesc const & S, FaceT<float> const & L, FaceT<float> const & R, double, int AX)
{
  const float ulN = AX == 0 ? L.u : (AX == 1 ? L.v : L.w);
  const float ulT1 = AX == 1 ? L.u : L.v;
  const float ulT2 = AX == 2 ? L.u : L.w;
  const float urN = AX == 0 ? R.u : (AX == 1 ? R.v : R.w);
  const float urT1 = AX == 1 ? R.u : R.v;
  const float urT2 = AX == 2 ? R.u : R.w;

  const float smallr = S.smallrf, smallp = S.smallpf, smallc = S.smallcf;

  // total energies in FP64 from the FP32 face states: rho e = p xi + Pi, plus the kinetic energy
  auto etot_of = [&](float r, float p, float a, float un, float t1, float t2) {
    const double ad = a, ud = un, t1d = t1, t2d = t2;
    double       k = ud * ud + t1d * t1d;
    if (DIM == 3)
      k += t2d * t2d;
    return (static_cast<double>(p) * eos_xi(S, ad, 1.0 - ad) + eos_pi(S, ad, 1.0 - ad)) +
           0.5 * static_cast<double>(r) * k;
  };

  const float rl = fmaxf(L.r0 + L.r1, smallr);
  const float pl = fmaxf(L.p, rl * smallp);
  const float ul = ulN;
  const float cl = eos_sound_speed(S, rl, pl, L.a, 1.0f - L.a);

  const float rr = fmaxf(R.r0 + R.r1, smallr);
  const float pr = fmaxf(R.p, rr * smallp);
  const float ur = urN;
  const float cr = eos_sound_speed(S, rr, pr, R.a, 1.0f - R.a);

  const float cfastl = fmaxf(cl, smallc);
  const float cfastr = fmaxf(cr, smallc);
  const float SL = fminf(ul, ur) - fmaxf(cfastl, cfastr);
  const float SR = fmaxf(ul, ur) + fmaxf(cfastl, cfastr);

  // star velocity and pressure written as corrections of the left state, so that a face with
  // equal velocities and pressures on both sides gets u* = u and p* = p exactly in FP32 (the
  // HLLC contact property survives the rounding: an interface advected at uniform u and p stays
  // in equilibrium)
  const float rcl = rl * (ul - SL);
  const float rcr = rr * (SR - ur);
  const float inv_rc = 1.0f / (rcr + rcl);
  const float ustar0 = ul + (rcr * (ur - ul) + (pl - pr)) * inv_rc;
  const float ptotstar = pl + (rcl * (pr - pl) + rcl * rcr * (ul - ur)) * inv_rc;

  float  ro, ro_0, ro_1, uo, ptoto;
  double etoto;
  if constexpr (DIM == 3)
  {
    // sampling at x/t = 0 without branches: the fan side K (left when SL > 0 or u* > 0), then its
    // star state or, for a supersonic side, its own state; the total energy of side K alone
    // etot* = f etot + (p* u* - p u)/(S - u*), f = (S - u)/(S - u*) (exactly 1 when u* = u)
    const bool   left = SL > 0.0f or ustar0 > 0.0f;
    const bool   star = left ? !(SL > 0.0f) : SR > 0.0f;
    const float  SK = left ? SL : SR, uK = left ? ul : ur, rK = left ? rl : rr, pK = left ? pl : pr;
    const float  r0K = left ? L.r0 : R.r0, r1K = left ? L.r1 : R.r1, aK = left ? L.a : R.a;
    const double eK = etot_of(rK, pK, aK, uK, left ? ulT1 : urT1, left ? ulT2 : urT2);
    const float  inv = 1.0f / (SK - ustar0);
    const float  factor = ustar0 == uK ? 1.0f : (SK - uK) * inv;
    ro_0 = star ? r0K * factor : r0K;
    ro_1 = star ? r1K * factor : r1K;
    ro = star ? rK * factor : rK;
    uo = star ? ustar0 : uK;
    ptoto = star ? ptotstar : pK;
    etoto =
      star ? static_cast<double>(factor) * eK + static_cast<double>((ptotstar * ustar0 - pK * uK) * inv) : eK;
  }
  else
  {
    // 2-D: the four branches (the selects of the branch-free form cost the 2-D step registers
    // and a wavefront per SIMD)
    const double etotl = etot_of(rl, pl, L.a, ul, ulT1, ulT2);
    const double etotr = etot_of(rr, pr, R.a, ur, urT1, urT2);
    if (SL > 0.0f)
    {
      ro_0 = L.r0;
      ro_1 = L.r1;
      ro = rl;
      uo = ul;
      ptoto = pl;
      etoto = etotl;
    }
    else if (ustar0 > 0.0f)
    {
      // etot* = f etot + (p* u* - p u)/(S - u*), f = (S - u)/(S - u*) (exactly 1 when u* = u)
      const float inv = 1.0f / (SL - ustar0);
      const float factor = ustar0 == ul ? 1.0f : (SL - ul) * inv;
      ro_0 = L.r0 * factor;
      ro_1 = L.r1 * factor;
      ro = rl * factor;
      uo = ustar0;
      ptoto = ptotstar;
      etoto = static_cast<double>(factor) * etotl + static_cast<double>((ptotstar * ustar0 - pl * ul) * inv);
    }
    else if (SR > 0.0f)
    {
      const float inv = 1.0f / (SR - ustar0);
      const float factor = ustar0 == ur ? 1.0f : (SR - ur) * inv;
      ro_0 = R.r0 * factor;
      ro_1 = R.r1 * factor;
      ro = rr * factor;
      uo = ustar0;
      ptoto = ptotstar;
      etoto = static_cast<double>(factor) * etotr + static_cast<double>((ptotstar * ustar0 - pr * ur) * inv);
    }
    else
    {
      ro_0 = R.r0;
      ro_1 = R.r1;
      ro = rr;
      uo = ur;
      ptoto = pr;
      etoto = etotr;
    }
  }

  const bool   up = (ro_0 + ro_1) * uo > 0.0f;
  FluxT<float> F;
  F.ro0 = ro_0;
  F.ro1 = ro_1;
  F.ro = ro;
  F.uo = uo;
  F.pt = ptoto;
  F.t1 = up ? ulT1 : urT1;
  F.t2 = DIM == 3 ? (up ? ulT2 : urT2) : 0.0f;
  F.phi0 = up ? L.a : R.a;
  F.etot = etoto;
  return F;
}

template <int DIM, bool FAST = false>
G5D Flux
riemann(SolveDesc const & S, Face const & L, Face const & R, double dtdx, int AX)
{
  if constexpr (FAST)
    return hllc_fast<DIM>(S, L, R, dtdx, AX);
  else
    return hllc<DIM>(S, L, R, dtdx, AX);
}
template <int DIM, bool FAST = false>
G5D FluxT<float>
riemann(SolveDesc const & S, FaceT<float> const & L, FaceT<float> const & R, double dtdx, int AX)
{
  return hllc_mixed<DIM>(S, L, R, dtdx, AX);
}

// ---------------------------------------------------------------------------------------------
// scheme/ReadFluxesAndConservativeUpdateFunctor.cpp:146-212 (same-level branch) + update_U
// (.h:319-393): accumulate left minus right flux with the Allaire-Kokh alpha source term (time-n
// alpha of the cell), then add to the running U2 with the 1e-12 clamps.
// ---------------------------------------------------------------------------------------------
// st0, st1: the source term factors (1 for ALLAIRE_KOKH, source_factors for MILLER_PUCKETT).
// FAST (the reassociated forms): the differences FL - FR with the source term on the difference of
// u* dt/dx, instead of the reference's accumulation from 0 face by face.
template <int DIM, bool FAST = false>
G5D void
apply_flux_pair(Cons & u2, Flux const & FL, Flux const & FR, double ua0, double ua1, double st0 = 1.0,
                double st1 = 1.0)
{
  const double eps = 1e-12;
  if constexpr (FAST)
  {
    const double dus = FL.us - FR.us;
    u2.ar0 += FL.ar0 - FR.ar0;
    if (u2.ar0 < eps)
      u2.ar0 = 0.0;
    u2.ar1 += FL.ar1 - FR.ar1;
    if (u2.ar1 < eps)
      u2.ar1 = 0.0;
    u2.a0 += (FL.a0 - FR.a0) - st0 * ua0 * dus;
    if (u2.a0 < eps)
      u2.a0 = 0.0;
    if (u2.a0 > 1.0 - eps)
      u2.a0 = 1.0;
    u2.a1 += (FL.a1 - FR.a1) - st1 * ua1 * dus;
    if (u2.a1 < eps)
      u2.a1 = 0.0;
    if (u2.a1 > 1.0 - eps)
      u2.a1 = 1.0;
    u2.e += FL.e - FR.e;
    u2.mu += FL.mu - FR.mu;
    u2.mv += FL.mv - FR.mv;
    if (DIM == 3)
      u2.mw += FL.mw - FR.mw;
    return;
  }
  double       acc_r0 = 0.0 + FL.ar0;
  double       acc_r1 = 0.0 + FL.ar1;
  double       acc_a0 = 0.0 + FL.a0;
  double       acc_a1 = 0.0 + FL.a1;
  double       acc_e = 0.0 + FL.e;
  double       acc_mu = 0.0 + FL.mu;
  double       acc_mv = 0.0 + FL.mv;
  double       acc_mw = 0.0 + FL.mw;
  acc_a0 -= st0 * FL.us * ua0;
  acc_a1 -= st1 * FL.us * ua1;
  acc_r0 = acc_r0 - FR.ar0;
  acc_r1 = acc_r1 - FR.ar1;
  acc_a0 = acc_a0 - FR.a0;
  acc_a1 = acc_a1 - FR.a1;
  acc_e = acc_e - FR.e;
  acc_mu = acc_mu - FR.mu;
  acc_mv = acc_mv - FR.mv;
  acc_mw = acc_mw - FR.mw;
  acc_a0 += st0 * FR.us * ua0;
  acc_a1 += st1 * FR.us * ua1;

  u2.ar0 += acc_r0;
  if (u2.ar0 < eps)
    u2.ar0 = 0.0;
  u2.ar1 += acc_r1;
  if (u2.ar1 < eps)
    u2.ar1 = 0.0;
  u2.a0 += acc_a0;
  if (u2.a0 < eps)
    u2.a0 = 0.0;
  if (u2.a0 > 1.0 - eps)
    u2.a0 = 1.0;
  u2.a1 += acc_a1;
  if (u2.a1 < eps)
    u2.a1 = 0.0;
  if (u2.a1 > 1.0 - eps)
    u2.a1 = 1.0;
  u2.e += acc_e;
  u2.mu += acc_mu;
  u2.mv += acc_mv;
  if (DIM == 3)
    u2.mw += acc_mw;
}

// Mixed path: the FP64 face flux (scaled by dt/dx, rotated back to x, y, z momentum) from the
// FP32 star state of an AX face; the operations of the flux part of hllc_fast.
template <int DIM>
G5D Flux
assemble_flux(FluxT<float> const & s, double dtdx, int AX)
{
  const double ro0 = s.ro0, ro1 = s.ro1, ro = s.ro, uo = s.uo, pt = s.pt;
  const double uod = uo * dtdx; // the fluxes times dt/dx through u* dt/dx
  const double fdd = (ro0 + ro1) * uod;
  const double fn = (ro * uo) * uod + pt * dtdx;
  const double ft1 = fdd * static_cast<double>(s.t1);
  const double ft2 = DIM == 3 ? fdd * static_cast<double>(s.t2) : 0.0;
  const double phi0 = s.phi0;
  const double phi1 = 1.0 - phi0;
  Flux         F;
  F.ar0 = ro0 * uod;
  F.ar1 = ro1 * uod;
  F.a0 = phi0 * uod;
  F.a1 = phi1 * uod;
  F.e = (s.etot + pt) * uod;
  F.mu = AX == 0 ? fn : (AX == 1 ? ft1 : ft2);
  F.mv = AX == 1 ? fn : ft1;
  F.mw = DIM == 3 ? (AX == 2 ? fn : ft2) : 0.0;
  F.us = uod;
  return F;
}

// the update with the face records of either path (dtdx and AX are used by the mixed path)
template <int DIM, bool FAST = false>
G5D void
apply_face_pair(Cons & u2, Flux const & FL, Flux const & FR, double ua0, double ua1, double, int, double st0 = 1.0,
                double st1 = 1.0)
{
  apply_flux_pair<DIM, FAST>(u2, FL, FR, ua0, ua1, st0, st1);
}
template <int DIM, bool FAST = false>
G5D void
apply_face_pair(Cons & u2, FluxT<float> const & FL, FluxT<float> const & FR, double ua0, double ua1, double dtdx,
                int AX, double st0 = 1.0, double st1 = 1.0)
{
  apply_flux_pair<DIM, FAST>(u2, assemble_flux<DIM>(FL, dtdx, AX), assemble_flux<DIM>(FR, dtdx, AX), ua0, ua1, st0,
                             st1);
}

// scheme/ReadFluxesAndConservativeUpdateFunctor.h:412-443 get_source_term_factor for
// MILLER_PUCKETT, from the cell's time-n state U: the pressure as compute_mixture_pressure
// (models/utils.h:160-182: density alpha_rho0 + alpha_rho1, kinetic energy with U[ID],
// alpha1 = 1 - alpha0), the material bulk moduli of the stiffened gas K_k = gamma_k (p + pinf_k),
// the mixture bulk modulus by Wood's formula K = 1 / (alpha0/K0 + alpha1/K1) with the stored
// alpha1, and st_k = 1 - K/K_k.  (The reference's bulk moduli come from its EOS class, which is
// not in the reference tree: these two formulas are the textbook ones for a stiffened gas at
// pressure equilibrium.)
template <int DIM>
G5D void
source_factors(SolveDesc const & S, Cons const & U, double & st0, double & st1)
{
  const double d = fmax(U.ar0 + U.ar1, S.smallr);
  const double dk = fmax(U.d, S.smallr);
  const double ux = U.mu / dk;
  const double uy = U.mv / dk;
  double       ekin;
  if (DIM == 2)
    ekin = 0.5 * (ux * ux + uy * uy);
  else
  {
    const double uz = U.mw / dk;
    ekin = 0.5 * (ux * ux + uy * uy + uz * uz);
  }
  const double eint = U.e / d - ekin;
  const double p = eos_pressure(S, d, eint, U.a0, 1.0 - U.a0);
  const double k0 = S.gamma[0] * (p + S.pinf[0]);
  const double k1 = S.gamma[1] * (p + S.pinf[1]);
  const double kbar = 1.0 / (U.a0 / k0 + U.a1 / k1);
  st0 = 1.0 - kbar / k0;
  st1 = 1.0 - kbar / k1;
}

// scheme/FixVolumeFractionsNormalization.cpp:58-70 (FAST -- the reassociated FP64 and the mixed
// paths: the reciprocal by rcp64)
template <bool FAST = false>
G5D void
fix_volume_fractions(Cons & u)
{
  const double sum_alpha = u.a0 + u.a1;
  const double sum_alpha_rho = u.ar0 + u.ar1;
  const double inv = FAST ? rcp64(sum_alpha) : 1.0 / sum_alpha;
  u.a0 *= inv;
  u.a1 *= inv;
  u.d = sum_alpha_rho;
}

// the solve's descriptor as a block of the given level sees it: the fine blocks (level 1) of a
// two-level solve have half the cell length (compute_cell_length(level, bx), halved per level;
// both scalings are exact)
G5D SolveDesc
at_level(SolveDesc S, int level)
{
  if (level != 0)
  {
    S.dx *= 0.5;
    S.inv_dx_p2 *= 2.0;
  }
  return S;
}

// scheme/ComputeDtFiveEqFunctor.cpp:85-134 compute_cfl (one cell's contribution; FAST: with the
// reassociated conversion)
template <int DIM, bool FAST = false>
G5D double
cell_inv_dt(SolveDesc const & S, Cons const & u, double & p)
{
  Prim         q;
  const double c = to_prim<DIM, FAST>(S, u, q);
  p = q.p;
  const double vx = c + fabs(q.u);
  const double vy = c + fabs(q.v);
  const double vz = DIM == 3 ? c + fabs(q.w) : 0.0;
  if constexpr (FAST)
  {
    // products by 1/dx (exact for dx = 2^k; otherwise within an ulp of the quotients), with no
    // branch in the step's loop: the reciprocal does not depend on the cell
    const double r = S.inv_dx_p2 > 0.0 ? S.inv_dx_p2 : 1.0 / S.dx;
    return DIM == 2 ? vx * r + vy * r : vx * r + vy * r + vz * r;
  }
  if (S.inv_dx_p2 > 0.0) // dx = 2^k: the quotients by dx are the exact products by 2^-k
  {
    const double r = S.inv_dx_p2;
    return DIM == 2 ? vx * r + vy * r : vx * r + vy * r + vz * r;
  }
  if (DIM == 2)
    return vx / S.dx + vy / S.dx;
  return vx / S.dx + vy / S.dx + vz / S.dx;
}

// mixed path: FP32 sound speed and velocities from the FP64-recovered pressure; the rate and its
// max reduction are FP64
template <int DIM>
G5D double
cell_inv_dt_mixed(SolveDesc const & S, Cons const & u, double & p)
{
  PrimT<float> q;
  const float  c = cons_to_prim_mixed<DIM>(S, u, q);
  p = q.p;
  float        v = (c + fabsf(q.u)) + (c + fabsf(q.v));
  if (DIM == 3)
    v += c + fabsf(q.w);
  return S.inv_dx_p2 > 0.0 ? static_cast<double>(v) * S.inv_dx_p2 : static_cast<double>(v) / S.dx;
}

// the CFL rate of a cell of the path of T (and FAST); p receives the cell's mixture pressure
template <int DIM, typename T, bool FAST = false>
G5D double
cell_rate(SolveDesc const & S, Cons const & u, double & p)
{
  if constexpr (sizeof(T) == sizeof(double))
    return cell_inv_dt<DIM, FAST>(S, u, p);
  else
    return cell_inv_dt_mixed<DIM>(S, u, p);
}

// G5_FAIL_* bits of one cell of a new state: a conservative value that is not finite (the
// conversion clamps the density and the volume fraction with fmax/fmin, which drop a NaN, so the
// state itself is tested: the difference s - s of the sum of its values is 0 only when the sum is
// finite; NaN compares false), and with the solve's check a negative pressure.  A cell whose
// sound speed is not real (c^2 < 0: the mixture EOS outside its domain, which needs p < 0) has a
// NaN CFL rate; the max reductions drop it, as the reference's do, and the solve goes on unless
// the negative-pressure check is on (a solve with no finite rate at all fails at its dt)
G5D uint32_t
cell_failure(SolveDesc const & S, Cons const & u, double, double p)
{
  const double sum = ((u.d + u.e) + (u.mu + u.mv) + (u.mw + u.a0)) + ((u.ar0 + u.a1) + u.ar1);
  const bool   ok = sum - sum == 0.0;
  uint32_t     f = ok ? 0u : static_cast<uint32_t>(G5_FAIL_NONFINITE);
  if (S.abort_neg_p != 0 and p < 0.0)
    f |= static_cast<uint32_t>(G5_FAIL_NEGATIVE_PRESSURE);
  return f;
}

// OR of the failure bits of a wavefront's lanes into the solve's control record: two ballots,
// and one atomic by the first failing lane only when a lane failed (every lane of the wavefront
// calls this)
G5D void
report_failure(SolveCtl * C, uint32_t f)
{
  const unsigned long long b1 = __ballot((f & 1u) != 0u);
  const unsigned long long b2 = __ballot((f & 2u) != 0u);
  const unsigned long long any = b1 | b2;
  if (any != 0ull and static_cast<int>(threadIdx.x & 63) == __ffsll(static_cast<long long>(any)) - 1)
    atomicOr(&C->failure, (b1 != 0ull ? 1u : 0u) | (b2 != 0ull ? 2u : 0u));
}

// Blocks of the launch list handled by this workgroup: [b, end) with the given stride.  One
// workgroup per block, or persistent workgroups looping over a range; with remap, the
// workgroups of XCD x (workgroup ids are dealt round-robin to the 8 XCDs) take the contiguous
// chunk [x * per_xcd, (x + 1) * per_xcd), so neighbouring blocks, which re-read each other's
// halo, share an L2.
G5D void
block_range(StepArgs const & A, int & b, int & end, int & stride)
{
  const int w = blockIdx.x;
  if (A.wg_per_xcd == 0)
  {
    b = A.remap ? (w & 7) * A.per_xcd + (w >> 3) : w;
    end = min(b + 1, A.nblocks);
    stride = 1;
  }
  else if (A.remap)
  {
    const int x = w & 7;
    b = x * A.per_xcd + (w >> 3);
    end = min((x + 1) * A.per_xcd, A.nblocks);
    stride = A.wg_per_xcd;
  }
  else
  {
    b = w;
    end = A.nblocks;
    stride = static_cast<int>(gridDim.x);
  }
}

// ---------------------------------------------------------------------------------------------
// ghost/border resolution.  Block N (brick coordinates, already wrapped in periodic directions,
// possibly one step outside the brick) and in-block cell c map to (offset of the source cell in the
// state buffer through the solve's slot table, negated momentum axis): same-level neighbour lookup of StencilHelper plus the outside-block fill
// of border_conditions/FillOutside.cpp:69-170 -- the inside block is the outside block moved back
// along its outside normal, the faces are visited in order xmin..zmax and the last one that
// applies defines the copy (ZERO_GRADIENT: edge cell; WALL: mirrored cell, normal momentum
// negated).
// ---------------------------------------------------------------------------------------------
// I: the inside block (N itself when N is inside the domain); cc: the cell of I to read.
// INLET false: the caller knows that no solve it serves has an inflow border.
template <int DIM, int B, bool INLET = true>
G5D void
remap_block_cell(SolveDesc const & S, int const (&N)[3], int const (&c)[3], int (&I)[3], int (&cc)[3], int & neg)
{
  int  n[3];
  bool outside = false;
  neg = -1;
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    n[d] = d < DIM ? (N[d] < 0 ? -1 : (N[d] >= S.nb[d] ? 1 : 0)) : 0;
    outside = outside or (n[d] != 0);
    I[d] = N[d] - n[d];
  }
  cc[0] = c[0];
  cc[1] = c[1];
  cc[2] = c[2];
  if (outside)
  {
#pragma unroll
    for (int face = 0; face < 2 * DIM; ++face)
    {
      const int  d = face / 2;
      const bool left = (face % 2) == 0;
      const bool at_border = left ? (I[d] == 0) : (I[d] == S.nb[d] - 1);
      if (at_border and n[d] != 0)
      {
        if (S.bc[face] == G5_BC_ZERO_GRADIENT)
        {
          cc[0] = c[0];
          cc[1] = c[1];
          cc[2] = c[2];
          cc[d] = left ? 0 : B - 1;
          neg = -1;
        }
        else if (S.bc[face] == G5_BC_WALL)
        {
          cc[0] = c[0];
          cc[1] = c[1];
          cc[2] = c[2];
          cc[d] = B - 1 - c[d];
          neg = d;
        }
      }
    }
  }
  // inflow override (SolverGodunovFiveEq.cpp:997-1018, after the border fill): outside cells with
  // x < xmin, for the jet those whose centre is inside it; neg = -2 marks them
  if (INLET and S.inlet != G5_INLET_NONE and N[0] < 0)
  {
    bool in = true;
    if (S.inlet == G5_INLET_JET)
    {
      const double y = S.origin[1] + (static_cast<double>(N[1] * B + c[1]) + 0.5) * S.dx;
      if (DIM == 2)
        in = y > S.jet[0] - S.jet[2] / 2 and y < S.jet[0] + S.jet[2] / 2;
      else
      {
        const double z = S.origin[2] + (static_cast<double>(N[2] * B + c[2]) + 0.5) * S.dx;
        const double r2 = (y - S.jet[0]) * (y - S.jet[0]) + (z - S.jet[1]) * (z - S.jet[1]);
        in = r2 < S.jet[2] * S.jet[2] / 4;
      }
    }
    if (in)
      neg = -2;
  }
}

template <int DIM, int B>
G5D int64_t
cell_offset(int blk, int const (&cc)[3])
{
  const int         cell = cc[0] + B * (cc[1] + (DIM == 3 ? B * cc[2] : 0));
  constexpr int64_t NV = DIM == 2 ? 8 : 9;
  constexpr int64_t NC = DIM == 2 ? B * B : B * B * B;
  return static_cast<int64_t>(blk) * NV * NC + cell;
}

template <int DIM, int B>
G5D int64_t
resolve_block_cell(SolveDesc const & S, int32_t const * __restrict__ slots, int const (&N)[3], int const (&c)[3],
                   int & neg)
{
  int I[3], cc[3];
  remap_block_cell<DIM, B>(S, N, c, I, cc, neg);
  return cell_offset<DIM, B>(slots[S.base + I[0] + S.nb[0] * (I[1] + S.nb[1] * I[2])], cc);
}

// the cell at tile coordinates g (in [-2, B+1]^d) of owned block (bx,by,bz)
// neighbour block N (periodic wrap applied) of block b in direction dir
template <int DIM>
G5D void
neighbour_block(SolveDesc const & S, int const (&b3)[3], int const (&dir)[3], int (&N)[3])
{
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    if (d >= DIM)
    {
      N[d] = 0;
      continue;
    }
    int Nd = b3[d] + dir[d];
    if (S.periodic[d])
      Nd = Nd < 0 ? Nd + S.nb[d] : (Nd >= S.nb[d] ? Nd - S.nb[d] : Nd);
    N[d] = Nd;
  }
}

template <int DIM, int B>
G5D void
tile_cell_dir(int gi, int gj, int gk, int (&dir)[3], int (&c)[3])
{
  const int g[3] = { gi, gj, gk };
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    dir[d] = d < DIM ? (g[d] < 0 ? -1 : (g[d] >= B ? 1 : 0)) : 0;
    c[d] = d < DIM ? g[d] - dir[d] * B : 0;
  }
}

template <int DIM, int B>
G5D int64_t
resolve_cell(SolveDesc const & S, int32_t const * __restrict__ slots, int bx, int by, int bz, int gi, int gj, int gk,
             int & neg)
{
  const int b3[3] = { bx, by, bz };
  int       dir[3], c[3], N[3];
  tile_cell_dir<DIM, B>(gi, gj, gk, dir, c);
  neighbour_block<DIM>(S, b3, dir, N);
  return resolve_block_cell<DIM, B>(S, slots, N, c, neg);
}

// The slots of the 3^DIM blocks a tile reads, resolved once per block into the table
// tab[(dx+1) + 3(dy+1) + 9(dz+1)] (entry k by thread k, k < 3^DIM): the inside block of the
// neighbour in direction (dx,dy,dz) does not depend on the cell.
template <int DIM>
G5D void
fill_slot_table(SolveDesc const & S, int32_t const * __restrict__ slots, int bx, int by, int bz, int k, int * tab)
{
  if (k >= (DIM == 2 ? 9 : 27))
    return;
  const int b3[3] = { bx, by, bz };
  const int dir[3] = { k % 3 - 1, (k / 3) % 3 - 1, DIM == 3 ? k / 9 - 1 : 0 };
  const int c[3] = { 0, 0, 0 };
  int       N[3], I[3], cc[3], neg;
  neighbour_block<DIM>(S, b3, dir, N);
  remap_block_cell<DIM, 16, false>(S, N, c, I, cc, neg); // the inside block only
  tab[k] = slots[S.base + I[0] + S.nb[0] * (I[1] + S.nb[1] * I[2])];
}

// resolve_cell with the block slot taken from a table made by fill_slot_table
template <int DIM, int B, bool INLET = true>
G5D int64_t
resolve_cell_tab(SolveDesc const & S, int const * tab, int bx, int by, int bz, int gi, int gj, int gk, int & neg)
{
  const int b3[3] = { bx, by, bz };
  int       dir[3], c[3], N[3], I[3], cc[3];
  tile_cell_dir<DIM, B>(gi, gj, gk, dir, c);
  neighbour_block<DIM>(S, b3, dir, N);
  remap_block_cell<DIM, B, INLET>(S, N, c, I, cc, neg);
  return cell_offset<DIM, B>(tab[(dir[0] + 1) + 3 * (dir[1] + 1) + 9 * (dir[2] + 1)], cc);
}

// the mirrored cell of a WALL border: momentum component neg negated (neg = -1: none)
G5D void
negate_momentum(Cons & u, int neg)
{
  u.mu = neg == 0 ? u.mu * -1.0 : u.mu;
  u.mv = neg == 1 ? u.mv * -1.0 : u.mv;
  u.mw = neg == 2 ? u.mw * -1.0 : u.mw;
}

// the outside-cell transform a loaded source cell still needs: the inflow state (neg = -2) or the
// WALL momentum sign (neg = 0, 1, 2)
template <int DIM, bool INLET = true>
G5D void
apply_border(SolveDesc const & S, Cons & u, int neg)
{
  if (INLET and neg == -2)
  {
    u.d = S.inlet_u[0];
    u.e = S.inlet_u[1];
    u.mu = S.inlet_u[2];
    u.mv = S.inlet_u[3];
    u.mw = DIM == 3 ? S.inlet_u[4] : 0.0;
    u.a0 = S.inlet_u[DIM == 3 ? 5 : 4];
    u.ar0 = S.inlet_u[DIM == 3 ? 6 : 5];
    u.a1 = S.inlet_u[DIM == 3 ? 7 : 6];
    u.ar1 = S.inlet_u[DIM == 3 ? 8 : 7];
  }
  else
    negate_momentum(u, neg);
}

// load the conservative state of one cell: base = offset of (block, var 0, cell)
template <int DIM, int B>
G5D Cons
load_cons(double const * __restrict__ U, int64_t base, int neg)
{
  constexpr int NC = DIM == 2 ? B * B : B * B * B;
  Cons          u;
  u.d = U[base + 0 * NC];
  u.e = U[base + 1 * NC];
  u.mu = U[base + 2 * NC];
  u.mv = U[base + 3 * NC];
  if (DIM == 3)
  {
    u.mw = U[base + 4 * NC];
    u.a0 = U[base + 5 * NC];
    u.ar0 = U[base + 6 * NC];
    u.a1 = U[base + 7 * NC];
    u.ar1 = U[base + 8 * NC];
  }
  else
  {
    u.mw = 0.0;
    u.a0 = U[base + 4 * NC];
    u.ar0 = U[base + 5 * NC];
    u.a1 = U[base + 6 * NC];
    u.ar1 = U[base + 7 * NC];
  }
  negate_momentum(u, neg);
  return u;
}

// non-temporal stores: the new state is read back only by the next launch, so it is written
// past L2, leaving L2 to the halo reads (measured 1% faster on the 4096-solve batch and the mixed
// 3-D step, neutral elsewhere)
template <int DIM, int B>
G5D void
store_cons(double * __restrict__ U, int64_t base, Cons const & u)
{
  constexpr int NC = DIM == 2 ? B * B : B * B * B;
  __builtin_nontemporal_store(u.d, &U[base + 0 * NC]);
  __builtin_nontemporal_store(u.e, &U[base + 1 * NC]);
  __builtin_nontemporal_store(u.mu, &U[base + 2 * NC]);
  __builtin_nontemporal_store(u.mv, &U[base + 3 * NC]);
  if (DIM == 3)
  {
    __builtin_nontemporal_store(u.mw, &U[base + 4 * NC]);
    __builtin_nontemporal_store(u.a0, &U[base + 5 * NC]);
    __builtin_nontemporal_store(u.ar0, &U[base + 6 * NC]);
    __builtin_nontemporal_store(u.a1, &U[base + 7 * NC]);
    __builtin_nontemporal_store(u.ar1, &U[base + 8 * NC]);
  }
  else
  {
    __builtin_nontemporal_store(u.a0, &U[base + 4 * NC]);
    __builtin_nontemporal_store(u.ar0, &U[base + 5 * NC]);
    __builtin_nontemporal_store(u.a1, &U[base + 6 * NC]);
    __builtin_nontemporal_store(u.ar1, &U[base + 7 * NC]);
  }
}

// wavefront max (butterfly over the 64 lanes; fmax is exact so the order does not matter)
G5D double
wave_max(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

} // namespace g5
