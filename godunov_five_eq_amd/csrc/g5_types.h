// g5_types.h -- descriptors shared by the host runtime and the HIP kernels.
#pragma once

#include <stdint.h>

#include <cmath>

#include "../../include/godunov_five_eq.h"

namespace g5
{

// One solve (one brick of blocks) as the kernels see it.  Many solves share one launch: their
// blocks are flattened into one block list (BlockDesc) and each block carries its solve id.  The
// state of block (x,y,z) of a solve lives in memory slot slots[base + x + nb0*(y + nb1*z)], so
// solves can be admitted and released without moving other solves' data.
struct SolveDesc
{
  int32_t base;        // index of the solve's first entry in the slot table
  int32_t nb[3];       // blocks per direction
  int32_t periodic[3]; // periodicity per direction
  int32_t bc[6];       // G5_BC_* per face
  int32_t integrator;  // G5_TI_*
  int32_t inlet;       // G5_INLET_*
  int32_t source_term; // G5_SOURCE_*
  int32_t abort_neg_p; // 1: a negative pressure in a new state fails the solve
  // two-level static mesh (2-D): slot tables of level 0 at base (nb[0] x nb[1], -1 where refined)
  // and of level 1 at base1 (2 nb[0] x 2 nb[1], -1 outside the refined blocks); base1 < 0: one level
  int32_t base1;
  int32_t prolong; // coarse-to-fine ghost fill: 0 SIMPLE_COPY, 1 EXTRAPOLATE_LINEAR_MINMOD,
                   // 2 / 3 CONSERVATIVE_INTERPOLATION_ORDER_2 / _4
  double  dx;
  double  cfl;
  double  slope_type;
  double  smallr;
  double  smallc;
  double  smallp;     // smallc^2 (models/HydroSettings.cpp:24)
  double  g1[2];      // 1/(gamma_k - 1)
  double  gp[2];      // gamma_k pinf_k/(gamma_k - 1)
  double  thinc_beta;
  double  thinc_eps;
  double  inv_dx_p2;  // 1/dx when dx is a power of two (x/dx == x*inv_dx_p2 exactly), else 0
  double  inlet_u[9]; // conservative inflow state of the overridden outside cells
  double  origin[3];  // domain origin
  double  jet[3];     // jet centre y, z and width
  double  gamma[2];   // gamma_k and p_inf_k (bulk moduli of the Miller-Puckett source term)
  double  pinf[2];
  // FP32 copies for the mixed-precision path
  float g1f[2];
  float gpf[2];
  float smallrf;
  float smallcf;
  float smallpf;
  float thinc_betaf;
  float thinc_epsf;
  float padf;
};

// per-solve time-stepping control, resident on the device
struct SolveCtl
{
  unsigned long long invdt_bits; // running max of 1/dt of the next state (bit pattern of a double >= 0)
  double             dt;         // dt of the step about to run
  double             t;          // current time
  double             tend;       // end time
  int32_t            active;     // 1 while t < tend and steps < max_steps
  int32_t            steps;      // steps done
  int32_t            max_steps;  // step budget
  int32_t            parity;     // which of the two state buffers holds the current state
  uint32_t           failure;    // G5_FAIL_* bits, set by the kernels that check a new state
  int32_t            pad;
};

// block of the launch list: solve id, memory slot, brick coordinates at the block's level, level
// (0, or 1 for the fine blocks of a two-level mesh)
struct BlockDesc
{
  int32_t solve;
  int32_t slot;
  int16_t bx, by, bz, level;
};

struct StepArgs
{
  double *          buf0;    // state buffer 0
  double *          buf1;    // state buffer 1
  const double *    U_in;    // non-null: read the step input here instead of buf[parity]
  double *          U_out;   // non-null: write the step output here instead of buf[1 - parity]
  const BlockDesc * blocks;
  const SolveDesc * solves;
  SolveCtl *        ctl;
  const int32_t *   slots;   // slot table
  int32_t           nblocks;
  int32_t           per_xcd;    // blocks per XCD chunk, ceil(nblocks / 8)
  int32_t           wg_per_xcd; // 0: one workgroup per block (grid 8 * per_xcd); > 0: persistent
                                // workgroups per XCD, each looping over its XCD's chunk
  int32_t           remap;      // 1: XCD x works on blocks [x * per_xcd, (x + 1) * per_xcd)
  double            dt_fixed;   // > 0: use this dt for every solve (host-driven step)
  // two-level meshes: face records of the fine blocks' faces that border coarse blocks, per slot
  // [side xmin, xmax, ymin, ymax][face 0..15][9 values], written by the fine blocks' launch and
  // read by their coarse neighbours in the next launch
  double *          iflux;
  // 3-D z-columns: the list holds columns of col_stride consecutive z blocks of one solve (padded
  // with solve = -1 entries), and work item w (nblocks of them) marches col_len of them: column
  // w / (col_stride / col_len), part w % (col_stride / col_len); a plain list has both 1
  int32_t           col_stride;
  int32_t           col_len;
};
constexpr int kIfluxDoubles = 4 * 16 * 8; // interface records per slot (2-D)
// 3-D: per coarse slot, [side xmin, xmax, ymin, ymax, zmin, zmax][face cell 0..255][fine face 0..3]
// [ar0, ar1, a0, a1, e, mu, mv, mw, us]
constexpr int kIflux3Doubles = 6 * 256 * 4 * 9;
// 3-D z-columns: at most this many blocks per column of the launch list (StepArgs.col_stride)
constexpr int kMaxColumn = 8;

// refine-flag computation (g5_refine_flags): parameters and the per-slot outputs
struct RefineArgs
{
  int32_t   indicator; // G5_INDICATOR_*
  int32_t   level_min, level_max;
  int32_t   base_level; // level of the solve's level-0 blocks
  double    refine_th, coarsen_th, eps;
  int32_t * flags; // per slot, max-accumulated (G5_AMR_*)
  double *  imax;  // per slot: largest indicator value (bit patterns max-accumulated; starts at 0)
};

// launch geometry of the step kernels (chosen by the autotuner or set by the caller); every
// choice computes the same bits
struct LaunchCfg
{
  int32_t wg_per_cu; // 0: one workgroup per block; > 0: persistent, this many workgroups per CU
  int32_t remap;     // XCD-contiguous block ranges
  int32_t zsplit;    // 3-D: z Riemann solve split around the x/y exchange
  int32_t ring_wave; // 2-D, one workgroup per block: a fifth wavefront for the ring work
  int32_t col_len;   // 3-D: blocks marched in z by one workgroup (1, 2, 4 or 8)
};

// filling the host-side SolveDesc from a g5_config
inline void
make_solve_desc(g5_config const & c, int32_t base, SolveDesc & s)
{
  s.base = base;
  for (int d = 0; d < 3; ++d)
  {
    s.nb[d] = d < c.dim ? c.nb[d] : 1;
    s.periodic[d] = d < c.dim ? c.periodic[d] : 0;
  }
  for (int f = 0; f < 6; ++f)
    s.bc[f] = c.bc[f];
  s.integrator = c.integrator;
  s.inlet = c.inlet;
  s.source_term = c.source_term;
  s.abort_neg_p = c.abort_negative_pressure != 0 ? 1 : 0;
  s.base1 = -1;
  s.prolong = 0;
  for (int k = 0; k < 9; ++k)
    s.inlet_u[k] = c.inlet_u[k];
  for (int d = 0; d < 3; ++d)
  {
    s.origin[d] = c.xmin[d];
    s.jet[d] = c.jet[d];
  }
  s.dx = c.dx;
  s.cfl = c.cfl;
  s.slope_type = c.slope_type;
  s.smallr = c.smallr;
  s.smallc = c.smallc;
  s.smallp = c.smallc * c.smallc;
  for (int k = 0; k < 2; ++k)
  {
    s.g1[k] = 1.0 / (c.gamma[k] - 1.0);
    s.gp[k] = c.gamma[k] * c.pinf[k] / (c.gamma[k] - 1.0);
    s.gamma[k] = c.gamma[k];
    s.pinf[k] = c.pinf[k];
  }
  s.thinc_beta = c.thinc_beta;
  s.thinc_eps = c.thinc_epsilon;
  int          e2 = 0;
  const double m2 = std::frexp(c.dx, &e2); // dx = m2 * 2^e2, m2 in [0.5, 1)
  s.inv_dx_p2 = (m2 == 0.5 and e2 > -1000 and e2 < 1000) ? std::ldexp(1.0, 1 - e2) : 0.0;
  for (int k = 0; k < 2; ++k)
  {
    s.g1f[k] = static_cast<float>(s.g1[k]);
    s.gpf[k] = static_cast<float>(s.gp[k]);
  }
  s.smallrf = static_cast<float>(s.smallr);
  s.smallcf = static_cast<float>(s.smallc);
  s.smallpf = static_cast<float>(s.smallp);
  s.thinc_betaf = static_cast<float>(s.thinc_beta);
  s.thinc_epsf = static_cast<float>(s.thinc_eps);
  s.padf = 0.0f;
}

// Initial-condition description for the device generator (InitShockBubble / InitTriplePoint):
// kind 0: region 0 for x <= x_front else region 1, with a region-2 sphere (circle) inclusion;
// kind 1: same with an axis-aligned box (rectangle) inclusion;
// kind 2: triple point, region 0 for x <= x_front, else region 1 where y <= yd
//         (sqrt(y^2+z^2) <= yd in 3-D) and region 2 elsewhere.
struct IcDesc
{
  int32_t kind;
  int32_t pad;
  double  x_front;
  double  yd;
  double  center[3];
  double  radius;
  double  half[3];
  double  xmin[3];
  double  S[3][9];  // conservative region states
  double  eint_pre; // InitShockBubble.cpp:183-207 specific energies of regions 1 and 2
  double  eint_bub;
};

} // namespace g5
