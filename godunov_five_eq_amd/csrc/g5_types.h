// g5_types.h -- descriptors shared by the host runtime and the HIP kernels.
#pragma once

#include <stdint.h>

#include "../../include/godunov_five_eq.h"

namespace g5
{

// One solve (one brick of blocks) as the kernels see it.  Several solves can share one launch:
// their blocks are flattened into one block list (BlockDesc) and each block carries its solve id.
struct SolveDesc
{
  int32_t base;        // index of the solve's first block in the flattened block list
  int32_t nb[3];       // blocks per direction
  int32_t periodic[3]; // periodicity per direction
  int32_t bc[6];       // G5_BC_* per face
  int32_t integrator;  // G5_TI_*
  int32_t pad;
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
  int32_t            parity;     // index (0/1) of the buffer holding the latest state
};

// block descriptor: solve id and brick coordinates
struct BlockDesc
{
  int32_t solve;
  int32_t bx, by, bz;
};

struct StepArgs
{
  const double *    U;   // state at time n (all blocks)
  double *          U2;  // state at time n+1
  const BlockDesc * blocks;
  const SolveDesc * solves;
  SolveCtl *        ctl;
  int32_t           nblocks;
  int32_t           per_xcd; // workgroups per XCD chunk (grid = 8 * per_xcd)
  double            dt_fixed; // > 0: use this dt for every solve (host-driven step)
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
  s.pad = 0;
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
  }
  s.thinc_beta = c.thinc_beta;
  s.thinc_eps = c.thinc_epsilon;
}

} // namespace g5
