// g5_kernels.h -- host-side launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "g5_types.h"

namespace g5
{

// fused step (primitives, slopes, fluxes, update, normalisation, next-dt contribution)
void launch_step2d(StepArgs const & A, bool thinc, bool hancock, hipStream_t stream);
void launch_step3d(StepArgs const & A, bool thinc, bool hancock, hipStream_t stream);

// max over owned cells of the CFL rate of U (ComputeDtFiveEqFunctor) into ctl[solve].invdt_bits
void launch_inv_dt(int dim, double const * U, BlockDesc const * blocks, SolveDesc const * solves,
                   SolveCtl * ctl, int nblocks, hipStream_t stream);

// dt of the first step from invdt_bits (clamped to tend); marks solves active
void launch_start(SolveCtl * ctl, SolveDesc const * solves, int nsolves, hipStream_t stream);

// after a step: t += dt, steps++, next dt from invdt_bits, invdt_bits = 0
void launch_finalize(SolveCtl * ctl, SolveDesc const * solves, int nsolves, hipStream_t stream);

// zero the invdt accumulators
void launch_reset_invdt(SolveCtl * ctl, int nsolves, hipStream_t stream);

// gather the outside ring (FillOutsideCellFunctor) into ring[n_out][nvar][cells]
void launch_fill_outside(int dim, double const * U, SolveDesc const * solve, int32_t const * coords,
                         int n_out, double * ring, hipStream_t stream);

// U2 = a*U + b*U2 over n doubles (RK2_SSP LinearCombination)
void launch_lincomb(double a, double const * U, double b, double * U2, size_t n, hipStream_t stream);

} // namespace g5
