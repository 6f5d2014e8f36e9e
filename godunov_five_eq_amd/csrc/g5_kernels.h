// g5_kernels.h -- host-side launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "g5_types.h"

namespace g5
{

// fused step (primitives, slopes, fluxes, update, normalisation, next-dt contribution); mixed
// selects the FP32 primitive/face path
// precision: G5_PREC_* (FP64 bit-exact, mixed, reassociated FP64)
void launch_step2d(StepArgs const & A, bool thinc, bool hancock, int precision, bool ring_wave, bool mp,
                   hipStream_t stream);
void launch_step3d(StepArgs const & A, bool thinc, bool hancock, int precision, int zsplit, bool inlet, bool mp,
                   hipStream_t stream);
// the reassociated FP64 instances (separate translation units)
void launch_step2d_fast(StepArgs const & A, bool thinc, bool hancock, bool ring_wave, bool mp, hipStream_t stream);
void launch_step3d_fast(StepArgs const & A, bool thinc, bool hancock, bool zsplit, bool inlet, bool mp,
                        hipStream_t stream);
// the mixed-precision instances (separate translation units, FMA contraction within expressions)
void launch_step2d_mixed(StepArgs const & A, bool thinc, bool hancock, bool ring_wave, bool mp, hipStream_t stream);
void launch_step3d_mixed(StepArgs const & A, bool thinc, bool hancock, bool zsplit, bool inlet, bool mp,
                         hipStream_t stream);

// the 2-D step over a launch list with blocks of two-level solves (FP64 or reassociated FP64): the
// first n_first blocks (the fine blocks that border coarse blocks) in one launch, the rest of the
// list in a second one
void launch_step2d_two_level(StepArgs const & A, int n_first, bool thinc, bool hancock, int precision, bool mp,
                             hipStream_t stream);
void launch_step2d_two_level_fast(StepArgs const & A, int n_first, bool thinc, bool hancock, bool mp,
                                  hipStream_t stream);
void launch_step3d_two_level(StepArgs const & A, int n_first, bool thinc, bool hancock, int precision,
                             hipStream_t stream);
void launch_step3d_two_level_fast(StepArgs const & A, int n_first, bool thinc, bool hancock, hipStream_t stream);

// refine flags and largest indicator values of the listed blocks into R.flags / R.imax by slot
void launch_refine_flags(int dim, double const * buf0, double const * buf1, BlockDesc const * blocks,
                         SolveDesc const * solves, SolveCtl const * ctl, int32_t const * slots, RefineArgs const & R,
                         int nblocks, hipStream_t stream);

// max over the listed blocks of the CFL rate of the current state into ctl[solve].invdt_bits
// (the rate of the path of `precision`, G5_PREC_*)
void launch_inv_dt(int dim, int precision, double * buf0, double * buf1, BlockDesc const * blocks, SolveDesc const * solves,
                   SolveCtl * ctl, int nblocks, hipStream_t stream);

// dt of the first step of solves ids[0..n) (all when ids is null) from invdt_bits; marks active
void launch_start(SolveCtl * ctl, SolveDesc const * solves, int32_t const * ids, int n, hipStream_t stream);

// after a device-dt step: t += dt, steps++, parity flip, next dt, invdt_bits = 0
void launch_finalize(SolveCtl * ctl, SolveDesc const * solves, int nsolves, hipStream_t stream, bool flip = true);

// after a host-dt step of every solve: parity flip
void launch_flip(SolveCtl * ctl, int nsolves, hipStream_t stream);

// zero the invdt accumulators
void launch_reset_invdt(SolveCtl * ctl, int nsolves, hipStream_t stream);

// copy a solve's blocks between their slots and a contiguous array (to_slots: array -> slots)
void launch_gather(double * buf0, double * buf1, SolveCtl const * ctl, int solve, int32_t const * slots, int base,
                   int nblk, int block_doubles, double * out, bool to_slots, hipStream_t stream);

// gather the outside ring of one solve (FillOutsideCellFunctor) into ring[n_out][nvar][cells]
void launch_fill_outside(int dim, double * buf0, double * buf1, SolveCtl const * ctl, int solve,
                         SolveDesc const * solves, int32_t const * slots, int32_t const * coords, int n_out,
                         double * ring, hipStream_t stream);

// U2 = a*U + b*U2 over n doubles (RK2_SSP LinearCombination)
void launch_lincomb(double a, double const * U, double b, double * U2, size_t n, hipStream_t stream);
void launch_rk2_combine(double * buf0, double * buf1, BlockDesc const * blocks, SolveCtl const * ctl, int nblocks,
                        int block_doubles, hipStream_t stream);

// initial condition of one solve into its slots of the current buffer
void launch_init_ic(int dim, double * buf0, double * buf1, SolveCtl const * ctl, int solve, SolveDesc const * solves,
                    int32_t const * slots, IcDesc const * ic, int nblk, hipStream_t stream);

} // namespace g5
