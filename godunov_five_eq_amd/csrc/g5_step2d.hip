// g5_step2d.hip -- the FP64 (bit-exact) instances of the fused 2-D step (the body is in
// g5_step2d_body.h; the reassociated FP64 instances are in g5_step2d_fast.hip, the mixed-precision
// ones in g5_step2d_mixed.hip).
#include "g5_step2d_body.h"

namespace g5
{

void
launch_step2d(StepArgs const & A, bool thinc, bool hancock, int precision, bool ring_wave, bool mp, hipStream_t stream)
{
  if (precision == G5_PREC_MIXED)
    launch_step2d_mixed(A, thinc, hancock, ring_wave, mp, stream);
  else if (precision == G5_PREC_FP64_FAST)
    launch_step2d_fast(A, thinc, hancock, ring_wave, mp, stream);
  else
    mp ? launch2<double, true, false>(A, thinc, hancock, ring_wave, stream)
       : launch2<double, false, false>(A, thinc, hancock, ring_wave, stream);
}

void
launch_step2d_two_level(StepArgs const & A, int n_first, bool thinc, bool hancock, int precision, bool mp,
                        hipStream_t stream)
{
  if (precision == G5_PREC_FP64_FAST)
    launch_step2d_two_level_fast(A, n_first, thinc, hancock, mp, stream);
  else
    mp ? launch2_two_level<true, false>(A, n_first, thinc, hancock, stream)
       : launch2_two_level<false, false>(A, n_first, thinc, hancock, stream);
}

} // namespace g5
