// g5_step2d_fast.hip -- the reassociated FP64 instances (G5_PREC_FP64_FAST) of the fused 2-D step
// of g5_step2d_body.h (see g5_step3d_fast.hip for what the reassociation changes).
#include "g5_step2d_body.h"

namespace g5
{

void
launch_step2d_fast(StepArgs const & A, bool thinc, bool hancock, bool ring_wave, bool mp, hipStream_t stream)
{
  mp ? launch2<double, true, true>(A, thinc, hancock, ring_wave, stream)
     : launch2<double, false, true>(A, thinc, hancock, ring_wave, stream);
}

void
launch_step2d_two_level_fast(StepArgs const & A, int n_first, bool thinc, bool hancock, bool mp, hipStream_t stream)
{
  mp ? launch2_two_level<true, true>(A, n_first, thinc, hancock, stream)
     : launch2_two_level<false, true>(A, n_first, thinc, hancock, stream);
}

} // namespace g5
