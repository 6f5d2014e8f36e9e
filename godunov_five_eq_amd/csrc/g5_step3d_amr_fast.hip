// g5_step3d_amr_fast.hip -- the reassociated FP64 instances (G5_PREC_FP64_FAST) of the 3-D step
// over launch lists with blocks of two-level solves.
#include "g5_step3d_body.h"

namespace g5
{

void
launch_step3d_two_level_fast(StepArgs const & A, int n_first, bool thinc, bool hancock, hipStream_t stream)
{
  launch3_two_level<true>(A, n_first, thinc, hancock, stream);
}

} // namespace g5
