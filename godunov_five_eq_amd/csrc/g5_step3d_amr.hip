// g5_step3d_amr.hip -- the 3-D step over launch lists with blocks of two-level solves, FP64 in the
// reference's operation order (the body is in g5_step3d_body.h).
#include "g5_step3d_body.h"

namespace g5
{

void
launch_step3d_two_level(StepArgs const & A, int n_first, bool thinc, bool hancock, int precision, hipStream_t stream)
{
  if (precision == G5_PREC_FP64_FAST)
    launch_step3d_two_level_fast(A, n_first, thinc, hancock, stream);
  else
    launch3_two_level<false>(A, n_first, thinc, hancock, stream);
}

} // namespace g5
