// g5_step3d_fast.hip -- the reassociated FP64 instances (G5_PREC_FP64_FAST) of the fused 3-D
// step of g5_step3d_body.h: FP64 throughout, divisions by repeated divisors as products by one
// refined reciprocal, refined hardware square roots, no e(p) -> p(e) round trip in HLLC.  Its own
// translation unit so that it compiles beside the bit-exact and mixed instances.
#include "g5_step3d_body.h"

namespace g5
{

template <bool THINC, bool HANCOCK, bool ZSPLIT, bool INLET, bool MP>
__global__ void __launch_bounds__(256)
k_step3d_fast(StepArgs A)
{
  step3d_blocks<THINC, HANCOCK, ZSPLIT, INLET, MP, double, true>(A);
}

template <bool ZSPLIT, bool INLET, bool MP>
static void
launch3f(StepArgs const & A, bool thinc, bool hancock, hipStream_t stream)
{
  const dim3 grid(8 * A.per_xcd), block(256);
  if (thinc and hancock)
    hipLaunchKernelGGL((k_step3d_fast<true, true, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
  else if (thinc)
    hipLaunchKernelGGL((k_step3d_fast<true, false, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
  else if (hancock)
    hipLaunchKernelGGL((k_step3d_fast<false, true, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
  else
    hipLaunchKernelGGL((k_step3d_fast<false, false, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
}

void
launch_step3d_fast(StepArgs const & A, bool thinc, bool hancock, bool split, bool inlet, bool mp, hipStream_t stream)
{
  if (mp) // the split z solve (same results as the other placement)
    inlet ? launch3f<true, true, true>(A, thinc, hancock, stream) : launch3f<true, false, true>(A, thinc, hancock, stream);
  else if (split)
    inlet ? launch3f<true, true, false>(A, thinc, hancock, stream)
          : launch3f<true, false, false>(A, thinc, hancock, stream);
  else
    inlet ? launch3f<false, true, false>(A, thinc, hancock, stream)
          : launch3f<false, false, false>(A, thinc, hancock, stream);
}

} // namespace g5
