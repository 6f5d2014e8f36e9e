// g5_step3d.hip -- the FP64 (bit-exact) instances of the fused 3-D step (the body is in
// g5_step3d_body.h; the reassociated FP64 instances are in g5_step3d_fast.hip, the mixed-precision
// ones in g5_step3d_mixed.hip).
#include "g5_step3d_body.h"

namespace g5
{

// INLET: the launch list holds a solve with an inflow border (the resolver then marks inflow
// cells, whose state the conversion replaces); without it the border code is the WALL /
// ZERO_GRADIENT / periodic resolution alone.  MP: a solve with the MILLER_PUCKETT source term
// (instantiated with the split z solve only).
template <bool THINC, bool HANCOCK, bool ZSPLIT, bool INLET, bool MP>
__global__ void __launch_bounds__(256)
k_step3d_v3(StepArgs A)
{
  step3d_blocks<THINC, HANCOCK, ZSPLIT, INLET, MP, double>(A);
}

template <bool ZSPLIT, bool INLET, bool MP>
static void
launch3v3(StepArgs const & A, bool thinc, bool hancock, hipStream_t stream)
{
  const dim3 grid(8 * A.per_xcd), block(256);
  if (thinc and hancock)
    hipLaunchKernelGGL((k_step3d_v3<true, true, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
  else if (thinc)
    hipLaunchKernelGGL((k_step3d_v3<true, false, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
  else if (hancock)
    hipLaunchKernelGGL((k_step3d_v3<false, true, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
  else
    hipLaunchKernelGGL((k_step3d_v3<false, false, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
}

void
launch_step3d(StepArgs const & A, bool thinc, bool hancock, int precision, int zsplit, bool inlet, bool mp,
              hipStream_t stream)
{
  const bool split = zsplit != 0;
  if (precision == G5_PREC_MIXED)
    launch_step3d_mixed(A, thinc, hancock, split, inlet, mp, stream);
  else if (precision == G5_PREC_FP64_FAST)
    launch_step3d_fast(A, thinc, hancock, split, inlet, mp, stream);
  else if (mp) // the split z solve (same bits as the other placement)
    inlet ? launch3v3<true, true, true>(A, thinc, hancock, stream) : launch3v3<true, false, true>(A, thinc, hancock, stream);
  else if (split)
    inlet ? launch3v3<true, true, false>(A, thinc, hancock, stream) : launch3v3<true, false, false>(A, thinc, hancock, stream);
  else
    inlet ? launch3v3<false, true, false>(A, thinc, hancock, stream)
          : launch3v3<false, false, false>(A, thinc, hancock, stream);
}

} // namespace g5
