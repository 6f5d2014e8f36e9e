// g5_step3d.hip -- the FP64 (bit-exact) and mixed-precision instances of the fused 3-D step
// (the body is in g5_step3d_body.h; the reassociated FP64 instances are in g5_step3d_fast.hip).
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

template <bool THINC, bool HANCOCK, bool ZSPLIT, bool INLET, bool MP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_step3d_mixed(StepArgs A)
{
  step3d_blocks<THINC, HANCOCK, ZSPLIT, INLET, MP, float>(A);
}

template <bool ZSPLIT, bool INLET, bool MP, typename T>
static void
launch3v3(StepArgs const & A, bool thinc, bool hancock, hipStream_t stream)
{
  const dim3 grid(8 * A.per_xcd), block(256);
  if constexpr (sizeof(T) == sizeof(double))
  {
    if (thinc and hancock)
      hipLaunchKernelGGL((k_step3d_v3<true, true, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
    else if (thinc)
      hipLaunchKernelGGL((k_step3d_v3<true, false, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
    else if (hancock)
      hipLaunchKernelGGL((k_step3d_v3<false, true, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
    else
      hipLaunchKernelGGL((k_step3d_v3<false, false, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
  }
  else
  {
    if (thinc and hancock)
      hipLaunchKernelGGL((k_step3d_mixed<true, true, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
    else if (thinc)
      hipLaunchKernelGGL((k_step3d_mixed<true, false, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
    else if (hancock)
      hipLaunchKernelGGL((k_step3d_mixed<false, true, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
    else
      hipLaunchKernelGGL((k_step3d_mixed<false, false, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
  }
}

template <typename T>
static void
launch3p(StepArgs const & A, bool thinc, bool hancock, bool split, bool inlet, bool mp, hipStream_t stream)
{
  if (mp) // the split z solve (same bits as the other placement)
  {
    if (inlet)
      launch3v3<true, true, true, T>(A, thinc, hancock, stream);
    else
      launch3v3<true, false, true, T>(A, thinc, hancock, stream);
  }
  else if (split)
  {
    if (inlet)
      launch3v3<true, true, false, T>(A, thinc, hancock, stream);
    else
      launch3v3<true, false, false, T>(A, thinc, hancock, stream);
  }
  else
  {
    if (inlet)
      launch3v3<false, true, false, T>(A, thinc, hancock, stream);
    else
      launch3v3<false, false, false, T>(A, thinc, hancock, stream);
  }
}

void
launch_step3d(StepArgs const & A, bool thinc, bool hancock, int precision, bool split, bool inlet, bool mp,
              hipStream_t stream)
{
  if (precision == G5_PREC_MIXED)
    launch3p<float>(A, thinc, hancock, split, inlet, mp, stream);
  else if (precision == G5_PREC_FP64_FAST)
    launch_step3d_fast(A, thinc, hancock, split, inlet, mp, stream);
  else
    launch3p<double>(A, thinc, hancock, split, inlet, mp, stream);
}

} // namespace g5
