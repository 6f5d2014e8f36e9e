// g5_step3d_mixed.hip -- the mixed-precision instances of the fused 3-D step (body in
// g5_step3d_body.h): FP32 primitives, face states and wave speeds, FP64 state, pressure recovery,
// energy flux and dt.  Their own translation unit so that the compiler may contract a*b+c into FMA
// (within an expression) here: the mixed path is held to its accuracy bar against the FP64
// oracle, not to the reference's operation order, which the bit-exact FP64 instances of
// g5_step3d.hip keep.  They take the reassociated forms of the device functions (FAST: products by
// reciprocals, the face axis's term alone in the MUSCL extrapolation, minmod as a median).
#include "g5_step3d_body.h"

namespace g5
{

// at most 256 registers and 72 KB of LDS per workgroup: two wavefronts per SIMD
template <bool THINC, bool HANCOCK, bool ZSPLIT, bool INLET, bool MP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_step3d_mixed(StepArgs A)
{
  step3d_blocks<THINC, HANCOCK, ZSPLIT, INLET, MP, float, true>(A);
}

template <bool ZSPLIT, bool INLET, bool MP>
static void
launch3m(StepArgs const & A, bool thinc, bool hancock, hipStream_t stream)
{
  const dim3 grid(8 * A.per_xcd), block(256);
  if (thinc and hancock)
    hipLaunchKernelGGL((k_step3d_mixed<true, true, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
  else if (thinc)
    hipLaunchKernelGGL((k_step3d_mixed<true, false, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
  else if (hancock)
    hipLaunchKernelGGL((k_step3d_mixed<false, true, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
  else
    hipLaunchKernelGGL((k_step3d_mixed<false, false, ZSPLIT, INLET, MP>), grid, block, 0, stream, A);
}

void
launch_step3d_mixed(StepArgs const & A, bool thinc, bool hancock, bool split, bool inlet, bool mp, hipStream_t stream)
{
  if (mp) // the split z solve (same results as the other placement)
    inlet ? launch3m<true, true, true>(A, thinc, hancock, stream) : launch3m<true, false, true>(A, thinc, hancock, stream);
  else if (split)
    inlet ? launch3m<true, true, false>(A, thinc, hancock, stream) : launch3m<true, false, false>(A, thinc, hancock, stream);
  else
    inlet ? launch3m<false, true, false>(A, thinc, hancock, stream)
          : launch3m<false, false, false>(A, thinc, hancock, stream);
}

} // namespace g5
