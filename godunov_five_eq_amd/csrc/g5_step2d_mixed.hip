// g5_step2d_mixed.hip -- the mixed-precision instances of the fused 2-D step (body in
// g5_step2d_body.h).  Their own translation unit so that the compiler may contract a*b+c into FMA
// (within an expression) here: the mixed path is held to its accuracy bar against the FP64
// oracle, not to the reference's operation order, which the bit-exact FP64 instances of
// g5_step2d.hip keep.  They take the reassociated forms of the device functions (FAST: products by
// reciprocals, the face axis's term alone in the MUSCL extrapolation, minmod as a median).
#include "g5_step2d_body.h"

namespace g5
{

void
launch_step2d_mixed(StepArgs const & A, bool thinc, bool hancock, bool ring_wave, bool mp, hipStream_t stream)
{
  mp ? launch2<float, true, true>(A, thinc, hancock, ring_wave, stream)
     : launch2<float, false, true>(A, thinc, hancock, ring_wave, stream);
}

} // namespace g5
