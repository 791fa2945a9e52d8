// FP64-local instantiation of the 5D HJI stage kernel.  This translation unit
// is compiled with -ffp-contract=off so that every double operation rounds
// exactly as in the reference CPU solver (reversible/native_relative.cpp,
// built by g++ without FMA contraction); fields then match it bit for bit.
#include "hji5d_stage.h"

namespace pe {

void hji5d_launch_f64(const Hji5dGrid &g, const StageArgs<double> &a,
                      bool blend, bool final_stage, hipStream_t stream) {
  const dim3 grid((unsigned)(g.per_xcd * 8)), block(256);
  if (blend)
    hipLaunchKernelGGL((hji5d_stage<double, true, true>), grid, block, 0,
                       stream, a, g, (int)final_stage);
  else
    hipLaunchKernelGGL((hji5d_stage<double, true, false>), grid, block, 0,
                       stream, a, g, (int)final_stage);
}

}  // namespace pe
