// gr_launch_bf16.hip -- sweep_kernel instantiations for float fields with __hip_bfloat16
// coefficients (see gr_launch.hpp).
#define GR_LAUNCH_IMPL 1
#include "gr_launch.hpp"

namespace gr {

hipError_t launch_sweep_bf16(int kid, bool fast, bool rocs, const SweepArgs<float, __hip_bfloat16> &a, const LaunchCfg &lc,
                          hipStream_t st, bool nt) {
  return launch_sweep_k<float, __hip_bfloat16>(kid, fast, rocs, a, lc, st, nt);
}

}  // namespace gr
