// gr_launch_f32.hip -- sweep_kernel instantiations for float fields with float
// coefficients (see gr_launch.hpp).
#define GR_LAUNCH_IMPL 1
#include "gr_launch.hpp"

namespace gr {

hipError_t launch_sweep_f32(int kid, bool fast, bool rocs, const SweepArgs<float, float> &a, const LaunchCfg &lc,
                          hipStream_t st, bool nt) {
  return launch_sweep_k<float, float>(kid, fast, rocs, a, lc, st, nt);
}

}  // namespace gr
