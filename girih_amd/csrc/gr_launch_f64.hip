// gr_launch_f64.hip -- sweep_kernel instantiations for double fields with double
// coefficients (see gr_launch.hpp).
#define GR_LAUNCH_IMPL 1
#include "gr_launch.hpp"

namespace gr {

hipError_t launch_sweep_f64(int kid, bool fast, bool rocs, const SweepArgs<double, double> &a, const LaunchCfg &lc,
                          hipStream_t st, bool nt) {
  return launch_sweep_k<double, double>(kid, fast, rocs, a, lc, st, nt);
}

}  // namespace gr
