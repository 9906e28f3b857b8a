// k_euler instances: the plain instances (no recorded increments, antithetic halves or groups) in FP64, with fused row moments (euler_kernel.hpp).
#include "euler_kernel.hpp"

namespace sdegpu {

SDEGPU_EULER_SET(euler_set_plain_mom_f64, false, double, true) {
  return pick_kernel_mom<false, double, true>(model, m, d, defer);
}

}  // namespace sdegpu
