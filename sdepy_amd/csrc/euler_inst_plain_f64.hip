// k_euler instances: the plain instances (no recorded increments, antithetic halves or groups) in FP64 (euler_kernel.hpp).
#include "euler_kernel.hpp"

namespace sdegpu {

SDEGPU_EULER_SET(euler_set_plain_f64, false, double, false) {
  return pick_kernel_mom<false, double, false>(model, m, d, defer);
}

}  // namespace sdegpu
