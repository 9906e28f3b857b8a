// k_euler instances: the plain instances (no recorded increments, antithetic halves or groups) in FP32 (euler_kernel.hpp).
#include "euler_kernel.hpp"

namespace sdegpu {

SDEGPU_EULER_SET(euler_set_plain_f32, false, float, false) {
  return pick_kernel_mom<false, float, false>(model, m, d, defer);
}

}  // namespace sdegpu
