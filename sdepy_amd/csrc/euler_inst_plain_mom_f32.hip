// k_euler instances: the plain instances (no recorded increments, antithetic halves or groups) in FP32, with fused row moments (euler_kernel.hpp).
#include "euler_kernel.hpp"

namespace sdegpu {

SDEGPU_EULER_SET(euler_set_plain_mom_f32, false, float, true) {
  return pick_kernel_mom<false, float, true>(model, m, d, defer);
}

}  // namespace sdegpu
