// k_euler instances: the general instances (recorded increments, antithetic halves, component groups) in FP32 (euler_kernel.hpp).
#include "euler_kernel.hpp"

namespace sdegpu {

SDEGPU_EULER_SET(euler_set_gen_f32, true, float, false) {
  return pick_kernel_mom<true, float, false>(model, m, d, defer);
}

}  // namespace sdegpu
