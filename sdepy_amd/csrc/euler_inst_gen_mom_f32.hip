// k_euler instances: the general instances (recorded increments, antithetic halves, component groups) in FP32, with fused row moments (euler_kernel.hpp).
#include "euler_kernel.hpp"

namespace sdegpu {

SDEGPU_EULER_SET(euler_set_gen_mom_f32, true, float, true) {
  return pick_kernel_mom<true, float, true>(model, m, d, defer);
}

}  // namespace sdegpu
