// k_euler instances: the general instances with per-path parameters in FP32
// (euler_kernel.hpp).
#include "euler_kernel.hpp"

namespace sdegpu {

SDEGPU_EULER_SET(euler_set_pp_f32, true, float, false) {
  return pick_kernel_mom<true, float, false, true>(model, m, d, defer);
}

}  // namespace sdegpu
