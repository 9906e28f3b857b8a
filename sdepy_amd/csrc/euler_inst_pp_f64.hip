// k_euler instances: the general instances with per-path parameters in FP64
// (euler_kernel.hpp).
#include "euler_kernel.hpp"

namespace sdegpu {

SDEGPU_EULER_SET(euler_set_pp_f64, true, double, false) {
  return pick_kernel_mom<true, double, false, true>(model, m, d, defer);
}

}  // namespace sdegpu
