// k_euler instances: the general instances (recorded increments, antithetic halves, component groups) in FP64 (euler_kernel.hpp).
#include "euler_kernel.hpp"

namespace sdegpu {

SDEGPU_EULER_SET(euler_set_gen_f64, true, double, false) {
  return pick_kernel_mom<true, double, false>(model, m, d, defer);
}

}  // namespace sdegpu
