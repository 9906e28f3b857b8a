// k_euler instances: the general instances (recorded increments, antithetic halves, component groups) in FP64, with fused row moments (euler_kernel.hpp).
#include "euler_kernel.hpp"

namespace sdegpu {

SDEGPU_EULER_SET(euler_set_gen_mom_f64, true, double, true) {
  return pick_kernel_mom<true, double, true>(model, m, d, defer);
}

}  // namespace sdegpu
