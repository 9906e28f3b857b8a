// k_euler_wide and k_wiener_fill_wide instances in FP32 (euler_wide.hpp).
#include "euler_wide.hpp"

namespace sdegpu {

SDEGPU_EULER_WIDE_SET(float)

}  // namespace sdegpu
