// The Euler path kernels without recorded increments and antithetic halves
// (the GEN = false instances of k_euler, see euler_kernel.hpp), compiled in a
// translation unit of their own so that both sets build in parallel.
#include "euler_kernel.hpp"

namespace sdegpu {

EulerKernel<float> euler_plain_f32(int model, int m, int d, bool defer) {
  return pick_kernel<false, float>(model, m, d, defer);
}

EulerKernel<double> euler_plain_f64(int model, int m, int d, bool defer) {
  return pick_kernel<false, double>(model, m, d, defer);
}

}  // namespace sdegpu
