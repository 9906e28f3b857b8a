// Shared host-side plumbing of libsdegpu: status codes, the per-thread error
// message, HIP error checks and launch-shape helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/sdegpu.h"

namespace sdegpu {

// Message of the last failure on this thread (read by sdegpu_last_error).
void set_error(const char* fmt, ...);

// Returns SDEGPU_OK or records the HIP error and returns SDEGPU_ERUNTIME.
int hip_status(hipError_t e, const char* what);

// Checks the launch that was just issued.
inline int launch_status(const char* kernel) { return hip_status(hipGetLastError(), kernel); }

inline int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  set_error("%s", buf);
  return code;
}

constexpr int kBlock = 256;  // threads per block of the per-path kernels (4 waves)

inline unsigned blocks_for(int64_t n, int per_block = kBlock) {
  return (unsigned)((n + per_block - 1) / per_block);
}

// Largest path count a kernel accepts: the path index is the 32-bit c1 word
// of the Philox counter and grids are limited to 2^31-1 blocks.
constexpr int64_t kMaxPaths = 0xFFFFFFFFll;

}  // namespace sdegpu
