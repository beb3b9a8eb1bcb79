// Shared definitions of the diffractom MI355X kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/diffractom_hip.h"

namespace dfm {

// Thread-local last error, reported through dfm_last_error().
void set_error(const char* fmt, ...);

#define DFM_CHECK_ARG(cond, ...)                 \
  do {                                           \
    if (!(cond)) {                               \
      ::dfm::set_error(__VA_ARGS__);             \
      return DFM_ERR_ARG;                        \
    }                                            \
  } while (0)

#define DFM_LAUNCH_CHECK(name)                                                   \
  do {                                                                           \
    hipError_t _e = hipGetLastError();                                           \
    if (_e != hipSuccess) {                                                      \
      ::dfm::set_error("%s: launch failed: %s", name, hipGetErrorString(_e));    \
      return DFM_ERR_LAUNCH;                                                     \
    }                                                                            \
  } while (0)

inline unsigned grid_stride_blocks(int64_t n, int block, unsigned cap = 1u << 16) {
  int64_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  if (b > (int64_t)cap) b = cap;
  return (unsigned)b;
}

}  // namespace dfm
