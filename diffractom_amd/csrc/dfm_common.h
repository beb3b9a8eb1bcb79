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

// Streaming (non-temporal) accesses for data a kernel touches once: they do not displace what the XCD's L2 holds for
// reuse (sinogram slabs, image columns, geometry tables).
typedef float f4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_stream(const float4* p) {
  const f4_t v = __builtin_nontemporal_load(reinterpret_cast<const f4_t*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_stream(float4* p, const float4& v) {
  const f4_t t = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(t, reinterpret_cast<f4_t*>(p));
}

// data rows omegas[a] (row_len values each) (+)= sum over ks < ksplit of part[ks][a] (fixed order).
int launch_splitk_reduce(const float* part, float* data, const int* omegas, int n_ang, int64_t row_len, int ksplit,
                         int accumulate, hipStream_t st);

inline unsigned grid_stride_blocks(int64_t n, int block, unsigned cap = 1u << 16) {
  int64_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  if (b > (int64_t)cap) b = cap;
  return (unsigned)b;
}

}  // namespace dfm
