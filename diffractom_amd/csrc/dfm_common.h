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

// One pole-figure entry PF[r, kg, c, p] (create_pfo_matrix.py:108-199, pf_generate.cl:62-140 pf_entry): the sum over
// the quadrature nodes (s, se) of the angle and eta bin (weights qw, or the mean of the nodes) of the ring's pole axes'
// images count_a g(|n_{kg a} . v|), g(d) = exp(-(1 - d) / sigma^2) + exp(-(1 + d) / sigma^2), each term dropped
// where its exponent reaches cut, times 1 / (8 pi sigma^2). Directions (R, S, C, Se, P, 3) and poles (K, A, 3) are
// float64 and so is the arithmetic: near a pole 1 - |n.v| is a difference of nearly equal numbers and the kernel
// amplifies its error by 1/sigma^2 (3300 at sigma = 1 deg), so float32 inputs alone would cost ~1e-4 of relative
// accuracy. gfx950 runs float64 at full rate.
__device__ __forceinline__ float pf_entry(const double* __restrict__ coords, const double* __restrict__ poles,
                                          const int* __restrict__ axis_start, const float* __restrict__ axis_count,
                                          const double* __restrict__ inv_sigma2, const double* __restrict__ normf,
                                          const float* __restrict__ qw, int r, int kg, int c, int p, int C, int P,
                                          int A, int S, int Se, double cut, int quad_weighted) {
  const double is2 = inv_sigma2[kg];
  const int a0 = axis_start[p], a1 = axis_start[p + 1];
  const double* pk = poles + (int64_t)kg * A * 3;
  double total_w = 0.0;
  for (int s = 0; s < S; ++s) {
    for (int se = 0; se < Se; ++se) {
      const double* v = coords + (((((int64_t)r * S + s) * C + c) * Se + se) * P + p) * 3;
      const double vx = v[0], vy = v[1], vz = v[2];
      double node = 0.0;
      for (int a = a0; a < a1; ++a) {
        const double d = fabs(pk[a * 3 + 0] * vx + pk[a * 3 + 1] * vy + pk[a * 3 + 2] * vz);
        double term = 0.0;
        const double x1 = (1.0 - d) * is2;
        if (x1 < cut) term += exp(-x1);
        const double x2 = (1.0 + d) * is2;
        if (x2 < cut) term += exp(-x2);
        node += (double)axis_count[a] * term;
      }
      total_w += quad_weighted ? (double)qw[s * Se + se] * node : node;
    }
  }
  if (!quad_weighted) total_w = total_w / (double)(S * Se);
  return (float)(total_w * normf[kg]);
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
