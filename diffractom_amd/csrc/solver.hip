// Element-wise kernels and float64 reductions of the SPDHG-Poisson solver on gfx950.
#include "dfm_common.h"

namespace dfm {

__device__ __forceinline__ float pval(const float* __restrict__ a, int64_t p, int64_t i, float dflt) {
  return p ? a[i % p] : dflt;
}

// y' = prox_{s F*}(y + s Ax) on the rows of a subset, float64 arithmetic; d: A x in, y' - y out.
__global__ void __launch_bounds__(256) dual_update_sub_kernel(float* __restrict__ y, float* __restrict__ d,
                                                              const float* __restrict__ sig, const float* __restrict__ b,
                                                              const float* __restrict__ w, int64_t pw,
                                                              const float* __restrict__ g, int64_t pg,
                                                              const float* __restrict__ bg, int64_t pbg, float eps,
                                                              const int* __restrict__ omegas, int64_t row_len,
                                                              int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = t / row_len;
    const int64_t i = (int64_t)omegas[s] * row_len + (t - s * row_len);
    const float wi = pval(w, pw, i, 1.0f);
    const float si = sig[i];
    const float yi = y[i];
    float yn = 0.0f;
    if (wi != 0.0f && si != 0.0f) {
      const float gi = pval(g, pg, i, 1.0f);
      const double sc = (double)wi / (double)gi;
      const double r = (double)pval(bg, pbg, i, 0.0f) + (double)eps * gi;
      const double v = (double)yi + (double)si * (double)d[i] + (double)si * r;
      const double dd = v - sc;
      yn = (float)(0.5 * (v + sc - sqrt(dd * dd + 4.0 * (double)si * sc * (double)b[i])));
    }
    d[i] = yn - yi;
    y[i] = yn;
  }
}

__global__ void __launch_bounds__(256) spdhg_update_kernel(float* __restrict__ x, float* __restrict__ z,
                                                           const float* __restrict__ T, const float* __restrict__ dz,
                                                           float inv_p, float lam, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float zi = z[i] + dz[i];
    z[i] = zi;
    const float zb = zi + inv_p * dz[i];
    const float t = T[i];
    x[i] = t > 0.0f ? fmaxf(0.0f, x[i] - t * (zb + lam)) : 0.0f;
  }
}

// (pixel, channel) kernels over channel-fastest (npix, kstride) arrays: blockIdx.y runs over pixels.
__global__ void __launch_bounds__(256) tinit_kernel(float* __restrict__ T, int64_t npix, int kstride) {
  for (int64_t p = blockIdx.y; p < npix; p += gridDim.y)
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kstride; k += gridDim.x * blockDim.x)
      T[p * kstride + k] = INFINITY;
}

__global__ void __launch_bounds__(256) tmin_kernel(float* __restrict__ T, const float* __restrict__ colsum,
                                                   const unsigned char* __restrict__ supp, int64_t npix, int kstride,
                                                   int kvalid, float c) {
  for (int64_t p = blockIdx.y; p < npix; p += gridDim.y) {
    const bool inside = supp ? supp[p] != 0 : true;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kvalid; k += gridDim.x * blockDim.x) {
      const int64_t i = p * kstride + k;
      const float cs = colsum[i];
      if (inside && cs > 0.0f) T[i] = fminf(T[i], c / cs);
    }
  }
}

__global__ void __launch_bounds__(256) tfinal_kernel(float* __restrict__ T, int64_t npix, int kstride, int kvalid) {
  for (int64_t p = blockIdx.y; p < npix; p += gridDim.y)
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kstride; k += gridDim.x * blockDim.x) {
      const int64_t i = p * kstride + k;
      if (k >= kvalid || isinf(T[i])) T[i] = 0.0f;
    }
}

__global__ void __launch_bounds__(256) support_ones_kernel(float* __restrict__ x, const unsigned char* __restrict__ supp,
                                                           int64_t npix, int kstride, int kvalid) {
  for (int64_t p = blockIdx.y; p < npix; p += gridDim.y) {
    const float in = (supp ? supp[p] != 0 : true) ? 1.0f : 0.0f;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kstride; k += gridDim.x * blockDim.x)
      x[p * kstride + k] = k < kvalid ? in : 0.0f;
  }
}

__global__ void __launch_bounds__(256) project_support_kernel(float* __restrict__ x,
                                                              const unsigned char* __restrict__ supp, int64_t npix,
                                                              int kstride, int kvalid) {
  for (int64_t p = blockIdx.y; p < npix; p += gridDim.y) {
    const bool inside = supp ? supp[p] != 0 : true;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kstride; k += gridDim.x * blockDim.x) {
      const int64_t i = p * kstride + k;
      if (!inside || k >= kvalid || x[i] < 0.0f) x[i] = 0.0f;
    }
  }
}

__global__ void __launch_bounds__(256) dual_steps_kernel(float* __restrict__ sig, const float* __restrict__ rowsum,
                                                         const float* __restrict__ w, int64_t pw, float sigma0,
                                                         int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float rs = rowsum[i];
    sig[i] = (pval(w, pw, i, 1.0f) != 0.0f && rs > 0.0f) ? sigma0 / rs : 0.0f;
  }
}

__global__ void __launch_bounds__(256) set_mask_ones_kernel(float* __restrict__ u, const float* __restrict__ w,
                                                            int64_t pw, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    u[i] = pval(w, pw, i, 1.0f) != 0.0f ? 1.0f : 0.0f;
}

__global__ void __launch_bounds__(256) scale_kernel(float* __restrict__ a, float s, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] *= s;
}

// Block reduction of NV doubles (256 threads: wave shuffles, then LDS across the 4 waves).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* __restrict__ out, int stride) {
  __shared__ double red[NV][4];
#pragma unroll
  for (int q = 0; q < NV; ++q)
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) red[q][wv] = v[q];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q * stride + blockIdx.x] = red[q][0] + red[q][1] + red[q][2] + red[q][3];
}

__global__ void __launch_bounds__(256) deviance_kernel(const float* __restrict__ mu, const float* __restrict__ b,
                                                       const float* __restrict__ w, int64_t pw,
                                                       const float* __restrict__ g, int64_t pg,
                                                       const float* __restrict__ bg, int64_t pbg, float eps, int64_t n,
                                                       double* __restrict__ partial) {
  double s[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float wi = pval(w, pw, i, 1.0f);
    if (wi != 0.0f) {
      const float gi = pval(g, pg, i, 1.0f);
      const double bi = (double)b[i];
      const double sc = (double)wi / (double)gi;
      const double m = (double)fmaxf(mu[i], 0.0f) + (double)pval(bg, pbg, i, 0.0f) + (double)eps * gi;
      double v = m - bi;
      if (bi > 0.0) v += bi * log(bi / m);
      s[0] += sc * v;
    }
  }
  block_sum<1>(s, partial, gridDim.x);
}

__global__ void __launch_bounds__(256) mean_counted_kernel(const float* __restrict__ b, const float* __restrict__ w,
                                                           int64_t pw, int64_t n, double* __restrict__ partial) {
  double s[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (pval(w, pw, i, 1.0f) != 0.0f) {
      s[0] += (double)b[i];
      s[1] += 1.0;
    }
  }
  block_sum<2>(s, partial, gridDim.x);
}

__global__ void __launch_bounds__(256) dot_kernel(const float* __restrict__ a, const float* __restrict__ b, int64_t n,
                                                  double* __restrict__ partial) {
  double s[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s[0] += (double)a[i] * (double)b[i];
  block_sum<1>(s, partial, gridDim.x);
}

static unsigned reduce_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > DFM_MAX_PARTIALS) b = DFM_MAX_PARTIALS;
  return (unsigned)b;
}

static dim3 pix_grid(int64_t npix, int kstride) {
  int64_t gy = npix < 32768 ? npix : 32768;
  return dim3((unsigned)((kstride + 255) / 256), (unsigned)gy);
}

}  // namespace dfm

using namespace dfm;

extern "C" {

int dfm_dual_update_sub(float* y, float* d, const float* sig, const float* b, const float* w, int64_t pw,
                        const float* g, int64_t pg, const float* bg, int64_t pbg, float eps, const int32_t* omegas,
                        int32_t n_sub, int64_t row_len, dfm_stream_t stream) {
  DFM_CHECK_ARG(y && d && sig && b && omegas, "dfm_dual_update_sub: null pointer");
  DFM_CHECK_ARG((pw == 0 || w) && (pg == 0 || g) && (pbg == 0 || bg) && pw >= 0 && pg >= 0 && pbg >= 0,
                "dfm_dual_update_sub: periodic array missing");
  DFM_CHECK_ARG(n_sub >= 0 && row_len > 0, "dfm_dual_update_sub: bad sizes");
  const int64_t n = (int64_t)n_sub * row_len;
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(dual_update_sub_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, y, d,
                     sig, b, w, pw, g, pg, bg, pbg, eps, omegas, row_len, n);
  DFM_LAUNCH_CHECK("dfm_dual_update_sub");
  return DFM_OK;
}

int dfm_spdhg_update(float* x, float* z, const float* T, const float* dz, float inv_p, float lam, int64_t n,
                     dfm_stream_t stream) {
  DFM_CHECK_ARG(x && z && T && dz && n >= 0, "dfm_spdhg_update: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(spdhg_update_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, x, z, T,
                     dz, inv_p, lam, n);
  DFM_LAUNCH_CHECK("dfm_spdhg_update");
  return DFM_OK;
}

int dfm_tinit(float* T, int64_t npix, int32_t kstride, dfm_stream_t stream) {
  DFM_CHECK_ARG(T && npix > 0 && npix < (1 << 30) && kstride > 0, "dfm_tinit: bad arguments");
  hipLaunchKernelGGL(tinit_kernel, pix_grid(npix, kstride), dim3(256), 0, (hipStream_t)stream, T, npix, kstride);
  DFM_LAUNCH_CHECK("dfm_tinit");
  return DFM_OK;
}

int dfm_tmin(float* T, const float* colsum, const uint8_t* supp, int64_t npix, int32_t kstride, int32_t kvalid,
             float c, dfm_stream_t stream) {
  DFM_CHECK_ARG(T && colsum && npix > 0 && npix < (1 << 30) && kstride > 0 && kvalid <= kstride,
                "dfm_tmin: bad arguments");
  hipLaunchKernelGGL(tmin_kernel, pix_grid(npix, kstride), dim3(256), 0, (hipStream_t)stream, T, colsum, supp, npix,
                     kstride, kvalid, c);
  DFM_LAUNCH_CHECK("dfm_tmin");
  return DFM_OK;
}

int dfm_tfinal(float* T, int64_t npix, int32_t kstride, int32_t kvalid, dfm_stream_t stream) {
  DFM_CHECK_ARG(T && npix > 0 && npix < (1 << 30) && kstride > 0 && kvalid <= kstride, "dfm_tfinal: bad arguments");
  hipLaunchKernelGGL(tfinal_kernel, pix_grid(npix, kstride), dim3(256), 0, (hipStream_t)stream, T, npix, kstride, kvalid);
  DFM_LAUNCH_CHECK("dfm_tfinal");
  return DFM_OK;
}

int dfm_dual_steps(float* sig, const float* rowsum, const float* w, int64_t pw, float sigma0, int64_t n,
                   dfm_stream_t stream) {
  DFM_CHECK_ARG(sig && rowsum && (pw == 0 || w) && n >= 0, "dfm_dual_steps: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(dual_steps_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, sig,
                     rowsum, w, pw, sigma0, n);
  DFM_LAUNCH_CHECK("dfm_dual_steps");
  return DFM_OK;
}

int dfm_set_mask_ones(float* u, const float* w, int64_t pw, int64_t n, dfm_stream_t stream) {
  DFM_CHECK_ARG(u && (pw == 0 || w) && n >= 0, "dfm_set_mask_ones: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(set_mask_ones_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, u, w,
                     pw, n);
  DFM_LAUNCH_CHECK("dfm_set_mask_ones");
  return DFM_OK;
}

int dfm_support_ones_k(float* x, const uint8_t* supp, int64_t npix, int32_t kstride, int32_t kvalid,
                       dfm_stream_t stream) {
  DFM_CHECK_ARG(x && npix > 0 && npix < (1 << 30) && kstride > 0 && kvalid <= kstride,
                "dfm_support_ones_k: bad arguments");
  hipLaunchKernelGGL(support_ones_kernel, pix_grid(npix, kstride), dim3(256), 0, (hipStream_t)stream, x, supp, npix,
                     kstride, kvalid);
  DFM_LAUNCH_CHECK("dfm_support_ones_k");
  return DFM_OK;
}

int dfm_project_support_k(float* x, const uint8_t* supp, int64_t npix, int32_t kstride, int32_t kvalid,
                          dfm_stream_t stream) {
  DFM_CHECK_ARG(x && npix > 0 && npix < (1 << 30) && kstride > 0 && kvalid <= kstride,
                "dfm_project_support_k: bad arguments");
  hipLaunchKernelGGL(project_support_kernel, pix_grid(npix, kstride), dim3(256), 0, (hipStream_t)stream, x, supp,
                     npix, kstride, kvalid);
  DFM_LAUNCH_CHECK("dfm_project_support_k");
  return DFM_OK;
}

int dfm_deviance_partial(const float* mu, const float* b, const float* w, int64_t pw, const float* g, int64_t pg,
                         const float* bg, int64_t pbg, float eps, int64_t n, double* partial, int32_t* n_partial,
                         dfm_stream_t stream) {
  DFM_CHECK_ARG(mu && b && partial && n_partial && (pw == 0 || w) && (pg == 0 || g) && (pbg == 0 || bg) && n >= 0,
                "dfm_deviance_partial: bad arguments");
  const unsigned nb = reduce_blocks(n);
  *n_partial = (int32_t)nb;
  hipLaunchKernelGGL(deviance_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, mu, b, w, pw, g, pg, bg, pbg, eps,
                     n, partial);
  DFM_LAUNCH_CHECK("dfm_deviance_partial");
  return DFM_OK;
}

int dfm_mean_counted_partial(const float* b, const float* w, int64_t pw, int64_t n, double* partial,
                             int32_t* n_partial, dfm_stream_t stream) {
  DFM_CHECK_ARG(b && partial && n_partial && (pw == 0 || w) && n >= 0, "dfm_mean_counted_partial: bad arguments");
  const unsigned nb = reduce_blocks(n);
  *n_partial = (int32_t)nb;
  hipLaunchKernelGGL(mean_counted_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, b, w, pw, n, partial);
  DFM_LAUNCH_CHECK("dfm_mean_counted_partial");
  return DFM_OK;
}

int dfm_dot_partial(const float* a, const float* b, int64_t n, double* partial, int32_t* n_partial,
                    dfm_stream_t stream) {
  DFM_CHECK_ARG(a && b && partial && n_partial && n >= 0, "dfm_dot_partial: bad arguments");
  const unsigned nb = reduce_blocks(n);
  *n_partial = (int32_t)nb;
  hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, a, b, n, partial);
  DFM_LAUNCH_CHECK("dfm_dot_partial");
  return DFM_OK;
}

int dfm_scale(float* a, float s, int64_t n, dfm_stream_t stream) {
  DFM_CHECK_ARG(a && n >= 0, "dfm_scale: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(scale_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, a, s, n);
  DFM_LAUNCH_CHECK("dfm_scale");
  return DFM_OK;
}

}  // extern "C"
