// The Fourier-slice projector's device steps (operators/fourier_radon.py:1-311, fourier_radon.cl), in the
// operator's orientation-fastest layout: every array is (points, channels) with the channels fastest, so each step
// is coalesced over the channels and the FFTs (a library call between the steps) run over the leading dimensions.
//
//   forward:  pad (pixels / psi into the oversampled grid) -> 2-D FFT -> interp (the polar samples: W x W taps of
//             the separable "exponential of semicircle" kernel, times the bin response, phase and scale) -> inverse
//             real FFT along the detector -> the first Ns bins
//   backward: the transpose of every step: real FFT of the zero-padded sinogram lines, spread (the interpolation's
//             transpose as a gather over grid points: a fixed order, no atomics) -> inverse 2-D FFT -> crop (real
//             part / psi at the pixels)
// The interpolation and its transpose are sparse matrices built once on the host (CSR over samples and over grid
// points, the same float32 weights in both).
#include "dfm_common.h"

namespace {

__global__ void __launch_bounds__(256) nufft_pad_kernel(const float* __restrict__ img, int kstride, int k_off,
                                                        int kn, const float* __restrict__ deapod,
                                                        const int* __restrict__ gidx, int64_t npix,
                                                        float2* __restrict__ grid) {
  const int64_t total = npix * kn;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / kn;
    const int k = (int)(i - p * kn);
    grid[(int64_t)gidx[p] * kn + k] = make_float2(img[p * kstride + k_off + k] * deapod[p], 0.f);
  }
}

// spec[(ai H + j) kn + k] = rowfac[r] * sum_e w[e] G[col[e] kn + k], r = angles[ai] H + j
__global__ void __launch_bounds__(256) nufft_interp_kernel(const float2* __restrict__ G, const int* __restrict__ ptr,
                                                           const int* __restrict__ col, const float* __restrict__ w,
                                                           const float2* __restrict__ rowfac,
                                                           const int* __restrict__ angles, int n_ang, int H, int kn,
                                                           float2* __restrict__ spec) {
  const int64_t total = (int64_t)n_ang * H * kn;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = i / kn;
    const int k = (int)(i - row * kn);
    const int ai = (int)(row / H), j = (int)(row - (int64_t)ai * H);
    const int r = angles[ai] * H + j;
    float ax = 0.f, ay = 0.f;
    for (int e = ptr[r]; e < ptr[r + 1]; ++e) {
      const float2 g = G[(int64_t)col[e] * kn + k];
      ax = fmaf(w[e], g.x, ax);
      ay = fmaf(w[e], g.y, ay);
    }
    const float2 f = rowfac[r];
    spec[i] = make_float2(f.x * ax - f.y * ay, f.x * ay + f.y * ax);
  }
}

// G[t kn + k] = sum_e tw[e] colfac[r] Z[(apos[a] H + j) kn + k] over the samples r = a H + j = smp[e] with a tap on
// grid point t whose angle is in the call's list (apos[a] >= 0)
__global__ void __launch_bounds__(256) nufft_spread_kernel(const float2* __restrict__ Z, const int* __restrict__ tptr,
                                                           const int* __restrict__ smp, const float* __restrict__ tw,
                                                           const float2* __restrict__ colfac,
                                                           const int* __restrict__ apos, int H, int kn,
                                                           int64_t n_grid, float2* __restrict__ G) {
  const int64_t total = n_grid * kn;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = i / kn;
    const int k = (int)(i - t * kn);
    float ax = 0.f, ay = 0.f;
    for (int e = tptr[t]; e < tptr[t + 1]; ++e) {
      const int r = smp[e];
      const int a = r / H;
      const int ap = apos[a];
      if (ap < 0) continue;
      const float2 z = Z[((int64_t)ap * H + (r - a * H)) * kn + k];
      const float2 f = colfac[r];
      const float wf = tw[e];
      ax = fmaf(wf, f.x * z.x - f.y * z.y, ax);
      ay = fmaf(wf, f.x * z.y + f.y * z.x, ay);
    }
    G[i] = make_float2(ax, ay);
  }
}

// out[p kstride + k_off + k] (+)= scale * Re(g[gidx[p] kn + k]) * deapod[p]
__global__ void __launch_bounds__(256) nufft_crop_kernel(const float2* __restrict__ g, const int* __restrict__ gidx,
                                                         const float* __restrict__ deapod, int64_t npix, int kn,
                                                         float scale, float* __restrict__ out, int kstride, int k_off,
                                                         int accumulate) {
  const int64_t total = npix * kn;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / kn;
    const int k = (int)(i - p * kn);
    const float v = scale * g[(int64_t)gidx[p] * kn + k].x * deapod[p];
    float* o = out + p * kstride + k_off + k;
    *o = accumulate ? *o + v : v;
  }
}

// x, z, T (rows, ld) and dz (rows, ld_dz), columns 0 .. cols - 1: the SPDHG update (spdhg_poisson.py:95-106) on a
// column block of a channel-fastest array
__global__ void __launch_bounds__(256) spdhg_update_2d_kernel(float* __restrict__ x, float* __restrict__ z,
                                                              const float* __restrict__ T,
                                                              const float* __restrict__ dz, int64_t rows, int cols,
                                                              int ld, int ld_dz, float inv_p, float lam) {
  const int64_t total = rows * cols;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t rr = i / cols;
    const int c = (int)(i - rr * cols);
    const int64_t o = rr * ld + c;
    const float d = dz[rr * ld_dz + c];
    const float zn = z[o] + d;
    z[o] = zn;
    const float t = T[o];
    x[o] = t > 0.0f ? fmaxf(0.0f, x[o] - t * ((zn + inv_p * d) + lam)) : 0.0f;
  }
}

}  // namespace

extern "C" {

int dfm_nufft_pad(const float* img_k, int32_t kstride, int32_t k_off, int32_t kn, const float* deapod,
                  const int32_t* gidx, int64_t npix, void* grid, dfm_stream_t stream) {
  DFM_CHECK_ARG(img_k && deapod && gidx && grid && kn >= 0 && k_off >= 0 && k_off + kn <= kstride && npix >= 0,
                "dfm_nufft_pad: bad arguments");
  if (npix == 0 || kn == 0) return DFM_OK;
  hipLaunchKernelGGL(nufft_pad_kernel, dim3(dfm::grid_stride_blocks(npix * kn, 256)), dim3(256), 0,
                     (hipStream_t)stream, img_k, kstride, k_off, kn, deapod, gidx, npix, (float2*)grid);
  DFM_LAUNCH_CHECK("dfm_nufft_pad");
  return DFM_OK;
}

int dfm_nufft_interp(const void* G, const int32_t* ptr, const int32_t* col, const float* w, const void* rowfac,
                     const int32_t* angles, int32_t n_ang, int32_t H, int32_t kn, void* spec, dfm_stream_t stream) {
  DFM_CHECK_ARG(G && ptr && col && w && rowfac && angles && spec && n_ang >= 0 && H > 0 && kn >= 0,
                "dfm_nufft_interp: bad arguments");
  if (n_ang == 0 || kn == 0) return DFM_OK;
  const int64_t total = (int64_t)n_ang * H * kn;
  hipLaunchKernelGGL(nufft_interp_kernel, dim3(dfm::grid_stride_blocks(total, 256)), dim3(256), 0,
                     (hipStream_t)stream, (const float2*)G, ptr, col, w, (const float2*)rowfac, angles, n_ang, H, kn,
                     (float2*)spec);
  DFM_LAUNCH_CHECK("dfm_nufft_interp");
  return DFM_OK;
}

int dfm_nufft_spread(const void* Z, const int32_t* tptr, const int32_t* smp, const float* tw, const void* colfac,
                     const int32_t* apos, int32_t H, int32_t kn, int64_t n_grid, void* G, dfm_stream_t stream) {
  DFM_CHECK_ARG(Z && tptr && smp && tw && colfac && apos && G && H > 0 && kn >= 0 && n_grid > 0,
                "dfm_nufft_spread: bad arguments");
  if (kn == 0) return DFM_OK;
  hipLaunchKernelGGL(nufft_spread_kernel, dim3(dfm::grid_stride_blocks(n_grid * kn, 256)), dim3(256), 0,
                     (hipStream_t)stream, (const float2*)Z, tptr, smp, tw, (const float2*)colfac, apos, H, kn, n_grid,
                     (float2*)G);
  DFM_LAUNCH_CHECK("dfm_nufft_spread");
  return DFM_OK;
}

int dfm_nufft_crop(const void* g, const int32_t* gidx, const float* deapod, int64_t npix, int32_t kn, float scale,
                   float* out_k, int32_t kstride, int32_t k_off, int32_t accumulate, dfm_stream_t stream) {
  DFM_CHECK_ARG(g && gidx && deapod && out_k && kn >= 0 && k_off >= 0 && k_off + kn <= kstride && npix >= 0,
                "dfm_nufft_crop: bad arguments");
  if (npix == 0 || kn == 0) return DFM_OK;
  hipLaunchKernelGGL(nufft_crop_kernel, dim3(dfm::grid_stride_blocks(npix * kn, 256)), dim3(256), 0,
                     (hipStream_t)stream, (const float2*)g, gidx, deapod, npix, kn, scale, out_k, kstride, k_off,
                     accumulate);
  DFM_LAUNCH_CHECK("dfm_nufft_crop");
  return DFM_OK;
}

int dfm_spdhg_update_2d(float* x, float* z, const float* T, const float* dz, int64_t rows, int32_t cols, int32_t ld,
                        int32_t ld_dz, float inv_p, float lam, dfm_stream_t stream) {
  DFM_CHECK_ARG(x && z && T && dz && rows >= 0 && cols >= 0 && ld >= cols && ld_dz >= cols,
                "dfm_spdhg_update_2d: bad arguments");
  if (rows == 0 || cols == 0) return DFM_OK;
  hipLaunchKernelGGL(spdhg_update_2d_kernel, dim3(dfm::grid_stride_blocks(rows * cols, 256)), dim3(256), 0,
                     (hipStream_t)stream, x, z, T, dz, rows, cols, ld, ld_dz, inv_p, lam);
  DFM_LAUNCH_CHECK("dfm_spdhg_update_2d");
  return DFM_OK;
}

}  // extern "C"
