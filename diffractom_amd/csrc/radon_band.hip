// Forward projection tiled by pixel bands (parallel_radon.py:92-124 forward weights, the geometry-table form of the
// projector):
//
//   part[b, a, s, k] = sum over the table entries (pixel p, weight w) of ray (a, s) with p in band b of w img[p, k]
//   sino[a, s, k]    = sum_b part[b, a, s, k]           (fixed band order)
//
// A band is a range of image rows; the entries of a ray are sorted by pixel index, so the entries of a band are one
// contiguous range per ray (band_ptr). A workgroup covers (64 channels, band b, angle a, 16 bins); workgroups are
// mapped so that one XCD works through the (channel block, band) pairs with blk % 8 == its index, i.e. with 8 bands
// each XCD streams one band of the image: the gathered image rows stay in its 4 MB L2 (64 channels x the band's
// pixels) instead of the whole image slice of a channel block, which misses L2 for large images.
#include "dfm_common.h"

namespace {

constexpr int N_XCD = 8;
constexpr int BCH = 128;   // entries per bin staged per pass

__device__ __forceinline__ bool xcd_split(int per_blk, int n_blk, int& blk, int& local) {
  const int i = blockIdx.x;
  const int xcd = i % N_XCD;
  const int j = i / N_XCD;
  blk = (j / per_blk) * N_XCD + xcd;
  local = j % per_blk;
  return blk < n_blk;
}

template <int UNR, int BQ>
__global__ void __launch_bounds__(256) radon_forward_band_kernel(const float4* __restrict__ img,
                                                                 float4* __restrict__ part,
                                                                 const int* __restrict__ band_ptr,
                                                                 const int* __restrict__ ent_pix,
                                                                 const float* __restrict__ ent_w, int ns, int na,
                                                                 int nb, int kstride4, int k4_off, int k4_n,
                                                                 int pkstride4, int64_t pband4) {
  constexpr int NBIN = 256 / BQ;                                   // bins per workgroup
  __shared__ int lp[NBIN][BCH + 1];
  __shared__ float lw[NBIN][BCH + 1];
  __shared__ int lmax[4];
  const int sblocks = (ns + NBIN - 1) / NBIN;
  const int n_cb = (k4_n + BQ - 1) / BQ;
  int blk, local;
  if (!xcd_split(na * sblocks, n_cb * nb, blk, local)) return;   // uniform over the workgroup
  const int cb = blk / nb, b = blk % nb;
  const int a = local / sblocks;
  const int sl = threadIdx.x / BQ;                                 // bin of the workgroup
  const int q = threadIdx.x & (BQ - 1);                            // channel quad of the block
  const int s = (local % sblocks) * NBIN + sl;
  const int kq = cb * BQ + q;
  const bool live = kq < k4_n && s < ns;
  const int r = a * ns + min(s, ns - 1);
  const int e0 = s < ns ? band_ptr[(int64_t)r * nb + b] : 0;
  const int e1 = s < ns ? band_ptr[(int64_t)r * nb + b + 1] : 0;
  const int len = e1 - e0;
  // longest bin range of the workgroup: the number of staging passes
  int m = len;
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) lmax[threadIdx.x >> 6] = m;
  __syncthreads();
  const int maxlen = max(max(lmax[0], lmax[1]), max(lmax[2], lmax[3]));
  const float4* col = img + (live ? kq : 0) + k4_off;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int c = 0; c < maxlen; c += BCH) {
    __syncthreads();
    for (int j = q; j < BCH; j += BQ) {                            // the BQ threads of bin sl copy its entries
      const int e = e0 + c + j;
      if (e < e1) {
        lp[sl][j] = ent_pix[e];
        lw[sl][j] = ent_w[e];
      }
    }
    __syncthreads();
    const int n = min(BCH, len - c);
    int j = 0;
    for (; j + UNR <= n; j += UNR) {
      int p[UNR];
      float w[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        p[u] = lp[sl][j + u];
        w[u] = lw[sl][j + u];
      }
      float4 v[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) v[u] = col[(int64_t)p[u] * kstride4];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        acc.x += w[u] * v[u].x;
        acc.y += w[u] * v[u].y;
        acc.z += w[u] * v[u].z;
        acc.w += w[u] * v[u].w;
      }
    }
    for (; j < n; ++j) {
      const float w = lw[sl][j];
      const float4 v = col[(int64_t)lp[sl][j] * kstride4];
      acc.x += w * v.x;
      acc.y += w * v.y;
      acc.z += w * v.z;
      acc.w += w * v.w;
    }
  }
  if (live) part[(int64_t)b * pband4 + (int64_t)r * pkstride4 + kq] = acc;
}

// sino[r, kq] = sum_{b < nb} part[b, r, kq] for r < n_rows, kq < k4_n (float4 units).
__global__ void __launch_bounds__(256) band_reduce_kernel(const float4* __restrict__ part, float4* __restrict__ sino,
                                                          int64_t n_rows, int k4_n, int pkstride4, int64_t pband4,
                                                          int skstride4, int nb) {
  const int64_t total = n_rows * k4_n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / k4_n;
    const int kq = (int)(i - r * k4_n);
    const float4* p = part + r * pkstride4 + kq;
    float4 acc = dfm::ld_stream(p);
    for (int b = 1; b < nb; ++b) {
      const float4 v = dfm::ld_stream(p + (int64_t)b * pband4);
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
    dfm::st_stream(sino + r * skstride4 + kq, acc);
  }
}

}  // namespace

extern "C" {

int dfm_radon_forward_band(const float* img_k, float* sino, float* part, const int32_t* band_ptr,
                           const int32_t* ent_pix, const float* ent_w, int32_t ns, int32_t n_angles, int32_t nb,
                           int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride, int32_t variant,
                           dfm_stream_t stream) {
  DFM_CHECK_ARG(img_k && sino && part && band_ptr && ent_pix && ent_w, "dfm_radon_forward_band: null pointer");
  DFM_CHECK_ARG(ns > 0 && n_angles >= 0 && nb >= 1 && nb <= 64 && kstride % 4 == 0 && k_off % 4 == 0 &&
                    k_n % 4 == 0 && k_off >= 0 && k_n >= 0 && k_off + k_n <= kstride && skstride % 4 == 0 &&
                    skstride >= k_n,
                "dfm_radon_forward_band: bad sizes");
  // variant = entries ahead (4, 8) + 16 * w: 64 (w 0), 128 (w 1) or 256 (w 2) channels per workgroup
  const int unr = variant % 16, wsel = variant / 16;
  DFM_CHECK_ARG((unr == 4 || unr == 8) && wsel <= 2,
                "dfm_radon_forward_band: variant %d: entries ahead 4 or 8, + 16 w for 64 << w channels", variant);
  if (n_angles == 0 || k_n == 0) return DFM_OK;
  const int k4n = k_n / 4;
  const int BQ = 16 << wsel;
  const int per_blk = n_angles * ((ns + 256 / BQ - 1) / (256 / BQ));
  const int n_blk = (k4n + BQ - 1) / BQ * nb;
  const unsigned grid = (unsigned)((n_blk + N_XCD - 1) / N_XCD * N_XCD * per_blk);
  const int64_t n_rows = (int64_t)n_angles * ns;
  const int64_t pband4 = n_rows * (skstride / 4);
  const hipStream_t st = (hipStream_t)stream;
#define DFM_FB(U, Q)                                                                                              \
  hipLaunchKernelGGL((radon_forward_band_kernel<U, Q>), dim3(grid), dim3(256), 0, st, (const float4*)img_k,      \
                     (float4*)part, band_ptr, ent_pix, ent_w, ns, n_angles, nb, kstride / 4, k_off / 4, k4n,        \
                     skstride / 4, pband4)
#define DFM_FBQ(U)        \
  if (BQ == 16) {         \
    DFM_FB(U, 16);        \
  } else if (BQ == 32) {  \
    DFM_FB(U, 32);        \
  } else {                \
    DFM_FB(U, 64);        \
  }
  if (unr == 4) {
    DFM_FBQ(4)
  } else {
    DFM_FBQ(8)
  }
#undef DFM_FBQ
#undef DFM_FB
  DFM_LAUNCH_CHECK("dfm_radon_forward_band");
  const int64_t total = n_rows * k4n;
  hipLaunchKernelGGL(band_reduce_kernel, dim3(dfm::grid_stride_blocks(total, 256, 1u << 16)), dim3(256), 0, st,
                     (const float4*)part, (float4*)sino, n_rows, k4n, skstride / 4, pband4, skstride / 4, nb);
  DFM_LAUNCH_CHECK("dfm_radon_forward_band (band sum)");
  return DFM_OK;
}

}  // extern "C"
