// Kernels of the local-basis operator (operators/local_forward_operator.py, local_kernels.cl:38-319): orientation k
// carries coefficients only in rectangular patches (h x w boxes of the grid, row-major at coff[p] of a flat vector,
// a byte per coefficient flagging the allowed ones).
//
// At angle a a patch projects onto a strip of W[p] detector bins starting at bin off[a, p]; the strips of an
// orientation batch are stored per angle, strips[a * SW + pw_off[p] + j] (SW = sum of the batch's W), so the bins of
// one angle are contiguous across the batch's patches. The per-(angle, patch) record ptab[a * P + p] = (off, tb)
// holds the strip's first bin and tb = t(x0, y0) - off, the detector coordinate of the patch's corner pixel relative
// to the strip (computed in float64 on the host), so the kernels evaluate t = tb + X du + Y dv on small patch-local
// offsets du, dv in float32.
//
// The PF products are the operator's own (block-sparse matrix-core or CSR kernels) on the batch's orientation-fastest
// sinogram (n_ang, My, kstride):
//   forward: strips (dfm_local_strips) -> the dense batch sinogram, zero outside the strips (dfm_local_assemble)
//            -> PF product
//   adjoint: PF adjoint -> the strips' bins of the sinogram (dfm_local_extract) -> backprojection of every patch
//            coefficient from its strips, or the SPDHG primal update fused into it (dfm_local_backproject)
// Footprint: trapezoid min(max(hw - |t - s|, 0), top) / A (hat: hw = top = A = 1), the operator's geometry records
// (R x 8 floats: X, Y, T0, scale, hw, top, 0, 0; the forward's scale 1/A, the backward's w/A).
#include "dfm_common.h"

// LDS of a workgroup that stages a patch (the strip kernel) or a patch's strips (the backprojection); larger ones
// take the global-memory forms.
#define DFM_LOCAL_LDS_BYTES (64 * 1024)

namespace {

// The footprint-weighted sum over a patch's pixels for strip bin j at one angle: the pixels are visited along the
// axis the projection direction varies faster in (|X| >= |Y|: along image rows), on every row (minor index m) whose
// footprint band meets the box the pixels u with |u - uc(m)| < reach, uc the band's centre, at most NT of them
// (NT = ceil(2 reach) of the steepest angle: the loop is unrolled and predicated; NT 0: a counted loop instead).
// val(q) gives the allowed coefficient of patch pixel q (0 where not allowed).
template <int NT, typename Val>
__device__ __forceinline__ float strip_bin(const Val& val, int h, int w, float X, float Y, float hw, float top, float tb,
                                           int j) {
  const bool along_x = fabsf(X) >= fabsf(Y);
  const float Am = along_x ? X : Y, Bm = along_x ? Y : X;
  const int n_major = along_x ? w : h, n_minor = along_x ? h : w;
  const int s_major = along_x ? 1 : w, s_minor = along_x ? w : 1;
  const float iA = 1.0f / Am;
  const float reach = hw * fabsf(iA);          // the footprint's half-width along the major axis, in pixels
  const float S = Bm * iA;                     // |S| <= 1: the footprint's centre moves by -S per minor step
  const float fj = (float)j;
  const float c0 = (fj - tb) * iA;             // the centre (major-axis pixel) at minor index 0
  const float lo_c = -reach, hi_c = (float)(n_major - 1) + reach;
  int m0 = 0, m1 = n_minor - 1;
  if (fabsf(S) > 1e-6f) {
    const float r1 = (c0 - hi_c) / S, r2 = (c0 - lo_c) / S;
    m0 = max(m0, (int)floorf(fmaxf(fminf(r1, r2), -1.0f)));
    m1 = min(m1, (int)ceilf(fminf(fmaxf(r1, r2), (float)n_minor)));
  } else if (c0 <= lo_c - 1.0f || c0 >= hi_c + 1.0f) {
    m1 = -1;
  }
  float acc = 0.0f;
  for (int m = m0; m <= m1; ++m) {
    const float uc = c0 - S * (float)m;
    const float tm = fmaf(Bm, (float)m, tb) - fj;        // t - j at u = 0
    const int qm = m * s_minor;
    const int u0 = (int)floorf(fminf(fmaxf(uc - reach, -2.0f), (float)n_major)) + 1;   // first pixel inside
    if (NT > 0) {
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        const int u = u0 + k;
        const float wt = fminf(fmaxf(hw - fabsf(fmaf(Am, (float)u, tm)), 0.0f), top);
        const float v = val(qm + min(max(u, 0), n_major - 1) * s_major);
        acc += (u >= 0 && u < n_major) ? wt * v : 0.0f;
      }
    } else {
      const int u1 = min(n_major - 1, (int)ceilf(fminf(uc + reach, (float)n_major)));
      for (int u = max(u0, 0); u <= u1; ++u)
        acc = fmaf(fminf(fmaxf(hw - fabsf(fmaf(Am, (float)u, tm)), 0.0f), top), val(qm + u * s_major), acc);
    }
  }
  return acc;
}

// strips[a * SW + jg] for every angle a = omegas[as] and strip bin jg of the batch: the footprint-weighted sum of the
// patch's allowed coefficients, times scale / A. A thread per (angle, bin), the coefficients read from global memory
// (patches too large for the LDS).
template <int NT>
__global__ void __launch_bounds__(256) local_strips_kernel(const float* __restrict__ coef,
                                                           const uint8_t* __restrict__ mask,
                                                           const int4* __restrict__ box,
                                                           const int64_t* __restrict__ coff,
                                                           const int* __restrict__ pw_off, int SW,
                                                           const int* __restrict__ jp, const int2* __restrict__ ptab,
                                                           int P, const float4* __restrict__ geo,
                                                           const int* __restrict__ omegas, int n_ang, float scale,
                                                           float* __restrict__ strips) {
  const int64_t total = (int64_t)n_ang * SW;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int as = (int)(i / SW);
    const int jg = (int)(i - (int64_t)as * SW);
    const int p = jp[jg];
    const int a = omegas[as];
    const float4 g0 = geo[2 * a], g1 = geo[2 * a + 1];
    const int4 bx = box[p];
    const float* cp = coef + coff[p];
    const uint8_t* mp = mask + coff[p];
    const auto val = [&](int q) { return mp[q] ? cp[q] : 0.0f; };
    const float acc = strip_bin<NT>(val, bx.z, bx.w, g0.x, g0.y, g1.x, g1.y,
                                    __int_as_float(ptab[(int64_t)a * P + p].y), jg - pw_off[p]);
    strips[(int64_t)a * SW + jg] = scale * g0.w * acc;
  }
}

// The same with a workgroup per patch: the patch's allowed coefficients staged in LDS once (0 where not allowed),
// then the workgroup's threads run over the patch's (angle, bin) pairs.
template <int NT>
__global__ void __launch_bounds__(256) local_strips_lds_kernel(const float* __restrict__ coef,
                                                               const uint8_t* __restrict__ mask,
                                                               const int4* __restrict__ box,
                                                               const int64_t* __restrict__ coff,
                                                               const int* __restrict__ W,
                                                               const int* __restrict__ pw_off, int SW,
                                                               const int2* __restrict__ ptab, int P,
                                                               const float4* __restrict__ geo,
                                                               const int* __restrict__ omegas, int n_ang, float scale,
                                                               float* __restrict__ strips) {
  extern __shared__ float lp[];
  const int p = blockIdx.x;
  const int4 bx = box[p];
  const int n = bx.z * bx.w;
  const int64_t c0 = coff[p];
  for (int q = threadIdx.x; q < n; q += blockDim.x) lp[q] = mask[c0 + q] ? coef[c0 + q] : 0.0f;
  __syncthreads();
  const int Wp = W[p], base = pw_off[p];
  const auto val = [&](int q) { return lp[q]; };
  for (int it = threadIdx.x; it < n_ang * Wp; it += blockDim.x) {
    const int as = it / Wp, j = it - as * Wp;
    const int a = omegas[as];
    const float4 g0 = geo[2 * a], g1 = geo[2 * a + 1];
    const float acc = strip_bin<NT>(val, bx.z, bx.w, g0.x, g0.y, g1.x, g1.y,
                                    __int_as_float(ptab[(int64_t)a * P + p].y), j);
    strips[(int64_t)a * SW + base + j] = scale * g0.w * acc;
  }
}

// sino[(as * My + y) * kstride + k] = the sum over the patches of orientation k (in patch order) of their strip bin
// at detector line y, 0 where no strip reaches (and in the padding channels k >= k_batch). Channels run across the
// threads (coalesced rows), rows over blockIdx.y.
__global__ void __launch_bounds__(256) local_assemble_kernel(const float* __restrict__ strips, int SW,
                                                             const int* __restrict__ pw_off,
                                                             const int* __restrict__ W,
                                                             const int2* __restrict__ ptab, int P,
                                                             const int* __restrict__ kp_ptr, int k_batch,
                                                             const int* __restrict__ omegas, int n_ang, int My,
                                                             int kstride, float* __restrict__ sino) {
  const int64_t rows = (int64_t)n_ang * My;
  for (int64_t row = blockIdx.y; row < rows; row += gridDim.y) {
    const int as = (int)(row / My);
    const int y = (int)(row - (int64_t)as * My);
    const int a = omegas[as];
    const float* sa = strips + (int64_t)a * SW;
    const int2* pa = ptab + (int64_t)a * P;
    float* out = sino + row * kstride;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kstride; k += gridDim.x * blockDim.x) {
      float v = 0.0f;
      if (k < k_batch) {
        const int p1 = kp_ptr[k + 1];
        for (int p = kp_ptr[k]; p < p1; ++p) {
          const int jj = y - pa[p].x;
          if (jj >= 0 && jj < W[p]) v += sa[pw_off[p] + jj];
        }
      }
      out[k] = v;
    }
  }
}

// strips[a * SW + jg] = sino[(as * My + off + j) * kstride + pk[p]] (0 off the detector): the bins of every patch's
// strips in the batch's adjoint sinogram.
__global__ void __launch_bounds__(256) local_extract_kernel(const float* __restrict__ sino, int n_ang, int My,
                                                            int kstride, const int* __restrict__ omegas,
                                                            const int2* __restrict__ ptab, int P,
                                                            const int* __restrict__ pk, const int* __restrict__ jp,
                                                            const int* __restrict__ pw_off, int SW,
                                                            float* __restrict__ strips) {
  const int64_t total = (int64_t)n_ang * SW;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int as = (int)(i / SW);
    const int jg = (int)(i - (int64_t)as * SW);
    const int p = jp[jg];
    const int a = omegas[as];
    const int y = ptab[(int64_t)a * P + p].x + (jg - pw_off[p]);
    const float v = (y >= 0 && y < My) ? sino[((int64_t)as * My + y) * kstride + pk[p]] : 0.0f;
    strips[(int64_t)a * SW + jg] = v;
  }
}

// The backprojection of every allowed coefficient of the batch (coefficients c_begin .. c_begin + n, a thread each;
// pid: the patch of each): acc = sum_a (w / A)_a sum_taps footprint(t - s) strip[a, s]. mode 0: out = acc (0 where
// not allowed); mode 1: the SPDHG primal update with dz = acc (spdhg_poisson.py:95-106): z += dz,
// x = T > 0 ? max(0, x - T (z + inv_p dz + lam)) : 0 (x = 0 where not allowed: T is 0 there).
template <int NTAP>
__global__ void __launch_bounds__(256) local_backproject_kernel(const float* __restrict__ strips, int SW,
                                                                const int* __restrict__ pw_off,
                                                                const int* __restrict__ W,
                                                                const int2* __restrict__ ptab, int P,
                                                                const float4* __restrict__ geo,
                                                                const int* __restrict__ omegas, int n_ang,
                                                                const int4* __restrict__ box,
                                                                const int64_t* __restrict__ coff,
                                                                const uint8_t* __restrict__ mask,
                                                                const int* __restrict__ pid, int64_t c_begin,
                                                                int64_t n, int mode, float* __restrict__ out,
                                                                float* __restrict__ x, float* __restrict__ z,
                                                                const float* __restrict__ T, float inv_p, float lam) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = c_begin + l;
    if (!mask[i]) {
      if (mode == 0) out[i] = 0.0f;
      else x[i] = 0.0f;
      continue;
    }
    const int p = pid[l];
    const int4 bx = box[p];
    const int q = (int)(i - coff[p]);
    const int dv = q / bx.w, du = q - dv * bx.w;
    const float fdu = (float)du, fdv = (float)dv;
    const int Wp = W[p];
    const int base = pw_off[p];
    float acc = 0.0f;
    for (int as = 0; as < n_ang; ++as) {
      const int a = omegas[as];
      const float4 g0 = geo[2 * a], g1 = geo[2 * a + 1];
      const float t = fmaf(g0.x, fdu, fmaf(g0.y, fdv, __int_as_float(ptab[(int64_t)a * P + p].y)));
      const float hw = g1.x, top = g1.y;
      const int s0 = (int)floorf(t - hw) + 1;
      const float* sa = strips + (int64_t)a * SW + base;
      float v = 0.0f;
#pragma unroll
      for (int tap = 0; tap < NTAP; ++tap) {
        const int s = s0 + tap;
        const float wt = fminf(fmaxf(hw - fabsf(t - (float)s), 0.0f), top);
        v = fmaf(wt, sa[min(max(s, 0), Wp - 1)], v);
      }
      acc = fmaf(g0.w, v, acc);
    }
    if (mode == 0) {
      out[i] = acc;
    } else {
      const float zi = z[i] + acc;
      z[i] = zi;
      const float tt = T[i];
      x[i] = tt > 0.0f ? fmaxf(0.0f, x[i] - tt * ((zi + inv_p * acc) + lam)) : 0.0f;
    }
  }
}

// The same with a workgroup per 256 pixels of one patch (blk: (patch, first pixel) per workgroup): the patch's strip
// bins at the call's angles and the angles' geometry staged in LDS, so the taps are LDS reads.
template <int NTAP>
__global__ void __launch_bounds__(256) local_backproject_lds_kernel(const float* __restrict__ strips, int SW,
                                                                    const int* __restrict__ pw_off,
                                                                    const int* __restrict__ W,
                                                                    const int2* __restrict__ ptab, int P,
                                                                    const float4* __restrict__ geo,
                                                                    const int* __restrict__ omegas, int n_ang,
                                                                    const int4* __restrict__ box,
                                                                    const int64_t* __restrict__ coff,
                                                                    const uint8_t* __restrict__ mask,
                                                                    const int2* __restrict__ blk, int mode,
                                                                    float* __restrict__ out, float* __restrict__ x,
                                                                    float* __restrict__ z,
                                                                    const float* __restrict__ T, float inv_p,
                                                                    float lam) {
  extern __shared__ float4 lds4[];
  const int2 e = blk[blockIdx.x];
  const int p = e.x;
  const int Wp = W[p], base = pw_off[p];
  float4* lg = lds4;                                    // per angle: (X, Y, tb, w / A), (hw, top, -, -)
  float* ls = reinterpret_cast<float*>(lds4 + 2 * n_ang);   // (n_ang, Wp) strip bins
  for (int as = threadIdx.x; as < n_ang; as += blockDim.x) {
    const int a = omegas[as];
    const float4 g0 = geo[2 * a], g1 = geo[2 * a + 1];
    lg[2 * as] = make_float4(g0.x, g0.y, __int_as_float(ptab[(int64_t)a * P + p].y), g0.w);
    lg[2 * as + 1] = g1;
  }
  for (int it = threadIdx.x; it < n_ang * Wp; it += blockDim.x) {
    const int as = it / Wp;
    ls[it] = strips[(int64_t)omegas[as] * SW + base + (it - as * Wp)];
  }
  __syncthreads();
  const int4 bx = box[p];
  const int q = e.y + (int)threadIdx.x;
  if (q >= bx.z * bx.w) return;
  const int64_t i = coff[p] + q;
  if (!mask[i]) {
    if (mode == 0) out[i] = 0.0f;
    else x[i] = 0.0f;
    return;
  }
  const int dv = q / bx.w, du = q - dv * bx.w;
  const float fdu = (float)du, fdv = (float)dv;
  float acc = 0.0f;
  for (int as = 0; as < n_ang; ++as) {
    const float4 g0 = lg[2 * as], g1 = lg[2 * as + 1];
    const float t = fmaf(g0.x, fdu, fmaf(g0.y, fdv, g0.z));
    const int s0 = (int)floorf(t - g1.x) + 1;
    const float* sa = ls + as * Wp;
    float v = 0.0f;
#pragma unroll
    for (int tap = 0; tap < NTAP; ++tap) {
      const int s = s0 + tap;
      const float wt = fminf(fmaxf(g1.x - fabsf(t - (float)s), 0.0f), g1.y);
      v = fmaf(wt, sa[min(max(s, 0), Wp - 1)], v);
    }
    acc = fmaf(g0.w, v, acc);
  }
  if (mode == 0) {
    out[i] = acc;
  } else {
    const float zi = z[i] + acc;
    z[i] = zi;
    const float tt = T[i];
    x[i] = tt > 0.0f ? fmaxf(0.0f, x[i] - tt * ((zi + inv_p * acc) + lam)) : 0.0f;
  }
}

}  // namespace

extern "C" {

int dfm_local_strips(const float* coef, const uint8_t* mask, const int32_t* box, const int64_t* coff,
                     const int32_t* W, const int32_t* pw_off, int32_t sw, const int32_t* jp, const int32_t* ptab,
                     int32_t n_patches, int32_t max_area, int32_t n_reach, const float* geo, const int32_t* omegas,
                     int32_t n_ang, float scale, float* strips, dfm_stream_t stream) {
  DFM_CHECK_ARG(coef && mask && box && coff && W && pw_off && jp && ptab && geo && omegas && strips && sw >= 0 &&
                    n_patches > 0 && n_ang >= 0 && max_area > 0 && n_reach >= 0,
                "dfm_local_strips: bad arguments");
  const int64_t total = (int64_t)n_ang * sw;
  if (total == 0) return DFM_OK;
  const hipStream_t st = (hipStream_t)stream;
  const bool lds = (int64_t)max_area * 4 <= DFM_LOCAL_LDS_BYTES;
#define DFM_LOCAL_STRIPS(NT)                                                                                       \
  if (lds)                                                                                                         \
    hipLaunchKernelGGL(local_strips_lds_kernel<NT>, dim3(n_patches), dim3(256), (size_t)max_area * 4, st, coef,   \
                       mask, (const int4*)box, coff, W, pw_off, sw, (const int2*)ptab, n_patches,                 \
                       (const float4*)geo, omegas, n_ang, scale, strips);                                          \
  else                                                                                                             \
    hipLaunchKernelGGL(local_strips_kernel<NT>, dim3(dfm::grid_stride_blocks(total, 256)), dim3(256), 0, st, coef, \
                       mask, (const int4*)box, coff, pw_off, sw, jp, (const int2*)ptab, n_patches,                \
                       (const float4*)geo, omegas, n_ang, scale, strips)
  // n_reach: the most pixels a bin's footprint covers along a row (ceil(2 reach)); larger ones take the counted loop
  if (n_reach == 2) DFM_LOCAL_STRIPS(2);
  else if (n_reach == 3) DFM_LOCAL_STRIPS(3);
  else if (n_reach == 4) DFM_LOCAL_STRIPS(4);
  else DFM_LOCAL_STRIPS(0);
#undef DFM_LOCAL_STRIPS
  DFM_LAUNCH_CHECK("dfm_local_strips");
  return DFM_OK;
}

int dfm_local_assemble(const float* strips, int32_t sw, const int32_t* pw_off, const int32_t* W, const int32_t* ptab,
                       int32_t n_patches, const int32_t* kp_ptr, int32_t k_batch, const int32_t* omegas,
                       int32_t n_ang, int32_t my, int32_t kstride, float* sino, dfm_stream_t stream) {
  DFM_CHECK_ARG(strips && pw_off && W && ptab && kp_ptr && omegas && sino && n_patches > 0 && k_batch > 0 &&
                    kstride >= k_batch && my > 0 && n_ang >= 0,
                "dfm_local_assemble: bad arguments");
  const int64_t rows = (int64_t)n_ang * my;
  if (rows == 0) return DFM_OK;
  const unsigned gx = (unsigned)((kstride + 255) / 256);
  const unsigned gy = (unsigned)(rows < 65535 ? rows : 65535);
  hipLaunchKernelGGL(local_assemble_kernel, dim3(gx, gy), dim3(256), 0, (hipStream_t)stream, strips, sw, pw_off, W,
                     (const int2*)ptab, n_patches, kp_ptr, k_batch, omegas, n_ang, my, kstride, sino);
  DFM_LAUNCH_CHECK("dfm_local_assemble");
  return DFM_OK;
}

int dfm_local_extract(const float* sino, int32_t n_ang, int32_t my, int32_t kstride, const int32_t* omegas,
                      const int32_t* ptab, int32_t n_patches, const int32_t* pk, const int32_t* jp,
                      const int32_t* pw_off, int32_t sw, float* strips, dfm_stream_t stream) {
  DFM_CHECK_ARG(sino && omegas && ptab && pk && jp && pw_off && strips && n_patches > 0 && my > 0 && kstride > 0 &&
                    n_ang >= 0 && sw >= 0,
                "dfm_local_extract: bad arguments");
  const int64_t total = (int64_t)n_ang * sw;
  if (total == 0) return DFM_OK;
  hipLaunchKernelGGL(local_extract_kernel, dim3(dfm::grid_stride_blocks(total, 256)), dim3(256), 0,
                     (hipStream_t)stream, sino, n_ang, my, kstride, omegas, (const int2*)ptab, n_patches, pk, jp,
                     pw_off, sw, strips);
  DFM_LAUNCH_CHECK("dfm_local_extract");
  return DFM_OK;
}

int dfm_local_backproject(const float* strips, int32_t sw, const int32_t* pw_off, const int32_t* W,
                          const int32_t* ptab, int32_t n_patches, int32_t w_max, const float* geo,
                          const int32_t* omegas, int32_t n_ang, const int32_t* box, const int64_t* coff,
                          const uint8_t* mask, const int32_t* pid, int64_t c_begin, int64_t n, const int32_t* blk,
                          int32_t n_blk, int32_t n_taps, int32_t mode, float* out, float* x, float* z, const float* T,
                          float inv_p, float lam, dfm_stream_t stream) {
  DFM_CHECK_ARG(strips && pw_off && W && ptab && geo && omegas && box && coff && mask && pid && blk &&
                    n_patches > 0 && w_max > 0 && n_ang >= 0 && c_begin >= 0 && n >= 0 && n_blk >= 0 &&
                    n_taps >= 2 && n_taps <= 4 && (mode == 0 || mode == 1),
                "dfm_local_backproject: bad arguments");
  DFM_CHECK_ARG(mode == 0 ? out != nullptr : (x && z && T), "dfm_local_backproject: missing output arrays");
  if (n == 0) return DFM_OK;
  const hipStream_t st = (hipStream_t)stream;
  const int64_t lds = (int64_t)n_ang * (32 + 4 * (int64_t)w_max);
  if (n_ang > 0 && n_blk > 0 && lds <= DFM_LOCAL_LDS_BYTES) {
#define DFM_LOCAL_BPL(NT)                                                                                          \
  hipLaunchKernelGGL(local_backproject_lds_kernel<NT>, dim3(n_blk), dim3(256), (size_t)lds, st, strips, sw, pw_off, \
                     W, (const int2*)ptab, n_patches, (const float4*)geo, omegas, n_ang, (const int4*)box, coff,  \
                     mask, (const int2*)blk, mode, out, x, z, T, inv_p, lam)
    if (n_taps == 2) DFM_LOCAL_BPL(2);
    else if (n_taps == 3) DFM_LOCAL_BPL(3);
    else DFM_LOCAL_BPL(4);
#undef DFM_LOCAL_BPL
    DFM_LAUNCH_CHECK("dfm_local_backproject");
    return DFM_OK;
  }
  const dim3 grid(dfm::grid_stride_blocks(n, 256)), block(256);
#define DFM_LOCAL_BP(NT)                                                                                         \
  hipLaunchKernelGGL(local_backproject_kernel<NT>, grid, block, 0, st, strips, sw, pw_off, W, (const int2*)ptab, \
                     n_patches, (const float4*)geo, omegas, n_ang, (const int4*)box, coff, mask, pid, c_begin, n,  \
                     mode, out, x, z, T, inv_p, lam)
  if (n_taps == 2) DFM_LOCAL_BP(2);
  else if (n_taps == 3) DFM_LOCAL_BP(3);
  else DFM_LOCAL_BP(4);
#undef DFM_LOCAL_BP
  DFM_LAUNCH_CHECK("dfm_local_backproject");
  return DFM_OK;
}

}  // extern "C"
