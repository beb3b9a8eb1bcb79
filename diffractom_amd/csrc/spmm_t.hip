// Pole-figure CSR products on transposed operands (wavefront-segmented SpMM), both directions with one kernel:
//
//   out[s, a, row, y] = alpha * sum_{e in CSR row (r, row), col(e) in split s} val[e] * in[a, col(e), y]
//
// forward (rows = segments j, columns = orientations k):  in = sinogram transposed (a, k, y), out -> data[r, y, j]
// adjoint (rows = orientations k, columns = segments j):  in = data rows transposed (a, j, y), out -> sino[a, y, k]
// with r = omegas[a] the angle's data row (pf_kernels.cl:597-687 spmm_pf_forward_c / spmm_pf_adjoint_c;
// single_phase_forward_operator.py:38-123).
//
// A workgroup of 16 waves owns 16 R consecutive CSR rows of one angle and one column range (split s). It walks the
// range in chunks of kc columns: the chunk of `in` (kc rows of ys floats, contiguous) is staged in LDS, then every wave
// consumes, for each of its R rows, the row's entries that fall in the chunk. The lanes of a wave run over the
// detector positions y (YW groups of 64), so an entry is wave-uniform (its column read from the row's 64-entry window
// held in registers with v_readlane, its value likewise from the window's values) and its operand is one
// conflict-free LDS row read. Per row the column window moves on only when all 64 of its entries are consumed.
// Summation follows the CSR order, as in the row-per-thread kernels.
#include <algorithm>

#include "dfm_common.h"

namespace {

constexpr int SPT_WAVES = 16;
constexpr int SPT_LDS_FLOATS = 40960;   // 160 KiB

template <int R, int YW, int U>
__global__ void __launch_bounds__(1024) spmm_t_kernel(const float* __restrict__ inT, const int* __restrict__ row_ptr,
                                                      const uint16_t* __restrict__ col, const float* __restrict__ val,
                                                      float* __restrict__ outT, const int* __restrict__ omegas,
                                                      int my, int ys, int ncols, int nrows, int rows_out, int compact,
                                                      float alpha, int cspan, int kc, int n_angles, int64_t in_ld,
                                                      int64_t out_ld, int ys_out) {
  extern __shared__ float lin[];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int a = blockIdx.y, s = blockIdx.z;
  const int ra = omegas ? omegas[a] : a;
  const int rr = compact ? a : ra;
  const int row0 = (blockIdx.x * SPT_WAVES + w) * R;
  const int c_lo = s * cspan, c_hi = min(ncols, c_lo + cspan);

  // lane l < R: first entry with column >= c_lo and the end of row row0 + l
  int cur = 0, end = 0;
  {
    const int row = row0 + lane;
    if (lane < R && row < nrows) {
      const int64_t rb = (int64_t)rr * nrows + row;
      int lo = row_ptr[rb], hi = row_ptr[rb + 1];
      end = hi;
      if (c_lo > 0) {
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if ((int)col[mid] < c_lo) lo = mid + 1;
          else hi = mid;
        }
      }
      cur = lo;
    }
  }
  // per row: window start we (entry index of lane 0), consumed count wo, end ee; window columns cb and values vw
  // per lane (one coalesced vector load each; an entry's column and value are then read with v_readlane)
  int we[R], wo[R], ee[R], cb[R];
  float vw[R];
  float acc[R][YW];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    we[i] = __builtin_amdgcn_readlane(cur, i);
    ee[i] = __builtin_amdgcn_readlane(end, i);
    wo[i] = 0;
    const int idx = we[i] + lane;
    cb[i] = idx < ee[i] ? (int)col[idx] : 0x7fffffff;
    vw[i] = idx < ee[i] ? val[idx] : 0.0f;
#pragma unroll
    for (int h = 0; h < YW; ++h) acc[i][h] = 0.f;
  }

  const float* lbase = lin + lane;
  // in_ld > 0: the operand untransposed, (my, in_ld) per angle, transposed while it is staged (ys odd: the
  // transposing LDS writes, a row of ys floats apart, hit distinct banks); in_ld == 0: transposed, (ncols, ys)
  const float* src = in_ld ? inT + (int64_t)a * my * in_ld : inT + (int64_t)a * ncols * ys;
  for (int c0 = c_lo; c0 < c_hi; c0 += kc) {
    const int kn = min(kc, c_hi - c0);
    __syncthreads();
    constexpr int SU = YW >= 7 ? 4 : 8;                      // staging loads in flight per lane
    if (in_ld) {
      // an item is 4 rows x 64 columns: wave w copies row quads g = w + 16 m, lane (cq_lo + 8 yq + 32 cq_hi) the
      // float4 of columns 4 cq .. 4 cq + 3 (cq = cq_lo + 8 cq_hi) of row 4 g + yq. c0, in_ld and the base are
      // multiples of 4 floats, and the columns up to kn rounded up to 4 lie in the row and in the LDS chunk; with
      // ys odd the 4 transposing writes of a float4 hit distinct banks over every 32 lanes (4 (cq ys mod 8) + yq).
      const int yq = (lane >> 3) & 3, cq = (lane & 7) + ((lane >> 5) << 3);
      const int nj = (kn + 63) / 64, ng = (my + 3) / 4;
      const int items = (ng - w + SPT_WAVES - 1) / SPT_WAVES * nj;
      for (int it0 = 0; it0 < items; it0 += SU) {
        float4 v[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
          const int it = it0 + u, m = it / nj, kk = 64 * (it - m * nj) + 4 * cq;
          const int y = 4 * (w + SPT_WAVES * m) + yq;
          v[u] = (it < items && y < my && kk < kn)
                     ? *reinterpret_cast<const float4*>(src + (int64_t)y * in_ld + c0 + kk)
                     : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < SU; ++u) {
          const int it = it0 + u, m = it / nj, kk = 64 * (it - m * nj) + 4 * cq;
          const int y = 4 * (w + SPT_WAVES * m) + yq;
          if (it < items && y < my && kk < kn) {
            float* l = lin + kk * ys + y;
            l[0] = v[u].x;
            l[ys] = v[u].y;
            l[2 * ys] = v[u].z;
            l[3 * ys] = v[u].w;
          }
        }
      }
    } else {
      const float4* g = reinterpret_cast<const float4*>(src + (int64_t)c0 * ys);
      float4* l4 = reinterpret_cast<float4*>(lin);
      const int n4 = kn * ys / 4;
      for (int i0 = threadIdx.x; i0 < n4; i0 += SPT_WAVES * 64 * SU) {
        float4 v[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
          const int i = i0 + SPT_WAVES * 64 * u;
          v[u] = i < n4 ? g[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < SU; ++u) {
          const int i = i0 + SPT_WAVES * 64 * u;
          if (i < n4) l4[i] = v[u];
        }
      }
    }
    __syncthreads();
    const int c1 = c0 + kn;
#pragma unroll
    for (int i = 0; i < R; ++i) {
      while (true) {
        const int cnt = __popcll(__ballot(cb[i] < c1));     // entries of the window below c1 (a prefix)
        const int off = (cb[i] - c0) * ys;                   // operand row offset of each window entry
        const int vbits = __float_as_int(vw[i]);
        int t = wo[i];
        for (; t + U <= cnt; t += U) {                        // U operand rows in flight, summed in entry order
          float x[U][YW], v[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const float* lp = lbase + __builtin_amdgcn_readlane(off, t + u);
            v[u] = __int_as_float(__builtin_amdgcn_readlane(vbits, t + u));
#pragma unroll
            for (int h = 0; h < YW; ++h) x[u][h] = lp[64 * h];
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int u = 0; u < U; ++u)
#pragma unroll
            for (int h = 0; h < YW; ++h) acc[i][h] += v[u] * x[u][h];
        }
        for (; t < cnt; ++t) {
          const float* lp = lbase + __builtin_amdgcn_readlane(off, t);
          const float vt = __int_as_float(__builtin_amdgcn_readlane(vbits, t));
#pragma unroll
          for (int h = 0; h < YW; ++h) acc[i][h] += vt * lp[64 * h];
        }
        wo[i] = cnt;
        if (cnt < 64 || we[i] + 64 >= ee[i]) break;
        we[i] += 64;                                       // window used up: the next 64 entries
        wo[i] = 0;
        const int idx = we[i] + lane;
        cb[i] = idx < ee[i] ? (int)col[idx] : 0x7fffffff;
        vw[i] = idx < ee[i] ? val[idx] : 0.0f;
      }
    }
  }

  if (out_ld) {
    // out (n_angles, my, out_ld): per group of 64 y, the workgroup's RB = 16 R rows go through LDS as (y, row) with
    // the odd row stride RB + 1 (64 (RB + 1) floats), then out one y row of RB consecutive floats at a time
    constexpr int RB = SPT_WAVES * R, RW = RB + 1;
    const int r = threadIdx.x % RB, row = blockIdx.x * RB + r;
    float* dst = outT + (int64_t)a * my * out_ld;
#pragma unroll
    for (int h = 0; h < YW; ++h) {
      __syncthreads();
#pragma unroll
      for (int i = 0; i < R; ++i) lin[lane * RW + w * R + i] = alpha * acc[i][h];
      __syncthreads();
      if (row < rows_out)
        for (int yl = threadIdx.x / RB; yl < 64 && 64 * h + yl < my; yl += SPT_WAVES * 64 / RB)
          dst[(int64_t)(64 * h + yl) * out_ld + row] = lin[yl * RW + r];
    }
    return;
  }
  float* dst = outT + ((int64_t)s * n_angles + a) * rows_out * ys_out;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int row = row0 + i;
    if (row < rows_out) {
#pragma unroll
      for (int h = 0; h < YW; ++h) {
        const int y = 64 * h + lane;
        if (y < my) dst[(int64_t)row * ys_out + y] = alpha * acc[i][h];
      }
    }
  }
}

// out[ao, c, r] (+)= sum_{s < nsplit} in[s * split_stride + ai * in_mat + r * in_ld + c]: a batch of (nr x nc)
// transposes through 32 x 33 LDS tiles, ai = omegas[a] when the angle list selects input matrices (omega_out 0),
// ao = omegas[a] when it selects output matrices (omega_out 1), a otherwise.
__global__ void __launch_bounds__(256) transpose_sum_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                            const int* __restrict__ omegas, int omega_out, int nr,
                                                            int nc, int64_t in_ld, int64_t in_mat, int64_t out_ld,
                                                            int64_t out_mat, int nsplit, int64_t split_stride,
                                                            int accumulate) {
  __shared__ float tile[32][33];
  const int a = blockIdx.z;
  const int ai = (omegas && !omega_out) ? omegas[a] : a;
  const int ao = (omegas && omega_out) ? omegas[a] : a;
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const float* src = in + (int64_t)ai * in_mat;
  float* dst = out + (int64_t)ao * out_mat;
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + tx;
    float v = 0.f;
    if (r < nr && c < nc) {
      const float* p = src + (int64_t)r * in_ld + c;
      for (int q = 0; q < nsplit; ++q) v += p[(int64_t)q * split_stride];
    }
    tile[k][tx] = v;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int c = c0 + k, r = r0 + tx;
    if (c < nc && r < nr) {
      float* p = dst + (int64_t)c * out_ld + r;
      *p = accumulate ? *p + tile[tx][k] : tile[tx][k];
    }
  }
}

template <int R, int YW>
void launch_spmm_t(dim3 grid, size_t lds, hipStream_t st, const float* inT, const int* row_ptr, const uint16_t* col,
                   const float* val, float* outT, const int* omegas, int my, int ys, int ncols, int nrows,
                   int rows_out, int compact, float alpha, int cspan, int kc, int n_angles, int64_t in_ld,
                   int64_t out_ld, int ys_out) {
  hipLaunchKernelGGL((spmm_t_kernel<R, YW, (YW <= 4 ? 8 : 4)>), grid, dim3(SPT_WAVES * 64), lds, st, inT, row_ptr,
                     col, val, outT, omegas, my, ys, ncols, nrows, rows_out, compact, alpha, cspan, kc, n_angles,
                     in_ld, out_ld, ys_out);
}

// ys: the LDS row stride (the transposed input's row stride when in_ld == 0); ys_out: the transposed output's
int spmm_t_impl(const float* in, int64_t in_ld, const int32_t* row_ptr, const uint16_t* col, const float* val,
                float* out, int64_t out_ld, int32_t ys_out, const int32_t* omegas, int32_t n_angles, int32_t my,
                int32_t ys, int32_t ncols, int32_t nrows, int32_t rows_out, int32_t compact, float alpha,
                int32_t nsplit, int32_t kc, hipStream_t st) {
  const int kc_max = (SPT_LDS_FLOATS - 64) / ys;
  if (kc <= 0 || kc > kc_max) kc = kc_max;
  int cspan = (ncols + nsplit - 1) / nsplit;
  if (in_ld) {                       // float4 staging of the untransposed operand: chunks start at multiples of 4
    kc &= ~3;
    cspan = (cspan + 3) & ~3;
  }
  if (n_angles == 0 || rows_out == 0) return DFM_OK;
  const int yw = (my + 63) / 64;
  const int R = my <= 128 ? 16 : 8;   // 16 rows of 3 y groups would spill with the value windows
  const dim3 grid((unsigned)((rows_out + SPT_WAVES * R - 1) / (SPT_WAVES * R)), (unsigned)n_angles,
                  (unsigned)nsplit);
  size_t lds = ((size_t)kc * ys + 64) * sizeof(float);
  if (out_ld) lds = std::max(lds, (size_t)64 * (SPT_WAVES * R + 1) * sizeof(float));
#define DFM_SPT(RR, YY)                                                                                      \
  launch_spmm_t<RR, YY>(grid, lds, st, in, row_ptr, col, val, out, omegas, my, ys, ncols, nrows, rows_out,  \
                        compact, alpha, cspan, kc, n_angles, in_ld, out_ld, ys_out)
  switch (yw) {
    case 1: DFM_SPT(16, 1); break;
    case 2: DFM_SPT(16, 2); break;
    case 3: DFM_SPT(8, 3); break;
    case 4: DFM_SPT(8, 4); break;
    case 5: DFM_SPT(8, 5); break;
    case 6: DFM_SPT(8, 6); break;
    case 7: DFM_SPT(8, 7); break;
    default: DFM_SPT(8, 8); break;
  }
#undef DFM_SPT
  return DFM_OK;
}

}  // namespace

extern "C" {

int dfm_spmm_t_rows_per_block(int32_t my) {
  if (my <= 0 || my > 512) return 0;
  return SPT_WAVES * (my <= 128 ? 16 : 8);
}

int dfm_spmm_t(const float* inT, const int32_t* row_ptr, const uint16_t* col, const float* val, float* outT,
               const int32_t* omegas, int32_t n_angles, int32_t my, int32_t ys, int32_t ncols, int32_t nrows,
               int32_t rows_out, int32_t compact, float alpha, int32_t nsplit, int32_t kc, dfm_stream_t stream) {
  DFM_CHECK_ARG(inT && row_ptr && col && val && outT, "dfm_spmm_t: null pointer");
  DFM_CHECK_ARG(n_angles >= 0 && my > 0 && my <= 512 && ys >= my && ys % 4 == 0 && ncols > 0 && ncols <= 65536 &&
                    nrows >= 0 && rows_out >= nrows && nsplit >= 1 && nsplit <= ncols && n_angles <= 65535,
                "dfm_spmm_t: bad sizes (my %d ys %d ncols %d nrows %d rows_out %d nsplit %d)", my, ys, ncols, nrows,
                rows_out, nsplit);
  if (((uintptr_t)inT) % 16) {
    dfm::set_error("dfm_spmm_t: the transposed operand must be 16-byte aligned");
    return DFM_ERR_ARG;
  }
  spmm_t_impl(inT, 0, row_ptr, col, val, outT, 0, ys, omegas, n_angles, my, ys, ncols, nrows, rows_out, compact,
              alpha, nsplit, kc, (hipStream_t)stream);
  DFM_LAUNCH_CHECK("dfm_spmm_t");
  return DFM_OK;
}

int dfm_spmm_t_direct(const float* in, int64_t in_ld, int32_t in_ys, const int32_t* row_ptr, const uint16_t* col,
                      const float* val, float* out, int64_t out_ld, int32_t out_ys, const int32_t* omegas,
                      int32_t n_angles, int32_t my, int32_t ncols, int32_t nrows, int32_t rows_out, int32_t compact,
                      float alpha, int32_t nsplit, int32_t kc, dfm_stream_t stream) {
  DFM_CHECK_ARG(in && row_ptr && col && val && out, "dfm_spmm_t_direct: null pointer");
  DFM_CHECK_ARG(n_angles >= 0 && my > 0 && my <= 512 && ncols > 0 && ncols <= 65536 && nrows >= 0 &&
                    rows_out >= nrows && nsplit >= 1 && nsplit <= ncols && n_angles <= 65535 &&
                    (in_ld ? in_ld >= ncols : (in_ys >= my && in_ys % 4 == 0)) &&
                    (out_ld ? (out_ld >= rows_out && nsplit == 1) : out_ys >= my),
                "dfm_spmm_t_direct: bad sizes (my %d ncols %d nrows %d rows_out %d nsplit %d in_ld %lld in_ys %d "
                "out_ld %lld out_ys %d)", my, ncols, nrows, rows_out, nsplit, (long long)in_ld, in_ys,
                (long long)out_ld, out_ys);
  if (((uintptr_t)in) % 16 || in_ld % 4) {
    dfm::set_error("dfm_spmm_t_direct: the input must be 16-byte aligned, in_ld a multiple of 4");
    return DFM_ERR_ARG;
  }
  const int ys = in_ld ? (my | 1) : in_ys;     // odd LDS row stride for the transposing staging
  spmm_t_impl(in, in_ld, row_ptr, col, val, out, out_ld, out_ys, omegas, n_angles, my, ys, ncols, nrows, rows_out,
              compact, alpha, nsplit, kc, (hipStream_t)stream);
  DFM_LAUNCH_CHECK("dfm_spmm_t_direct");
  return DFM_OK;
}

int dfm_transpose_sum(const float* in, float* out, const int32_t* omegas, int32_t omega_out, int32_t n_mat,
                      int32_t nr, int32_t nc, int64_t in_ld, int64_t in_mat, int64_t out_ld, int64_t out_mat,
                      int32_t nsplit, int64_t split_stride, int32_t accumulate, dfm_stream_t stream) {
  DFM_CHECK_ARG(in && out && n_mat >= 0 && n_mat <= 65535 && nr > 0 && nc > 0 && in_ld >= nc && out_ld >= nr &&
                    nsplit >= 1 && (nr + 31) / 32 <= 65535,
                "dfm_transpose_sum: bad arguments");
  if (n_mat == 0) return DFM_OK;
  const dim3 grid((unsigned)((nc + 31) / 32), (unsigned)((nr + 31) / 32), (unsigned)n_mat);
  hipLaunchKernelGGL(transpose_sum_kernel, grid, dim3(256), 0, (hipStream_t)stream, in, out, omegas, omega_out, nr,
                     nc, in_ld, in_mat, out_ld, out_mat, nsplit, split_stride, accumulate);
  DFM_LAUNCH_CHECK("dfm_transpose_sum");
  return DFM_OK;
}

}  // extern "C"
