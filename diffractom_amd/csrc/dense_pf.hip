// Pole-figure contraction on the matrix cores (dense PF mode): BF16 operands, FP32 accumulation, optional
// residual correction (hi/lo splitting of both operands).
//
// Forward, per rotation angle r:  data[r, y, j] (+)= sum_k sino[a, y, k] * PF[r, k, j]     (M = y, N = j, K = k)
// Adjoint, per rotation angle r:  sino[a, y, k]    = alpha * sum_j data[r, y, j] * PF[r, k, j]   (M = y, N = k, K = j)
// with a the position of r in the angle list (single_phase_forward_operator.py:2063-2127, CLBlast
// gemmStridedBatched in the reference's dense mode).
//
// The PF matrix is stored twice, each copy already in the operand order of v_mfma_f32_16x16x32_bf16, so every
// fragment is one coalesced 1 KiB load (lane l holds 8 consecutive bf16 of the contraction index):
//   forward fragment (r, nt, kc): lane l holds PF[k = 32 kc + 8 (l / 16) + i][j = 16 nt + l % 16], i = 0..7
//   adjoint fragment (r, kt, jc): lane l holds PF[k = 16 kt + l % 16][j = 32 jc + 8 (l / 16) + i]
// at element offset ((r * n_tiles + tile) * n_chunks + chunk) * 512 + 8 l.  Entries beyond K or C*P are zero.
// The y-side operand (sinogram or data, FP32 in HBM) is split into bf16 hi (+ lo) while it is staged in LDS. With the
// residual the product is hi*hi + hi*lo + lo*hi (the lo*lo term, 2^-16 relative, is dropped).
#include "dfm_common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MT_MAX = 8;   // 16-row y tiles held by one wave
constexpr int N_XCD = 8;

__device__ __forceinline__ f32x4 mma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Logical workgroup index such that indices [x * nb / 8, (x + 1) * nb / 8) run on XCD x (workgroup i is dispatched
// to XCD i % 8); nb = gridDim.x is a multiple of 8.
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t nb = gridDim.x, i = blockIdx.x;
  return (i % N_XCD) * (nb / N_XCD) + i / N_XCD;
}

// row[j0 .. j0 + 7], zero at and beyond n; vec: row is 16-byte aligned.
__device__ __forceinline__ void load8(const float* __restrict__ row, int j0, int n, bool vec, float v[8]) {
  if (vec && j0 + 8 <= n) {
    const float4 p = *reinterpret_cast<const float4*>(row + j0);
    const float4 q = *reinterpret_cast<const float4*>(row + j0 + 4);
    v[0] = p.x; v[1] = p.y; v[2] = p.z; v[3] = p.w;
    v[4] = q.x; v[5] = q.y; v[6] = q.z; v[7] = q.w;
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = (j0 + i < n) ? row[j0 + i] : 0.f;
  }
}

// v ~= hi + lo with hi = bf16(v) (round to nearest even) and lo = bf16(v - hi); v - hi is exact in FP32.
template <bool RES>
__device__ __forceinline__ void split8(const float v[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const __bf16 h = (__bf16)v[i];
    hi[i] = h;
    if (RES) lo[i] = (__bf16)(v[i] - (float)h);
  }
}

// ---------------------------------------------------------------------------------------------- packing
__global__ void pack_fwd_kernel(const float* __restrict__ dense, bf16x8* __restrict__ hi, bf16x8* __restrict__ lo,
                                int R, int km, int kb, int cp, int n_nt, int n_kc) {
  const int64_t total = (int64_t)R * n_nt * n_kc * 64;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(idx & 63);
    const int64_t f = idx >> 6;
    const int kc = (int)(f % n_kc);
    const int64_t t = f / n_kc;
    const int nt = (int)(t % n_nt);
    const int r = (int)(t / n_nt);
    const int j = nt * 16 + (lane & 15);
    const int k0 = kc * 32 + (lane >> 4) * 8;
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
      v[i] = (k0 + i < kb && j < cp) ? dense[((int64_t)r * km + k0 + i) * cp + j] : 0.f;
    bf16x8 h, l;
    split8<true>(v, h, l);
    hi[idx] = h;
    if (lo) lo[idx] = l;
  }
}

__global__ void pack_adj_kernel(const float* __restrict__ dense, bf16x8* __restrict__ hi, bf16x8* __restrict__ lo,
                                int R, int km, int kb, int cp, int n_kt, int n_jc) {
  const int64_t total = (int64_t)R * n_kt * n_jc * 64;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(idx & 63);
    const int64_t f = idx >> 6;
    const int jc = (int)(f % n_jc);
    const int64_t t = f / n_jc;
    const int kt = (int)(t % n_kt);
    const int r = (int)(t / n_kt);
    const int k = kt * 16 + (lane & 15);
    const int j0 = jc * 32 + (lane >> 4) * 8;
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
      v[i] = (k < kb && j0 + i < cp) ? dense[((int64_t)r * km + k) * cp + j0 + i] : 0.f;
    bf16x8 h, l;
    split8<true>(v, h, l);
    hi[idx] = h;
    if (lo) lo[idx] = l;
  }
}

// ---------------------------------------------------------------------------------------------- products
// One kernel for both directions. A workgroup (4 waves) owns (angle a, k split ks, y group yg) and 4 NT tiles of the
// output columns (NT per wave). Per contraction chunk (32 orientations forward, 32 segments adjoint) the y-side
// operand (mt_n 16-row tiles x 32, FP32 in HBM) is loaded once by the workgroup, split into bf16 hi (+ lo) and
// stored in LDS in fragment order (double-buffered); each wave reads its A fragments from LDS (one ds_read_b128 per
// lane) and streams its PF fragments from HBM one chunk ahead.
// The same layout as dfm_mfma_desc of the C ABI (the grouped launch reads an array of them).
struct MfmaArgs {
  const float* src;        // forward: sino (n_ang, my, sks); adjoint: data (R, my, cp)
  const bf16x8* hi;
  const bf16x8* lo;
  const int* omegas;
  float* out;              // forward: data (R, my, cp) or the partial sums (ksplit, n_ang, my, cp); adjoint: sino
  int n_ang, my, cp, sks, kb;
  int n_tiles, n_chunks;   // forward: j tiles of 16, k chunks of 32; adjoint: k tiles of 16, j chunks of 32
  int mt_n, n_yg, n_tg;    // y tiles per group, y groups, groups of 4 NT output tiles
  int ksplit, ch_per;      // chunks split over ksplit workgroups (forward)
  int direct, accumulate;
  float alpha;
  int ncols;               // adjoint: sinogram columns written (< ncols; the row stride is sks)
};
static_assert(sizeof(MfmaArgs) == sizeof(dfm_mfma_desc), "MfmaArgs and dfm_mfma_desc differ");

// Workgroup w of one product (the decomposition below); lds_hi / lds_lo: 2 * MT_MAX * 64 fragments each.
template <int NT, bool RES, bool FWD>
__device__ __forceinline__ void mfma_pf_body(const MfmaArgs& p, int64_t w, bf16x8* __restrict__ lds_hi,
                                             bf16x8* __restrict__ lds_lo) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int tg = (int)(w % p.n_tg);
  w /= p.n_tg;
  const int yg = (int)(w % p.n_yg);
  w /= p.n_yg;
  const int ks = (int)(w % p.ksplit);
  w /= p.ksplit;
  if (w >= p.n_ang) return;                            // uniform over the workgroup
  const int a = (int)w;
  const int r = p.omegas ? p.omegas[a] : a;
  const int t0 = (tg * 4 + (tid >> 6)) * NT;           // this wave's first output tile
  const int y0 = yg * p.mt_n * 16;
  const int c0 = ks * p.ch_per;
  const int c1 = min(p.n_chunks, c0 + p.ch_per);
  if (c0 >= c1) return;                                // uniform

  // y-side operand slots of this thread: s = tid, tid + 256 < mt_n * 64 (fragment mt = s / 64, lane s % 64)
  const int n_slots = p.mt_n * 64;
  const float* srow[2];
  bool sok[2];
  int scol[2];
  const int64_t src_rows = FWD ? (int64_t)a * p.my : (int64_t)r * p.my;
  const int src_stride = FWD ? p.sks : p.cp;
  const int src_n = FWD ? p.kb : p.cp;
  const bool vec = (src_stride & 3) == 0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int s = tid + 256 * i;
    const int y = y0 + (s >> 6) * 16 + (s & 15);
    sok[i] = s < n_slots && y < p.my;
    srow[i] = p.src + (src_rows + (sok[i] ? y : 0)) * src_stride;
    scol[i] = ((s & 63) >> 4) * 8;
  }
  // PF fragment streams of this wave's tiles
  const bf16x8* bh_p[NT];
  const bf16x8* bl_p[NT];
  bool tok[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    tok[t] = t0 + t < p.n_tiles;
    const int64_t base = (((int64_t)r * p.n_tiles + (tok[t] ? t0 + t : 0)) * p.n_chunks) * 64 + lane;
    bh_p[t] = p.hi + base;
    bl_p[t] = RES ? p.lo + base : p.hi;
  }
  const bf16x8 zero8 = {};

  float raw[2][8];
  auto load_raw = [&](int c) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      if (sok[i]) {
        load8(srow[i], c * 32 + scol[i], src_n, vec, raw[i]);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) raw[i][e] = 0.f;
      }
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int s = tid + 256 * i;
      if (s < n_slots) {
        bf16x8 h, l;
        split8<RES>(raw[i], h, l);
        lds_hi[buf * MT_MAX * 64 + s] = h;
        if (RES) lds_lo[buf * MT_MAX * 64 + s] = l;
      }
    }
  };
  // PF fragments of chunk c (bh, bl), c + 1 (nh, nl) and, with AHEAD2, c + 2 (mh, ml) in flight: at 2 output tiles
  // per wave (146 VGPRs, 3 waves per SIMD) the second chunk ahead keeps the occupancy and doubles the bytes in flight
  constexpr bool AHEAD2 = NT <= 2;
  bf16x8 bh[NT], bl[NT], nh[NT], nl[NT], mh[NT], ml[NT];
  auto load_b = [&](int c, bf16x8 (&h)[NT], bf16x8 (&l)[NT]) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      h[t] = tok[t] ? __builtin_nontemporal_load(bh_p[t] + (int64_t)c * 64) : zero8;
      l[t] = (RES && tok[t]) ? __builtin_nontemporal_load(bl_p[t] + (int64_t)c * 64) : zero8;
    }
  };

  f32x4 acc[MT_MAX][NT];
#pragma unroll
  for (int mt = 0; mt < MT_MAX; ++mt)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[mt][t] = f32x4{0.f, 0.f, 0.f, 0.f};

  load_raw(c0);
  load_b(c0, bh, bl);
  store_lds(0);
  if (AHEAD2 && c0 + 1 < c1) load_b(c0 + 1, nh, nl);
  __syncthreads();
  int cur = 0;
  for (int c = c0; c < c1; ++c) {
    const bool more = c + 1 < c1;
    if (more) {
      load_raw(c + 1);
      if (AHEAD2) {
        if (c + 2 < c1) load_b(c + 2, mh, ml);
      } else {
        load_b(c + 1, nh, nl);
      }
    }
#pragma unroll
    for (int mt = 0; mt < MT_MAX; ++mt) {
      if (mt < p.mt_n) {
        const bf16x8 ah = lds_hi[cur * MT_MAX * 64 + mt * 64 + lane];
        const bf16x8 al = RES ? lds_lo[cur * MT_MAX * 64 + mt * 64 + lane] : zero8;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          acc[mt][t] = mma(ah, bh[t], acc[mt][t]);
          if (RES) {
            acc[mt][t] = mma(ah, bl[t], acc[mt][t]);
            acc[mt][t] = mma(al, bh[t], acc[mt][t]);
          }
        }
      }
    }
    if (more) store_lds(cur ^ 1);
    __syncthreads();
    if (more) {
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        bh[t] = nh[t];
        bl[t] = nl[t];
        if (AHEAD2) {
          nh[t] = mh[t];
          nl[t] = ml[t];
        }
      }
    }
    cur ^= 1;
  }

  // D fragment: lane l holds rows 4 (l / 16) + i, column l % 16 of each 16 x 16 tile
  const int row = lane & 15, grp = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < MT_MAX; ++mt) {
    if (mt >= p.mt_n) continue;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (!tok[t]) continue;
      const int col = (t0 + t) * 16 + row;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int y = y0 + mt * 16 + grp * 4 + i;
        if (y >= p.my) continue;
        const float v = acc[mt][t][i];
        if (FWD) {
          if (col >= p.cp) continue;
          if (p.direct) {
            float* d = p.out + ((int64_t)r * p.my + y) * p.cp + col;
            *d = p.accumulate ? *d + v : v;
          } else {
            p.out[(((int64_t)ks * p.n_ang + a) * p.my + y) * p.cp + col] = v;
          }
        } else {
          if (col < p.ncols) p.out[((int64_t)a * p.my + y) * p.sks + col] = p.alpha * v;
        }
      }
    }
  }
}

template <int NT, bool RES, bool FWD>
__global__ __launch_bounds__(256, 2) void mfma_pf_kernel(MfmaArgs p) {
  __shared__ bf16x8 lds_hi[2 * MT_MAX * 64];
  __shared__ bf16x8 lds_lo[RES ? 2 * MT_MAX * 64 : 1];
  mfma_pf_body<NT, RES, FWD>(p, xcd_block(), lds_hi, lds_lo);
}

// Grouped launch: workgroup w runs workgroup wg_map[w].y of product wg_map[w].x (products of different K, PF, column
// offsets and data rows in one launch; entries with x < 0 pad the grid).
template <int NT, bool RES, bool FWD>
__global__ __launch_bounds__(256, 2) void mfma_pf_grouped_kernel(const MfmaArgs* __restrict__ descs,
                                                                 const int2* __restrict__ wg_map, int64_t n_map) {
  __shared__ bf16x8 lds_hi[2 * MT_MAX * 64];
  __shared__ bf16x8 lds_lo[RES ? 2 * MT_MAX * 64 : 1];
  const int64_t w = xcd_block();
  if (w >= n_map) return;
  const int2 m = wg_map[w];
  if (m.x < 0) return;
  const MfmaArgs p = descs[m.x];
  mfma_pf_body<NT, RES, FWD>(p, m.y, lds_hi, lds_lo);
}

// Sum of the k-split partial results into the data rows of the angles (fixed order: deterministic).
__global__ void splitk_reduce_kernel(const float* __restrict__ part, float* __restrict__ data,
                                     const int* __restrict__ omegas, int n_ang, int64_t row_len, int ksplit,
                                     int accumulate) {
  const int64_t total = (int64_t)n_ang * row_len;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int ks = 0; ks < ksplit; ++ks) s += part[(int64_t)ks * total + i];
    const int a = (int)(i / row_len);
    const int64_t rem = i - (int64_t)a * row_len;
    const int r = omegas ? omegas[a] : a;
    float* d = data + (int64_t)r * row_len + rem;
    *d = accumulate ? *d + s : s;
  }
}

template <bool FWD>
void launch_grouped(const MfmaArgs* descs, const int2* wg_map, int64_t n_map, int variant, bool res,
                    hipStream_t st) {
  const dim3 grid((unsigned)((n_map + N_XCD - 1) / N_XCD * N_XCD)), block(256);
  if (variant == 4) {
    if (res) hipLaunchKernelGGL((mfma_pf_grouped_kernel<4, true, FWD>), grid, block, 0, st, descs, wg_map, n_map);
    else hipLaunchKernelGGL((mfma_pf_grouped_kernel<4, false, FWD>), grid, block, 0, st, descs, wg_map, n_map);
  } else {
    if (res) hipLaunchKernelGGL((mfma_pf_grouped_kernel<2, true, FWD>), grid, block, 0, st, descs, wg_map, n_map);
    else hipLaunchKernelGGL((mfma_pf_grouped_kernel<2, false, FWD>), grid, block, 0, st, descs, wg_map, n_map);
  }
}

template <bool FWD>
void launch_mfma(const MfmaArgs& p, int variant, bool res, int64_t n_wg, hipStream_t st) {
  const dim3 grid((unsigned)((n_wg + N_XCD - 1) / N_XCD * N_XCD)), block(256);
  if (variant == 4) {
    if (res) hipLaunchKernelGGL((mfma_pf_kernel<4, true, FWD>), grid, block, 0, st, p);
    else hipLaunchKernelGGL((mfma_pf_kernel<4, false, FWD>), grid, block, 0, st, p);
  } else {
    if (res) hipLaunchKernelGGL((mfma_pf_kernel<2, true, FWD>), grid, block, 0, st, p);
    else hipLaunchKernelGGL((mfma_pf_kernel<2, false, FWD>), grid, block, 0, st, p);
  }
}

// y tiling shared by both directions: n_yg groups of mt_n 16-row tiles (mt_n <= MT_MAX).
void y_tiling(int my, int& mt_n, int& n_yg) {
  const int tiles = (my + 15) / 16;
  n_yg = (tiles + MT_MAX - 1) / MT_MAX;
  mt_n = (tiles + n_yg - 1) / n_yg;
}


// The FP32 dense PF products (pf_mode "dense", pf_precision "fp32"; single_phase_forward_operator.py:2063-2127
// gemmStridedBatched), one angle per blockIdx.z: C[y][c] (+)= alpha sum_r A[y][r] B(r, c) with A row-major (lda),
// B(r, c) = NT ? Bm[c ldb + r] : Bm[r ldb + c], C row-major (ldc). The forward is NN (sino x PF: B's rows are the
// orientations), the adjoint NT (data x PF^T: the PF rows read along their segments), both straight from the stored
// (R, kstride, CP) batch: no transposed or angle-gathered copy of it. A 64 x 64 output tile per 256-thread workgroup,
// 4 x 4 outputs per thread, the reduction in chunks of 16 staged in LDS (transposed for A, so both operands are read
// as float4 rows of the chunk); FP32 FMA in a fixed order.
template <bool NT>
__global__ void __launch_bounds__(256) dense_f32_kernel(const float* __restrict__ A, int64_t a_sa, int lda, int a_om,
                                                        const float* __restrict__ Bm, int64_t b_sa, int ldb,
                                                        float* __restrict__ C, int64_t c_sa, int ldc, int c_om,
                                                        const int* __restrict__ omegas, int n_ang, int M, int N,
                                                        int R, int r_len, float alpha, int accumulate,
                                                        float* __restrict__ partial) {
  __shared__ float4 As[16][16 + 1];   // As[r][y / 4]: rows y of the tile, chunk column r
  __shared__ float4 Bs[16][16 + 1];   // Bs[r][c / 4]
  // blockIdx.z = split * n_ang + as: the split's reduction range r_len * split .. (partial sums with a split)
  const int as = blockIdx.z % n_ang, split = blockIdx.z / n_ang;
  const int a = omegas ? omegas[as] : as;
  const int r_beg = split * r_len, r_end = min(R, r_beg + r_len);
  const float* Ap = A + (a_om ? a : as) * a_sa;
  const float* Bp = Bm + (int64_t)a * b_sa;
  float* Cp = C + (c_om ? a : as) * c_sa;
  const int y0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const int t = threadIdx.x;
  const int ty = t >> 4, tx = t & 15;
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
  float* Asf = reinterpret_cast<float*>(As);
  float* Bsf = reinterpret_cast<float*>(Bs);
  for (int r0 = r_beg; r0 < r_end; r0 += 16) {
    // A chunk: row y = y0 + t / 4, columns r0 + 4 (t % 4) .. + 3, stored transposed
    {
      const int yl = t >> 2, rl = (t & 3) * 4;
      const int y = y0 + yl;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = r0 + rl + u;
        Asf[(rl + u) * (17 * 4) + yl] = (y < M && r < r_end) ? Ap[(int64_t)y * lda + r] : 0.0f;
      }
    }
    // B chunk: Bs[r][c]
    if (NT) {
      const int cl = t >> 2, rl = (t & 3) * 4;
      const int c = c0 + cl;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = r0 + rl + u;
        Bsf[(rl + u) * (17 * 4) + cl] = (c < N && r < r_end) ? Bp[(int64_t)c * ldb + r] : 0.0f;
      }
    } else {
      const int rl = t >> 4, cl = (t & 15) * 4;
      const int r = r0 + rl;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int c = c0 + cl + u;
        Bsf[rl * (17 * 4) + cl + u] = (c < N && r < r_end) ? Bp[(int64_t)r * ldb + c] : 0.0f;
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const float4 av = As[r][ty], bv = Bs[r][tx];
      const float ai[4] = {av.x, av.y, av.z, av.w}, bj[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(ai[i], bj[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int y = y0 + ty * 4 + i;
    if (y >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = c0 + tx * 4 + j;
      if (c >= N) continue;
      if (partial) {
        partial[((int64_t)blockIdx.z * M + y) * N + c] = acc[i][j];
      } else {
        float* o = Cp + (int64_t)y * ldc + c;
        *o = accumulate ? *o + alpha * acc[i][j] : alpha * acc[i][j];
      }
    }
  }
}

// C (+)= alpha sum over the splits (in order) of the partial sums of dense_f32_kernel.
__global__ void __launch_bounds__(256) dense_f32_reduce_kernel(const float* __restrict__ partial, int n_split,
                                                               float* __restrict__ C, int64_t c_sa, int ldc, int c_om,
                                                               const int* __restrict__ omegas, int n_ang, int M,
                                                               int N, float alpha, int accumulate) {
  const int64_t per = (int64_t)M * N, total = per * n_ang;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int as = (int)(i / per);
    const int64_t e = i - (int64_t)as * per;
    const int y = (int)(e / N), c = (int)(e - (int64_t)y * N);
    float v = 0.0f;
    for (int sp = 0; sp < n_split; ++sp) v += partial[((int64_t)sp * n_ang + as) * per + e];
    const int a = omegas ? omegas[as] : as;
    float* o = C + (c_om ? a : as) * c_sa + (int64_t)y * ldc + c;
    *o = accumulate ? *o + alpha * v : alpha * v;
  }
}
}  // namespace

namespace dfm {
int launch_splitk_reduce(const float* part, float* data, const int* omegas, int n_ang, int64_t row_len, int ksplit,
                         int accumulate, hipStream_t st) {
  const int64_t total = (int64_t)n_ang * row_len;
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3(grid_stride_blocks(total, 256, 1u << 16)), dim3(256), 0, st, part, data,
                     omegas, n_ang, row_len, ksplit, accumulate);
  DFM_LAUNCH_CHECK("k-split reduce");
  return DFM_OK;
}
}  // namespace dfm

extern "C" {

int dfm_dense_f32(int32_t transpose_b, const float* a, int64_t a_stride, int32_t lda, int32_t a_by_angle,
                  const float* b, int64_t b_stride, int32_t ldb, float* c, int64_t c_stride, int32_t ldc,
                  int32_t c_by_angle, const int32_t* omegas, int32_t n_ang, int32_t m, int32_t n, int32_t r,
                  float alpha, int32_t accumulate, int32_t r_split, float* partial, dfm_stream_t stream) {
  DFM_CHECK_ARG(a && b && c && n_ang >= 0 && m >= 0 && n >= 0 && r >= 0 && lda >= 0 && ldb >= 0 && ldc >= n &&
                    a_stride >= 0 && b_stride >= 0 && c_stride >= 0 && r_split >= 1 && (r_split == 1 || partial) &&
                    (int64_t)n_ang * r_split < 65536,
                "dfm_dense_f32: bad arguments");
  if (n_ang == 0 || m == 0 || n == 0) return DFM_OK;
  const int r_len = r_split == 1 ? (r > 0 ? r : 1) : ((r + r_split - 1) / r_split + 15) / 16 * 16;
  const int n_split = r_split == 1 ? 1 : (r + r_len - 1) / r_len;
  float* part = n_split > 1 ? partial : nullptr;
  const dim3 grid((unsigned)((n + 63) / 64), (unsigned)((m + 63) / 64), (unsigned)(n_ang * n_split));
  const hipStream_t st = (hipStream_t)stream;
  if (transpose_b)
    hipLaunchKernelGGL(dense_f32_kernel<true>, grid, dim3(256), 0, st, a, a_stride, lda, a_by_angle, b, b_stride, ldb,
                       c, c_stride, ldc, c_by_angle, omegas, n_ang, m, n, r, r_len, alpha, accumulate, part);
  else
    hipLaunchKernelGGL(dense_f32_kernel<false>, grid, dim3(256), 0, st, a, a_stride, lda, a_by_angle, b, b_stride,
                       ldb, c, c_stride, ldc, c_by_angle, omegas, n_ang, m, n, r, r_len, alpha, accumulate, part);
  if (part)
    hipLaunchKernelGGL(dense_f32_reduce_kernel, dim3(dfm::grid_stride_blocks((int64_t)n_ang * m * n, 256)),
                       dim3(256), 0, st, part, n_split, c, c_stride, ldc, c_by_angle, omegas, n_ang, m, n, alpha,
                       accumulate);
  DFM_LAUNCH_CHECK("dfm_dense_f32");
  return DFM_OK;
}

int dfm_mfma_pack_fwd(const float* dense, void* hi, void* lo, int R, int km, int kb, int cp, int n_nt, int n_kc,
                      dfm_stream_t stream) {
  DFM_CHECK_ARG(dense && hi, "dfm_mfma_pack_fwd: null pointer");
  DFM_CHECK_ARG(R > 0 && kb > 0 && km >= kb && cp > 0, "dfm_mfma_pack_fwd: bad sizes");
  DFM_CHECK_ARG(n_nt * 16 >= cp && n_kc * 32 >= kb, "dfm_mfma_pack_fwd: %d x %d fragments do not cover %d x %d",
                n_kc, n_nt, kb, cp);
  const int64_t total = (int64_t)R * n_nt * n_kc * 64;
  hipLaunchKernelGGL(pack_fwd_kernel, dim3(dfm::grid_stride_blocks(total, 256, 1u << 20)), dim3(256), 0,
                     (hipStream_t)stream, dense, (bf16x8*)hi, (bf16x8*)lo, R, km, kb, cp, n_nt, n_kc);
  DFM_LAUNCH_CHECK("dfm_mfma_pack_fwd");
  return DFM_OK;
}

int dfm_mfma_pack_adj(const float* dense, void* hi, void* lo, int R, int km, int kb, int cp, int n_kt, int n_jc,
                      dfm_stream_t stream) {
  DFM_CHECK_ARG(dense && hi, "dfm_mfma_pack_adj: null pointer");
  DFM_CHECK_ARG(R > 0 && kb > 0 && km >= kb && cp > 0, "dfm_mfma_pack_adj: bad sizes");
  DFM_CHECK_ARG(n_kt * 16 >= kb && n_jc * 32 >= cp, "dfm_mfma_pack_adj: %d x %d fragments do not cover %d x %d",
                n_kt, n_jc, kb, cp);
  const int64_t total = (int64_t)R * n_kt * n_jc * 64;
  hipLaunchKernelGGL(pack_adj_kernel, dim3(dfm::grid_stride_blocks(total, 256, 1u << 20)), dim3(256), 0,
                     (hipStream_t)stream, dense, (bf16x8*)hi, (bf16x8*)lo, R, km, kb, cp, n_kt, n_jc);
  DFM_LAUNCH_CHECK("dfm_mfma_pack_adj");
  return DFM_OK;
}

int dfm_mfma_forward(const float* sino, int sks, int kb, const void* hi, const void* lo, const int* omegas,
                     int n_ang, int my, int cp, int n_nt, int n_kc, float* data, float* part, int ksplit,
                     int accumulate, int variant, dfm_stream_t stream) {
  DFM_CHECK_ARG(sino && hi && data, "dfm_mfma_forward: null pointer");
  DFM_CHECK_ARG(n_ang > 0 && my > 0 && cp > 0 && kb > 0 && sks >= kb && (sks & 3) == 0,
                "dfm_mfma_forward: bad sizes (n_ang %d, My %d, CP %d, Kb %d, stride %d)", n_ang, my, cp, kb, sks);
  DFM_CHECK_ARG(n_nt * 16 >= cp && n_kc * 32 >= kb, "dfm_mfma_forward: fragments do not cover the matrix");
  DFM_CHECK_ARG(ksplit >= 1 && ksplit <= n_kc, "dfm_mfma_forward: k split %d not in 1 .. %d", ksplit, n_kc);
  DFM_CHECK_ARG(ksplit == 1 || part, "dfm_mfma_forward: a k split needs the partial-sum workspace");
  DFM_CHECK_ARG(variant == 2 || variant == 4, "dfm_mfma_forward: variant (j tiles per wave) must be 2 or 4");
  MfmaArgs p{};
  p.src = sino;
  p.hi = (const bf16x8*)hi;
  p.lo = (const bf16x8*)lo;
  p.omegas = omegas;
  p.n_ang = n_ang;
  p.my = my;
  p.cp = cp;
  p.sks = sks;
  p.kb = kb;
  p.n_tiles = n_nt;
  p.n_chunks = n_kc;
  y_tiling(my, p.mt_n, p.n_yg);
  p.n_tg = (n_nt + 4 * variant - 1) / (4 * variant);
  p.ch_per = (n_kc + ksplit - 1) / ksplit;
  p.ksplit = (n_kc + p.ch_per - 1) / p.ch_per;     // every split holds at least one chunk
  p.direct = p.ksplit == 1;
  p.accumulate = accumulate;
  p.out = p.direct ? data : part;
  p.alpha = 1.f;
  hipStream_t st = (hipStream_t)stream;
  launch_mfma<true>(p, variant, lo != nullptr, (int64_t)n_ang * p.ksplit * p.n_yg * p.n_tg, st);
  DFM_LAUNCH_CHECK("dfm_mfma_forward");
  if (!p.direct) {
    const int64_t total = (int64_t)n_ang * my * cp;
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(dfm::grid_stride_blocks(total, 256, 1u << 16)), dim3(256), 0, st,
                       part, data, omegas, n_ang, (int64_t)my * cp, p.ksplit, accumulate);
    DFM_LAUNCH_CHECK("dfm_mfma_forward (reduce)");
  }
  return DFM_OK;
}

int dfm_mfma_adjoint(const float* data, const void* hi, const void* lo, const int* omegas, int n_ang, int my, int cp,
                     int n_kt, int n_jc, float* sino, int sks, float alpha, int variant, dfm_stream_t stream) {
  DFM_CHECK_ARG(data && hi && sino, "dfm_mfma_adjoint: null pointer");
  DFM_CHECK_ARG(n_ang > 0 && my > 0 && cp > 0 && sks > 0 && (sks & 3) == 0 && n_kt * 16 >= sks,
                "dfm_mfma_adjoint: bad sizes (n_ang %d, My %d, CP %d, stride %d, k tiles %d)", n_ang, my, cp, sks,
                n_kt);
  DFM_CHECK_ARG(n_jc * 32 >= cp, "dfm_mfma_adjoint: fragments do not cover the matrix");
  DFM_CHECK_ARG(variant == 2 || variant == 4, "dfm_mfma_adjoint: variant (k tiles per wave) must be 2 or 4");
  MfmaArgs p{};
  p.src = data;
  p.hi = (const bf16x8*)hi;
  p.lo = (const bf16x8*)lo;
  p.omegas = omegas;
  p.out = sino;
  p.n_ang = n_ang;
  p.my = my;
  p.cp = cp;
  p.sks = sks;
  p.kb = sks;
  p.n_tiles = n_kt;
  p.n_chunks = n_jc;
  y_tiling(my, p.mt_n, p.n_yg);
  p.n_tg = (n_kt + 4 * variant - 1) / (4 * variant);
  p.ksplit = 1;
  p.ch_per = n_jc;
  p.alpha = alpha;
  p.ncols = sks;
  launch_mfma<false>(p, variant, lo != nullptr, (int64_t)n_ang * p.n_yg * p.n_tg, (hipStream_t)stream);
  DFM_LAUNCH_CHECK("dfm_mfma_adjoint");
  return DFM_OK;
}

int dfm_mfma_grouped(int32_t forward, const dfm_mfma_desc* descs, const int32_t* wg_map, int64_t n_wg,
                     int32_t residual, int32_t variant, dfm_stream_t stream) {
  DFM_CHECK_ARG(descs && wg_map && n_wg > 0, "dfm_mfma_grouped: null pointer or no workgroups");
  DFM_CHECK_ARG(variant == 2 || variant == 4, "dfm_mfma_grouped: variant (tiles per wave) must be 2 or 4");
  DFM_CHECK_ARG(n_wg <= ((int64_t)1 << 31) - 8, "dfm_mfma_grouped: %lld workgroups", (long long)n_wg);
  if (forward)
    launch_grouped<true>((const MfmaArgs*)descs, (const int2*)wg_map, n_wg, variant, residual != 0,
                         (hipStream_t)stream);
  else
    launch_grouped<false>((const MfmaArgs*)descs, (const int2*)wg_map, n_wg, variant, residual != 0,
                          (hipStream_t)stream);
  DFM_LAUNCH_CHECK("dfm_mfma_grouped");
  return DFM_OK;
}

}  // extern "C"
