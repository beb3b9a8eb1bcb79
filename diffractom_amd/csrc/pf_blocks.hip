// Pole-figure contraction on the matrix cores over a block-sparse PF (the sparse PF mode with 16-bit operands).
//
// Forward, per rotation angle r:  data[r, y, j] (+)= sum_k sino[a, y, k] * PF[r, k, j]     (M = y, N = j, K = k)
// Adjoint, per rotation angle r:  sino[a, y, k]    = alpha * sum_j data[r, y, j] * PF[r, k, j]   (M = y, N = k, K = j)
// (single_phase_forward_operator.py:38-123 spmm_pf_forward_sub / spmm_pf_adjoint_sub, pf_kernels.cl:597-687).
//
// Storage (one copy, read by both directions): the PF of an orientation batch is cut into 32 x 32 blocks
// (angle r, orientation block kb of 32, segment block jb of 32), block index (r * n_kb + kb) * n_jb + jb, and each
// block lists its non-zeros: ptr[block] .. ptr[block + 1] - 1 into pos (uint16: k % 32 in bits 0-4, j % 32 in bits
// 5-9) and val. Value encodings: bf16 pairs (uint32: bf16(v) in the low half, bf16(v - hi) in the high half: the
// hi/lo residual split, exact in FP32), or fp16 (uint16) of v times a power-of-two scale of its ring. At 6 % fill a
// non-zero costs 6 (bf16 pairs) or 4 (fp16) bytes of HBM instead of the 2 x 32 x 32 x 2 bytes per block of dense
// fragments.
//
// A wave owns one column block (32 output columns: two 16-wide MFMA tiles) and walks the contraction in chunks of
// 32. Per chunk it expands its block into dense MFMA fragments in a private LDS area (clear, scatter the block's
// entries, read back in operand order: lane l holds 8 consecutive contraction indices of column l % 16), while the
// workgroup's four waves share the y-side operand chunk (FP32 from HBM, split into bf16 hi/lo or scaled to fp16 in
// LDS, double buffered). v_mfma_f32_16x16x32_{bf16,f16}, FP32 accumulation; the residual mode adds hi*hi + hi*lo +
// lo*hi (lo*lo, 2^-16 relative, dropped). In the fp16 mode every (row, chunk) of the y-side operand is scaled by a
// power of two that puts its largest magnitude at 2^14 (fp16 keeps 11 significant bits from 2^-14 up), and the
// product of the chunk is scaled back before it is added to the FP32 accumulator; the ring scale of the PF is
// undone on the output columns (forward) or folded into the staged data (adjoint).
#include "dfm_common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MT_MAX = 8;       // 16-row y tiles held by one wave
constexpr int N_XCD = 8;
constexpr int WAVES = 4;        // waves per workgroup (one column block each)
constexpr int THREADS = 64 * WAVES;

enum { MODE_BF16 = 0, MODE_BF16_RES = 1, MODE_F16 = 2 };

__device__ __forceinline__ int64_t xcd_block() {
  const int64_t nb = gridDim.x, i = blockIdx.x;
  return (i % N_XCD) * (nb / N_XCD) + i / N_XCD;
}

__device__ __forceinline__ f32x4 mma_bf16(const s16x8& a, const s16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}
__device__ __forceinline__ f32x4 mma_f16(const s16x8& a, const s16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0,
                                                0);
}

__device__ __forceinline__ unsigned short bf16_bits(float v) { return __builtin_bit_cast(unsigned short, (__bf16)v); }
__device__ __forceinline__ unsigned short f16_bits(float v) { return __builtin_bit_cast(unsigned short, (_Float16)v); }

// ---------------------------------------------------------------------------------------------- build
// One wave per block (r, kb, jb); element e = 64 i + lane of the block is (k = 32 kb + e / 32, j = 32 jb + e % 32),
// so each load instruction reads two 128-byte rows of the dense batch.
__global__ void __launch_bounds__(256) blk_count_kernel(const float* __restrict__ dense, int* __restrict__ counts,
                                                        int R, int km, int kb, int cp, int n_kb, int n_jb) {
  const int64_t nblk = (int64_t)R * n_kb * n_jb;
  const int lane = threadIdx.x & 63;
  for (int64_t blk = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; blk < nblk;
       blk += ((int64_t)gridDim.x * blockDim.x) >> 6) {
    const int jb = (int)(blk % n_jb);
    const int64_t t = blk / n_jb;
    const int kbi = (int)(t % n_kb);
    const int r = (int)(t / n_kb);
    int c = 0;
#pragma unroll 4
    for (int i = 0; i < 16; ++i) {
      const int e = 64 * i + lane;
      const int k = 32 * kbi + (e >> 5), j = 32 * jb + (e & 31);
      if (k < kb && j < cp && dense[((int64_t)r * km + k) * cp + j] != 0.f) ++c;
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if (lane == 0) counts[blk] = c;
  }
}

// Entries in element order (deterministic). mode MODE_F16: val16 = fp16(v * ring_scale[j % P]); else val32.
__global__ void __launch_bounds__(256) blk_fill_kernel(const float* __restrict__ dense, const int* __restrict__ ptr,
                                                       unsigned short* __restrict__ pos, unsigned* __restrict__ val32,
                                                       unsigned short* __restrict__ val16,
                                                       const float* __restrict__ ring_scale, int P, int R, int km,
                                                       int kb, int cp, int n_kb, int n_jb) {
  const int64_t nblk = (int64_t)R * n_kb * n_jb;
  const int lane = threadIdx.x & 63;
  for (int64_t blk = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; blk < nblk;
       blk += ((int64_t)gridDim.x * blockDim.x) >> 6) {
    const int jb = (int)(blk % n_jb);
    const int64_t t = blk / n_jb;
    const int kbi = (int)(t % n_kb);
    const int r = (int)(t / n_kb);
    int base = ptr[blk];
    for (int i = 0; i < 16; ++i) {
      const int e = 64 * i + lane;
      const int kl = e >> 5, jl = e & 31;
      const int k = 32 * kbi + kl, j = 32 * jb + jl;
      const float v = (k < kb && j < cp) ? dense[((int64_t)r * km + k) * cp + j] : 0.f;
      const unsigned long long m = __ballot(v != 0.f);
      if (v != 0.f) {
        const int o = base + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        pos[o] = (unsigned short)(kl | (jl << 5));
        if (val16) {
          val16[o] = f16_bits(v * ring_scale[j % P]);
        } else {
          const __bf16 h = (__bf16)v;
          const __bf16 l = (__bf16)(v - (float)h);
          val32[o] = (unsigned)__builtin_bit_cast(unsigned short, h) |
                     ((unsigned)__builtin_bit_cast(unsigned short, l) << 16);
        }
      }
      base += __popcll(m);
    }
  }
}

// ---------------------------------------------------------------------------------------------- products
struct BlkArgs {
  const float* src;         // forward: sino (n_ang, my, sks); adjoint: data (R, my, cp)
  const int* ptr;           // block pointers of the batch (R * n_kb * n_jb + 1)
  const unsigned short* pos;
  const void* val;          // uint32 bf16 pairs or uint16 fp16
  const int* omegas;
  const float* ring_inv;    // fp16 mode: 1 / ring scale per ring (P values)
  float* out;               // forward: data or the partial sums (ksplit, n_ang, my, cp); adjoint: sino
  int n_ang, my, cp, sks, kb, P;
  int n_kb, n_jb;           // stored blocks
  int n_cb, n_ch;           // column blocks of the output (waves) and contraction chunks
  int mt_n, n_yg, n_tg;     // y tiles per group, y groups, groups of WAVES column blocks
  int ksplit, ch_per;
  int direct, accumulate;
  float alpha;
  int ncols;                // adjoint: sinogram columns written
  int blk_local;            // blocks indexed by the position a in the angle list (generated PF) instead of omegas[a]
};

static_assert(sizeof(BlkArgs) == sizeof(dfm_pfblk_desc), "BlkArgs and dfm_pfblk_desc differ");

// Forward fragment position of entry (kl, jl) in the wave's two-tile area (elements of 16 bits):
// tile jl / 16, lane (jl % 16) + 16 (kl / 8), element kl % 8; adjoint: tile kl / 16, lane (kl % 16) + 16 (jl / 8),
// element jl % 8.
template <bool FWD>
__device__ __forceinline__ int frag_slot(unsigned p) {
  const int kl = p & 31, jl = (p >> 5) & 31;
  const int c = FWD ? jl : kl, q = FWD ? kl : jl;      // column within the block, contraction index
  return (c >> 4) * 512 + ((c & 15) + 16 * (q >> 3)) * 8 + (q & 7);
}

// Workgroup w of one product (decomposition: column-block group fastest, then y group, k split, angle). MTN > 0:
// the y tiles per group as a compile-time count (equal to p.mt_n), so the tile loops carry no run-time bound and the
// chunk's LDS reads and MFMAs form one straight-line block the scheduler can interleave.
template <int MODE, bool FWD, int WV, int MTN = 0>
__device__ __forceinline__ void blk_pf_body(const BlkArgs& p, int64_t w) {
  constexpr bool RES = MODE == MODE_BF16_RES;
  constexpr bool F16 = MODE == MODE_F16;
  constexpr int TH = 64 * WV;                                // threads of the workgroup
  // y-side operand, double buffered: fragment s = mt * 64 + lane (8 consecutive contraction indices of a row)
  __shared__ s16x8 a_hi[2 * MT_MAX * 64];
  __shared__ s16x8 a_lo[RES ? 2 * MT_MAX * 64 : 1];
  __shared__ __align__(16) float a_inv[F16 ? 2 * MT_MAX * 16 : 4];   // fp16 mode: 1 / scale of each chunk row
  // each wave's expanded PF block: [hi / lo][tile][lane] fragments
  __shared__ s16x8 b_frag[WV][RES ? 2 : 1][2][64];

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int tg = (int)(w % p.n_tg);
  w /= p.n_tg;
  const int yg = (int)(w % p.n_yg);
  w /= p.n_yg;
  const int ks = (int)(w % p.ksplit);
  w /= p.ksplit;
  if (w >= p.n_ang) return;                                  // uniform over the workgroup
  const int a = (int)w;
  const int r = p.omegas ? p.omegas[a] : a;
  const int cb = tg * WV + wv;                               // this wave's column block
  const bool cb_live = cb < p.n_cb;
  const bool cb_stored = FWD ? cb < p.n_jb : cb < p.n_kb;    // adjoint: blocks beyond the batch are empty (zeros)
  const int mt_n = MTN > 0 ? MTN : p.mt_n;
  const int y0 = yg * mt_n * 16;
  const int c0 = ks * p.ch_per;
  const int c1 = min(p.n_ch, c0 + p.ch_per);
  if (c0 >= c1) return;                                      // uniform

  // ---- y-side operand slots of this thread: s = tid + TH i < mt_n * 64
  const int n_slots = mt_n * 64;
  constexpr int NS = (MT_MAX * 64 + TH - 1) / TH;
  const float* srow[NS];
  bool sok[NS];
  const int64_t src_rows = FWD ? (int64_t)a * p.my : (int64_t)r * p.my;
  const int src_stride = FWD ? p.sks : p.cp;
  const int src_n = FWD ? p.kb : p.cp;
  const bool vec = (src_stride & 3) == 0;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const int s = tid + TH * i;
    const int y = y0 + (s >> 6) * 16 + (s & 15);
    sok[i] = s < n_slots && y < p.my;
    srow[i] = p.src + (src_rows + (sok[i] ? y : 0)) * src_stride;
  }
  const int scol = ((tid & 63) >> 4) * 8;                    // the same for every slot of this thread
  // the ring of column j (adjoint fp16 mode): (c * 32 + scol) % P once per chunk, then stepped
  auto load_raw = [&](int c, float (&raw)[NS][8]) {
    const int j0 = c * 32 + scol;
    int jr = 0;
    if (F16 && !FWD) jr = j0 % p.P;
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      if (sok[i] && vec && j0 + 8 <= src_n) {
        const float4 u = *reinterpret_cast<const float4*>(srow[i] + j0);
        const float4 v = *reinterpret_cast<const float4*>(srow[i] + j0 + 4);
        raw[i][0] = u.x; raw[i][1] = u.y; raw[i][2] = u.z; raw[i][3] = u.w;
        raw[i][4] = v.x; raw[i][5] = v.y; raw[i][6] = v.z; raw[i][7] = v.w;
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) raw[i][e] = (sok[i] && j0 + e < src_n) ? srow[i][j0 + e] : 0.f;
      }
      if (F16 && !FWD) {
        // the ring scale of the PF folded into the data: column j = j0 + e, ring j % P
        int q = jr;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          raw[i][e] *= p.ring_inv[q];
          q = q + 1 == p.P ? 0 : q + 1;
        }
      }
    }
  };
  auto store_lds = [&](int buf, float (&raw)[NS][8]) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const int s = tid + TH * i;
      if (F16) {
        // row scale: the 4 slots of a row are lanes l, l ^ 16, l ^ 32, l ^ 48 of this wave (same i)
        float m = 0.f;
#pragma unroll
        for (int e = 0; e < 8; ++e) m = fmaxf(m, fabsf(raw[i][e]));
        m = fmaxf(m, __shfl_xor(m, 16, 64));
        m = fmaxf(m, __shfl_xor(m, 32, 64));
        // 2^(14 - e) with m = f 2^e, f in [0.5, 1): the largest magnitude lands in [2^13, 2^14)
        const int ex = m > 0.f ? ilogbf(m) + 1 : 0;
        const float sc = m > 0.f ? ldexpf(1.f, 14 - ex) : 1.f;
        s16x8 h;
#pragma unroll
        for (int e = 0; e < 8; ++e) h[e] = (short)f16_bits(raw[i][e] * sc);
        if (s < n_slots) {
          a_hi[buf * MT_MAX * 64 + s] = h;
          if ((s & 63) < 16) a_inv[buf * MT_MAX * 16 + (s >> 6) * 16 + (s & 15)] = 1.f / sc;
        }
      } else if (s < n_slots) {
        s16x8 h, l;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const __bf16 hb = (__bf16)raw[i][e];
          h[e] = (short)__builtin_bit_cast(unsigned short, hb);
          if (RES) l[e] = (short)bf16_bits(raw[i][e] - (float)hb);
        }
        a_hi[buf * MT_MAX * 64 + s] = h;
        if (RES) a_lo[buf * MT_MAX * 64 + s] = l;
      }
    }
  };

  // ---- this wave's PF blocks: block (r, kb, jb) with kb = chunk (forward) or the column block (adjoint)
  auto blk_index = [&](int c) -> int64_t {
    const int kbi = FWD ? c : cb, jbi = FWD ? cb : c;
    return ((int64_t)(p.blk_local ? a : r) * p.n_kb + kbi) * p.n_jb + jbi;
  };
  const bool have = cb_live && cb_stored;
  const unsigned* v32 = (const unsigned*)p.val;
  const unsigned short* v16 = (const unsigned short*)p.val;
  // the entries of chunk c's block: its range [lo, hi) and the first ER * 64 (ER per lane), fetched two chunks ahead
  // (a 32 x 32 block holds 64 entries at 6 % fill; blocks near poles hold several hundred)
  constexpr int ER = 4;
  struct Ent {
    int lo, hi;
    unsigned pp[ER], vv[ER];
  };
  // the block range of chunk c (two dependent loads: the pointers are fetched one chunk before the entries, so that
  // the wait for them never includes the loads issued in the same chunk)
  auto fetch_ptr = [&](int c, int2& q) {
    q = make_int2(0, 0);
    if (have && c < c1) {
      const int64_t bi = blk_index(c);
      q.x = p.ptr[bi];
      q.y = p.ptr[bi + 1];
    }
  };
  auto fetch = [&](const int2& q, Ent& E) {
    E.lo = q.x;
    E.hi = q.y;
#pragma unroll
    for (int u = 0; u < ER; ++u) {
      const int e = E.lo + 64 * u + lane;
      if (e < E.hi) {
        E.pp[u] = p.pos[e];
        E.vv[u] = F16 ? (unsigned)v16[e] : v32[e];
      }
    }
  };
  short* bf = reinterpret_cast<short*>(&b_frag[wv][0][0][0]);   // the element type of the s16x8 fragments
  auto scatter = [&](unsigned pp, unsigned vv) {
    const int sl = frag_slot<FWD>(pp);
    bf[sl] = (short)(vv & 0xffffu);
    if (RES) bf[1024 + sl] = (short)(vv >> 16);              // lo part: the second [hi / lo] plane
  };

  f32x4 acc[MT_MAX][2];
#pragma unroll
  for (int mt = 0; mt < MT_MAX; ++mt) acc[mt][0] = acc[mt][1] = f32x4{0.f, 0.f, 0.f, 0.f};
  const s16x8 zero8 = {};

  // chunk c: expand its block into the wave's fragments (in-order LDS accesses of one wave: the clear, the scatter
  // and the read-back need no barrier) and multiply with the y-side operand in LDS buffer buf
  auto compute = [&](int buf, const Ent& E) {
    if (!have) return;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      b_frag[wv][0][t][lane] = zero8;
      if (RES) b_frag[wv][RES ? 1 : 0][t][lane] = zero8;
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < ER; ++u)
      if (E.lo + 64 * u + lane < E.hi) scatter(E.pp[u], E.vv[u]);
    for (int e = E.lo + 64 * ER + lane; e - lane < E.hi; e += 64) {  // blocks of more than 64 ER entries
      if (e < E.hi) scatter(p.pos[e], F16 ? (unsigned)v16[e] : v32[e]);
    }
    __builtin_amdgcn_wave_barrier();
    s16x8 bh[2], bl[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      bh[t] = b_frag[wv][0][t][lane];
      bl[t] = RES ? b_frag[wv][RES ? 1 : 0][t][lane] : zero8;
    }
#pragma unroll
    for (int mt = 0; mt < MT_MAX; ++mt) {
      if (mt < mt_n) {
        const s16x8 ah = a_hi[buf * MT_MAX * 64 + mt * 64 + lane];
        if (F16) {
          const float4 iv = *reinterpret_cast<const float4*>(&a_inv[buf * MT_MAX * 16 + mt * 16 + (lane >> 4) * 4]);
#pragma unroll
          for (int t = 0; t < 2; ++t) {
            const f32x4 d = mma_f16(ah, bh[t], f32x4{0.f, 0.f, 0.f, 0.f});
            acc[mt][t][0] += d[0] * iv.x;
            acc[mt][t][1] += d[1] * iv.y;
            acc[mt][t][2] += d[2] * iv.z;
            acc[mt][t][3] += d[3] * iv.w;
          }
          __builtin_amdgcn_sched_barrier(0);   // one row tile's products at a time (register pressure)
        } else {
          const s16x8 al = RES ? a_lo[buf * MT_MAX * 64 + mt * 64 + lane] : zero8;
#pragma unroll
          for (int t = 0; t < 2; ++t) {
            acc[mt][t] = mma_bf16(ah, bh[t], acc[mt][t]);
            if (RES) {
              acc[mt][t] = mma_bf16(ah, bl[t], acc[mt][t]);
              acc[mt][t] = mma_bf16(al, bh[t], acc[mt][t]);
            }
          }
        }
      }
    }
  };

  // Software pipeline, two chunks deep: at chunk c the y-side rows of chunk c + 2 and the entries of its block are
  // fetched (its block range three chunks ahead), chunk c is computed, and the rows of chunk c + 1 (fetched at chunk
  // c - 1) go to the other LDS buffer.
  // The loop is unrolled by two so that the register sets alternate roles instead of being copied.
  float rawA[NS][8], rawB[NS][8];
  Ent E0, E1;
  int2 Q0, Q1;                                               // block ranges of chunks c + 2 and c + 3
  load_raw(c0, rawA);
  store_lds(0, rawA);
  if (c0 + 1 < c1) load_raw(c0 + 1, rawA);
  fetch_ptr(c0, Q0);
  fetch_ptr(c0 + 1, Q1);
  fetch(Q0, E0);
  fetch(Q1, E1);
  fetch_ptr(c0 + 2, Q0);
  __syncthreads();
  // chunk c: the ranges of chunk c + 3 into qn, the rows of chunk c + 2, compute chunk c, the entries of chunk
  // c + 2 (range qe, fetched one chunk ago), the rows of chunk c + 1 to the other LDS buffer
  auto step = [&](int c, int buf, float (&rin)[NS][8], float (&rout)[NS][8], Ent& E, const int2& qe, int2& qn) {
    fetch_ptr(c + 3, qn);
    if (c + 2 < c1) load_raw(c + 2, rout);
    compute(buf, E);
    fetch(qe, E);
    if (c + 1 < c1) store_lds(buf ^ 1, rin);
    __syncthreads();
  };
  for (int c = c0; c < c1; c += 2) {
    step(c, 0, rawA, rawB, E0, Q0, Q1);
    if (c + 1 < c1) step(c + 1, 1, rawB, rawA, E1, Q1, Q0);
  }

  if (!cb_live) return;
  // D fragment: lane l holds rows 4 (l / 16) + i, column l % 16 of each 16 x 16 tile
  const int row = lane & 15, grp = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < MT_MAX; ++mt) {
    if (mt >= mt_n) continue;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int col = cb * 32 + t * 16 + row;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int y = y0 + mt * 16 + grp * 4 + i;
        if (y >= p.my) continue;
        float v = acc[mt][t][i];
        if (FWD) {
          if (col >= p.cp) continue;
          if (F16) v *= p.ring_inv[col % p.P];
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

// WV waves per workgroup (WV column blocks share each staged y-side chunk): 4 (three workgroups per CU, two in the
// fp16 mode) or 8 (half the re-reads of the y-side rows, which every group of WV column blocks stages again; one
// workgroup per CU).
template <int MODE, bool FWD, int WV, int MTN>
__global__ __launch_bounds__(64 * WV, WV == 4 ? (MODE == MODE_F16 ? 2 : 3) : 2) void blk_pf_kernel(
    BlkArgs p) {
  blk_pf_body<MODE, FWD, WV, MTN>(p, xcd_block());
}

// the y-tile counts compiled in (MTN) for the 4-wave form: 7 (97-112 detector rows in one group) and 6 (groups of
// 6 tiles, e.g. 261 rows in three)

// Grouped launch: workgroup w runs workgroup wg_map[w].y of product wg_map[w].x (the slices of a multi-slice batch,
// each its own PF blocks, K, sinogram columns and data rows); entries with x < 0 pad the grid.
// (MTN > 0: every descriptor's y-tile count, compiled in)
template <int MODE, bool FWD, int MTN>
__global__ __launch_bounds__(THREADS, MODE == MODE_F16 ? 2 : 3) void blk_pf_grouped_kernel(const BlkArgs* __restrict__ descs,
                                                                    const int2* __restrict__ wg_map, int64_t n_map) {
  const int64_t w = xcd_block();
  if (w >= n_map) return;
  const int2 m = wg_map[w];
  if (m.x < 0) return;
  const BlkArgs p = descs[m.x];
  blk_pf_body<MODE, FWD, WAVES, MTN>(p, m.y);
}

template <bool FWD, int MTN>
void launch_blk_grouped_mt(const BlkArgs* descs, const int2* wg_map, int64_t n_map, int mode, hipStream_t st) {
  const dim3 grid((unsigned)((n_map + N_XCD - 1) / N_XCD * N_XCD)), block(THREADS);
  if (mode == MODE_BF16_RES)
    hipLaunchKernelGGL((blk_pf_grouped_kernel<MODE_BF16_RES, FWD, MTN>), grid, block, 0, st, descs, wg_map, n_map);
  else if (mode == MODE_BF16)
    hipLaunchKernelGGL((blk_pf_grouped_kernel<MODE_BF16, FWD, MTN>), grid, block, 0, st, descs, wg_map, n_map);
  else
    hipLaunchKernelGGL((blk_pf_grouped_kernel<MODE_F16, FWD, MTN>), grid, block, 0, st, descs, wg_map, n_map);
}

template <bool FWD>
void launch_blk_grouped(const BlkArgs* descs, const int2* wg_map, int64_t n_map, int mode, int mtn, hipStream_t st) {
  if (mtn == 5) launch_blk_grouped_mt<FWD, 5>(descs, wg_map, n_map, mode, st);
  else launch_blk_grouped_mt<FWD, 0>(descs, wg_map, n_map, mode, st);
}

template <bool FWD, int WV, int MTN>
void launch_blk_wv(const BlkArgs& p, int mode, hipStream_t st) {
  const int64_t n_wg = (int64_t)p.n_ang * p.ksplit * p.n_yg * p.n_tg;
  const dim3 grid((unsigned)((n_wg + N_XCD - 1) / N_XCD * N_XCD)), block(64 * WV);
  if (mode == MODE_BF16_RES) hipLaunchKernelGGL((blk_pf_kernel<MODE_BF16_RES, FWD, WV, MTN>), grid, block, 0, st, p);
  else if (mode == MODE_BF16) hipLaunchKernelGGL((blk_pf_kernel<MODE_BF16, FWD, WV, MTN>), grid, block, 0, st, p);
  else hipLaunchKernelGGL((blk_pf_kernel<MODE_F16, FWD, WV, MTN>), grid, block, 0, st, p);
}

template <bool FWD>
void launch_blk(const BlkArgs& p, int mode, int wv, bool fixed, hipStream_t st) {
  if (wv == 8) launch_blk_wv<FWD, 8, 0>(p, mode, st);
  else if (fixed && p.mt_n == 7) launch_blk_wv<FWD, 4, 7>(p, mode, st);
  else if (fixed && p.mt_n == 6) launch_blk_wv<FWD, 4, 6>(p, mode, st);
  else launch_blk_wv<FWD, 4, 0>(p, mode, st);
}

// waves per workgroup from bits 5-6 of the mode argument: 0 -> 4, 1 -> 8; bit 7: the 4-wave form with the y-tile
// count compiled in where it is 6 or 7 (the same results)
int mode_waves(int mode) { return 4 << ((mode >> 5) & 3); }

void y_tiling(int my, int& mt_n, int& n_yg) {
  const int tiles = (my + 15) / 16;
  n_yg = (tiles + MT_MAX - 1) / MT_MAX;
  mt_n = (tiles + n_yg - 1) / n_yg;
}

}  // namespace

extern "C" {

int dfm_pfblk_count(const float* dense, int32_t* counts, int32_t R, int32_t km, int32_t kb, int32_t cp,
                    dfm_stream_t stream) {
  DFM_CHECK_ARG(dense && counts, "dfm_pfblk_count: null pointer");
  DFM_CHECK_ARG(R > 0 && kb > 0 && km >= kb && cp > 0, "dfm_pfblk_count: bad sizes");
  const int n_kb = (kb + 31) / 32, n_jb = (cp + 31) / 32;
  const int64_t threads = (int64_t)R * n_kb * n_jb * 64;
  hipLaunchKernelGGL(blk_count_kernel, dim3(dfm::grid_stride_blocks(threads, 256, 1u << 20)), dim3(256), 0,
                     (hipStream_t)stream, dense, counts, R, km, kb, cp, n_kb, n_jb);
  DFM_LAUNCH_CHECK("dfm_pfblk_count");
  return DFM_OK;
}

int dfm_pfblk_fill(const float* dense, const int32_t* ptr, uint16_t* pos, void* val, int32_t fp16,
                   const float* ring_scale, int32_t P, int32_t R, int32_t km, int32_t kb, int32_t cp,
                   dfm_stream_t stream) {
  DFM_CHECK_ARG(dense && ptr && pos && val && (!fp16 || (ring_scale && P > 0)), "dfm_pfblk_fill: null pointer");
  DFM_CHECK_ARG(R > 0 && kb > 0 && km >= kb && cp > 0, "dfm_pfblk_fill: bad sizes");
  const int n_kb = (kb + 31) / 32, n_jb = (cp + 31) / 32;
  const int64_t threads = (int64_t)R * n_kb * n_jb * 64;
  hipLaunchKernelGGL(blk_fill_kernel, dim3(dfm::grid_stride_blocks(threads, 256, 1u << 20)), dim3(256), 0,
                     (hipStream_t)stream, dense, ptr, pos, fp16 ? nullptr : (unsigned*)val,
                     fp16 ? (unsigned short*)val : nullptr, ring_scale, P, R, km, kb, cp, n_kb, n_jb);
  DFM_LAUNCH_CHECK("dfm_pfblk_fill");
  return DFM_OK;
}

int dfm_pfblk_forward(const float* sino, int32_t sks, int32_t kb, const int32_t* ptr, const uint16_t* pos,
                      const void* val, int32_t mode, const float* ring_inv, int32_t P, const int32_t* omegas,
                      int32_t n_ang, int32_t my, int32_t cp, float* data, float* part, int32_t ksplit,
                      int32_t accumulate, dfm_stream_t stream) {
  DFM_CHECK_ARG(sino && ptr && pos && val && data, "dfm_pfblk_forward: null pointer");
  const int local = (mode >> 4) & 1;
  const int wv = mode_waves(mode);
  DFM_CHECK_ARG(wv <= 8, "dfm_pfblk_forward: waves code %d (0, 1: 4, 8 waves per workgroup)", (mode >> 5) & 3);
  const bool fixed = (mode >> 7) & 1;
  mode &= 15;
  DFM_CHECK_ARG(mode >= MODE_BF16 && mode <= MODE_F16 && (mode != MODE_F16 || (ring_inv && P > 0)),
                "dfm_pfblk_forward: mode %d (0 bf16, 1 bf16 residual, 2 fp16 with ring scales)", mode);
  DFM_CHECK_ARG(n_ang > 0 && my > 0 && cp > 0 && kb > 0 && sks >= kb && (sks & 3) == 0,
                "dfm_pfblk_forward: bad sizes (n_ang %d, My %d, CP %d, Kb %d, stride %d)", n_ang, my, cp, kb, sks);
  BlkArgs p{};
  p.blk_local = local;
  p.n_kb = (kb + 31) / 32;
  p.n_jb = (cp + 31) / 32;
  DFM_CHECK_ARG(ksplit >= 1 && ksplit <= p.n_kb, "dfm_pfblk_forward: k split %d not in 1 .. %d", ksplit, p.n_kb);
  DFM_CHECK_ARG(ksplit == 1 || part, "dfm_pfblk_forward: a k split needs the partial-sum workspace");
  p.src = sino;
  p.ptr = ptr;
  p.pos = pos;
  p.val = val;
  p.omegas = omegas;
  p.ring_inv = ring_inv;
  p.P = P;
  p.n_ang = n_ang;
  p.my = my;
  p.cp = cp;
  p.sks = sks;
  p.kb = kb;
  p.n_cb = p.n_jb;
  p.n_ch = p.n_kb;
  y_tiling(my, p.mt_n, p.n_yg);
  p.n_tg = (p.n_cb + wv - 1) / wv;
  p.ch_per = (p.n_ch + ksplit - 1) / ksplit;
  p.ksplit = (p.n_ch + p.ch_per - 1) / p.ch_per;
  p.direct = p.ksplit == 1;
  p.accumulate = accumulate;
  p.out = p.direct ? data : part;
  p.alpha = 1.f;
  const hipStream_t st = (hipStream_t)stream;
  launch_blk<true>(p, mode, wv, fixed, st);
  DFM_LAUNCH_CHECK("dfm_pfblk_forward");
  if (!p.direct) return dfm::launch_splitk_reduce(part, data, omegas, n_ang, (int64_t)my * cp, p.ksplit, accumulate, st);
  return DFM_OK;
}

int dfm_pfblk_adjoint(const float* data, const int32_t* ptr, const uint16_t* pos, const void* val, int32_t mode,
                      const float* ring_inv, int32_t P, const int32_t* omegas, int32_t n_ang, int32_t my, int32_t cp,
                      int32_t kb, float* sino, int32_t sks, float alpha, dfm_stream_t stream) {
  DFM_CHECK_ARG(data && ptr && pos && val && sino, "dfm_pfblk_adjoint: null pointer");
  const int local = (mode >> 4) & 1;
  const int wv = mode_waves(mode);
  DFM_CHECK_ARG(wv <= 8, "dfm_pfblk_adjoint: waves code %d (0, 1: 4, 8 waves per workgroup)", (mode >> 5) & 3);
  const bool fixed = (mode >> 7) & 1;
  mode &= 15;
  DFM_CHECK_ARG(mode >= MODE_BF16 && mode <= MODE_F16 && (mode != MODE_F16 || (ring_inv && P > 0)),
                "dfm_pfblk_adjoint: mode %d (0 bf16, 1 bf16 residual, 2 fp16 with ring scales)", mode);
  DFM_CHECK_ARG(n_ang > 0 && my > 0 && cp > 0 && kb > 0 && sks >= kb && (sks & 3) == 0,
                "dfm_pfblk_adjoint: bad sizes (n_ang %d, My %d, CP %d, Kb %d, stride %d)", n_ang, my, cp, kb, sks);
  BlkArgs p{};
  p.blk_local = local;
  p.n_kb = (kb + 31) / 32;
  p.n_jb = (cp + 31) / 32;
  p.src = data;
  p.ptr = ptr;
  p.pos = pos;
  p.val = val;
  p.omegas = omegas;
  p.ring_inv = ring_inv;
  p.P = P;
  p.out = sino;
  p.n_ang = n_ang;
  p.my = my;
  p.cp = cp;
  p.sks = sks;
  p.kb = sks;
  p.n_cb = (sks + 31) / 32;           // every sinogram column is written (those of no stored block with 0)
  p.n_ch = p.n_jb;
  y_tiling(my, p.mt_n, p.n_yg);
  p.n_tg = (p.n_cb + wv - 1) / wv;
  p.ksplit = 1;
  p.ch_per = p.n_ch;
  p.alpha = alpha;
  p.ncols = sks;
  launch_blk<false>(p, mode, wv, fixed, (hipStream_t)stream);
  DFM_LAUNCH_CHECK("dfm_pfblk_adjoint");
  return DFM_OK;
}

int dfm_pfblk_grouped(int32_t forward, const void* descs, const int32_t* wg_map, int64_t n_wg, int32_t mode,
                      dfm_stream_t stream) {
  DFM_CHECK_ARG(descs && wg_map && n_wg > 0, "dfm_pfblk_grouped: null pointer or no workgroups");
  // bits 8-11: every descriptor's y-tile count mt_n, when the caller knows it (5 has a compiled form; 0: unknown)
  const int mtn = (mode >> 8) & 15;
  mode &= 15;
  DFM_CHECK_ARG(mode >= MODE_BF16 && mode <= MODE_F16, "dfm_pfblk_grouped: mode %d", mode);
  DFM_CHECK_ARG(n_wg <= ((int64_t)1 << 31) - 8, "dfm_pfblk_grouped: %lld workgroups", (long long)n_wg);
  if (forward)
    launch_blk_grouped<true>((const BlkArgs*)descs, (const int2*)wg_map, n_wg, mode, mtn, (hipStream_t)stream);
  else
    launch_blk_grouped<false>((const BlkArgs*)descs, (const int2*)wg_map, n_wg, mode, mtn, (hipStream_t)stream);
  DFM_LAUNCH_CHECK("dfm_pfblk_grouped");
  return DFM_OK;
}

int dfm_pfblk_desc_size(void) { return (int)sizeof(BlkArgs); }

}  // extern "C"
