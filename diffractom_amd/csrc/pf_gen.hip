// The pole-figure matrix generated on the fly (the generated PF mode, pf_generate.py:218-303 with the kernels of
// pf_generate.cl:146-454), straight into the block-sparse form the matrix-core products read (pf_blocks.hip): for the
// angles of one call only, no stored PF.
//
// A workgroup owns one angle r and 32 orientations (an orientation block). It first marks the candidate entries in
// an LDS bitmap (one word per 32 segments and orientation, words in (segment block, orientation) order) from the eta
// windows of pf_cand.h. The bitmap's set bits, read in word then bit order, are the block-sparse form's element
// order (segment block, then orientation, then segment). They are expanded into an LDS list (at most LIST_CAP at a
// time) and evaluated one candidate per thread with the stored mode's own entry function (pf_entry: all axes of the
// ring, all nodes, float64), so the values equal the stored PF bit for bit; a workgroup-wide scan of the non-zero
// flags gives each non-zero its place. The blocks of one (angle, orientation block) are consecutive in the pointer
// array, so the workgroup's non-zeros form one run from the first block's pointer and the stores are coalesced.
// Count mode adds the non-zeros per block instead. The counts are taken once per angle set; the pointers of an
// angle list are their exclusive scan.
#include "pf_cand.h"

namespace {

constexpr int G_THREADS = 256;
constexpr int G_WAVES = G_THREADS / 64;
constexpr int LIST_CAP = 4096;            // candidates expanded per pass (16 KB of LDS)

struct GenArgs {
  const double* coords;      // probed directions (R, S, C, Se, P, 3)
  const double* poles;       // (K, A, 3)
  const int* axis_start;     // (P + 1,)
  const float* axis_count;   // (A,)
  const double* inv_sigma2;  // (K,)
  const double* normf;       // (K,)
  const float* qw;           // quadrature weights (S * Se) or unused
  const double* frames;      // (R, S, 9): R^T d0, R^T d90, R^T p0 of every omega node
  const double* sin_th;      // (P,)
  const double* cos_th;      // (P,)
  const int* omegas;         // the angles of the call (n_ang)
  const int* ptr;            // fill: block pointers of the call, (n_ang * n_kb * n_jb + 1)
  int* counts;               // count: non-zeros per block, indexed (a, kb, jb)
  unsigned short* pos;
  unsigned* val32;           // bf16 pairs
  unsigned short* val16;     // fp16 times the ring scale
  const float* ring_scale;
  int n_ang, k0, kb, n_kb, n_jb, C, P, A, S, Se, quad_weighted;
  double cut, eta0, deta, xlo, xhi;
  int mode;                  // 0 count, 1 fill
};

__device__ __forceinline__ unsigned short f16_bits(float v) { return __builtin_bit_cast(unsigned short, (_Float16)v); }

__global__ void __launch_bounds__(G_THREADS) pf_gen_kernel(GenArgs g) {
  extern __shared__ unsigned gen_lds[];
  unsigned* bm = gen_lds;                                   // word (jb, kl) at jb * 32 + kl
  unsigned* list = bm + 32 * g.n_jb;                        // LIST_CAP entries: word index << 5 | bit
  int* wsum = reinterpret_cast<int*>(list + LIST_CAP);      // G_WAVES wave totals
  int* cnt = wsum + G_WAVES;                                // count mode: non-zeros per segment block
  const int kbi = blockIdx.x % g.n_kb;
  const int a = blockIdx.x / g.n_kb;
  if (a >= g.n_ang) return;
  const int r = g.omegas[a];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int CP = g.C * g.P;
  const int W = 32 * g.n_jb;
  for (int i = tid; i < W; i += G_THREADS) bm[i] = 0u;
  if (g.mode == 0)
    for (int i = tid; i < g.n_jb; i += G_THREADS) cnt[i] = 0;
  __syncthreads();

  // ---- candidates
  const dfm::CandGeo cg{g.poles, g.frames, g.sin_th, g.cos_th, g.axis_start, g.inv_sigma2, g.k0, g.kb, g.C, g.P,
                        g.A, g.S, g.cut, g.eta0, g.deta, g.xlo, g.xhi};
  dfm::mark_candidates(cg, r, kbi, bm, tid, G_THREADS);
  __syncthreads();

  // ---- the candidates' order: each thread owns a run of words; exclusive scan of their bit counts
  const int per = (W + G_THREADS - 1) / G_THREADS;
  const int w0 = min(tid * per, W), w1 = min(w0 + per, W);
  int mine = 0;
  for (int w = w0; w < w1; ++w) mine += __popc(bm[w]);
  int incl = mine;
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) wsum[wv] = incl;
  __syncthreads();
  int before_w = 0, ncand = 0;
  for (int q = 0; q < G_WAVES; ++q) {
    const int t = wsum[q];
    if (q < wv) before_w += t;
    ncand += t;
  }
  const int cbase = before_w + incl - mine;                  // this thread's first candidate
  const int64_t blk0 = ((int64_t)a * g.n_kb + kbi) * g.n_jb;
  int64_t out = g.mode == 1 ? (int64_t)g.ptr[blk0] : 0;     // the workgroup's run of non-zeros
  __syncthreads();                                           // wsum is reused below

  for (int c0 = 0; c0 < ncand; c0 += LIST_CAP) {
    const int c1 = c0 + LIST_CAP;
    if (mine > 0 && cbase < c1 && cbase + mine > c0) {
      int idx = cbase;
      for (int w = w0; w < w1 && idx < c1; ++w) {
        unsigned bits = bm[w];
        const int nb = __popc(bits);
        if (idx + nb <= c0) {
          idx += nb;
          continue;
        }
        while (bits && idx < c1) {
          const int b = __builtin_ctz(bits);
          bits &= bits - 1u;
          if (idx >= c0) list[idx - c0] = ((unsigned)w << 5) | (unsigned)b;
          ++idx;
        }
      }
    }
    __syncthreads();
    const int n = min(LIST_CAP, ncand - c0);
    for (int i0 = 0; i0 < n; i0 += G_THREADS) {
      const int i = i0 + tid;
      float v = 0.f;
      int jb = 0, kl = 0, b = 0;
      if (i < n) {
        const unsigned e = list[i];
        const int w = (int)(e >> 5);
        b = (int)(e & 31u);
        jb = w >> 5;
        kl = w & 31;
        const int j = jb * 32 + b;
        const int kloc = kbi * 32 + kl;
        if (j < CP && kloc < g.kb)
          v = dfm::pf_entry(g.coords, g.poles, g.axis_start, g.axis_count, g.inv_sigma2, g.normf, g.qw, r,
                            g.k0 + kloc, j / g.P, j % g.P, g.C, g.P, g.A, g.S, g.Se, g.cut, g.quad_weighted);
      }
      const bool nzf = v != 0.f;
      const unsigned long long bal = __ballot(nzf);
      const int rank = __popcll(bal & ((1ull << lane) - 1ull));
      if (lane == 0) wsum[wv] = __popcll(bal);
      __syncthreads();
      int before = 0, tot = 0;
      for (int q = 0; q < G_WAVES; ++q) {
        const int t = wsum[q];
        if (q < wv) before += t;
        tot += t;
      }
      if (nzf) {
        if (g.mode == 0) {
          atomicAdd(&cnt[jb], 1);
        } else {
          const int64_t o = out + before + rank;
          g.pos[o] = (unsigned short)(kl | (b << 5));
          if (g.val16) {
            g.val16[o] = f16_bits(v * g.ring_scale[(jb * 32 + b) % g.P]);
          } else {
            const __bf16 hi = (__bf16)v;
            const __bf16 lo = (__bf16)(v - (float)hi);
            g.val32[o] = (unsigned)__builtin_bit_cast(unsigned short, hi) |
                         ((unsigned)__builtin_bit_cast(unsigned short, lo) << 16);
          }
        }
      }
      out += tot;
      __syncthreads();                                       // wsum (and, after the last, list) reused
    }
  }
  if (g.mode == 0)
    for (int i = tid; i < g.n_jb; i += G_THREADS) g.counts[blk0 + i] = cnt[i];
}

}  // namespace

extern "C" {

int dfm_pfgen(int32_t mode, const double* coords, const double* poles, const int32_t* axis_start,
              const float* axis_count, const double* inv_sigma2, const double* norm, const float* qw,
              int32_t quad_weighted, const double* frames, const double* sin_th, const double* cos_th,
              const int32_t* omegas, int32_t n_ang, int32_t k0, int32_t kb, int32_t C, int32_t P, int32_t A, int32_t S,
              int32_t Se, float pf_cut, double eta0, double deta, double xlo, double xhi,
              int32_t* counts, const int32_t* ptr, uint16_t* pos, void* val, int32_t fp16, const float* ring_scale,
              dfm_stream_t stream) {
  DFM_CHECK_ARG(coords && poles && axis_start && axis_count && inv_sigma2 && norm && frames && sin_th && cos_th &&
                    omegas && (!quad_weighted || qw),
                "dfm_pfgen: null pointer");
  DFM_CHECK_ARG(mode == 0 ? counts != nullptr : (ptr && pos && val && (!fp16 || ring_scale)),
                "dfm_pfgen: mode %d needs %s", mode, mode == 0 ? "counts" : "ptr, pos, val (and ring scales)");
  DFM_CHECK_ARG(n_ang >= 0 && kb > 0 && k0 >= 0 && C > 0 && P > 0 && A >= P && S > 0 && Se > 0 && deta > 0.0 &&
                    xlo <= xhi && C * P <= 65536,
                "dfm_pfgen: bad sizes");
  if (n_ang == 0) return DFM_OK;
  GenArgs g{};
  g.coords = coords;
  g.poles = poles;
  g.axis_start = axis_start;
  g.axis_count = axis_count;
  g.inv_sigma2 = inv_sigma2;
  g.normf = norm;
  g.qw = qw;
  g.frames = frames;
  g.sin_th = sin_th;
  g.cos_th = cos_th;
  g.omegas = omegas;
  g.ptr = ptr;
  g.counts = counts;
  g.pos = pos;
  g.val32 = fp16 ? nullptr : (unsigned*)val;
  g.val16 = fp16 ? (unsigned short*)val : nullptr;
  g.ring_scale = ring_scale;
  g.n_ang = n_ang;
  g.k0 = k0;
  g.kb = kb;
  g.n_kb = (kb + 31) / 32;
  g.n_jb = (C * P + 31) / 32;
  g.C = C;
  g.P = P;
  g.A = A;
  g.S = S;
  g.Se = Se;
  g.quad_weighted = quad_weighted;
  g.cut = (double)pf_cut;
  g.eta0 = eta0;
  g.deta = deta;
  g.xlo = xlo;
  g.xhi = xhi;
  g.mode = mode;
  const size_t lds = ((size_t)32 * g.n_jb + LIST_CAP + G_WAVES + g.n_jb) * sizeof(unsigned);
  DFM_CHECK_ARG(lds <= 163840, "dfm_pfgen: %d segment blocks exceed the LDS", g.n_jb);
  const int64_t n_wg = (int64_t)n_ang * g.n_kb;
  DFM_CHECK_ARG(n_wg < ((int64_t)1 << 31), "dfm_pfgen: too many workgroups");
  hipLaunchKernelGGL(pf_gen_kernel, dim3((unsigned)n_wg), dim3(G_THREADS), lds, (hipStream_t)stream, g);
  DFM_LAUNCH_CHECK("dfm_pfgen");
  return DFM_OK;
}

}  // extern "C"
