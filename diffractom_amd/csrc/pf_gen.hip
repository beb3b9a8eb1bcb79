// The pole-figure matrix generated on the fly (the generated PF mode, pf_generate.py:218-303 with the kernels of
// pf_generate.cl:146-454), straight into the block-sparse form the matrix-core products read (pf_blocks.hip): for the
// angles of one call only, no stored PF.
//
// A workgroup owns one angle r and 32 orientations (an orientation block). It first marks the candidate entries in
// an LDS bitmap (32 orientations x C*P segments): for every orientation k, pole axis a and omega node s the pole
// direction n = U_k h_a is taken into the detector frame of the node (x, y, z = n . R^T d0, n . R^T d90, n . R^T p0),
// where the probed directions of ring p form the cone v(eta) = cos(th_p) (cos(eta) R^T d0 + sin(eta) R^T d90) -
// sin(th_p) R^T p0, so that n . v = A cos(eta - phi) - B with A = cos(th_p) sqrt(x^2 + y^2), B = sin(th_p) z,
// phi = atan2(y, x). |n . v| > tau = 1 - cut sigma^2 - 1e-4 (a superset of the entries whose exponent
// (1 - |n . v|) / sigma^2 stays below the cut) holds on eta windows phi +- acos((tau + B) / A) (for n) and
// phi + pi +- acos((tau - B) / A) (for -n); every eta bin with a quadrature node in a window is a candidate of ring
// p. Then each wave takes 32 x 32 blocks (orientation block x 32 segments): a lane evaluates its row's candidates
// of the block with the stored mode's own entry function (pf_entry: all axes of the ring, all nodes, float64), so
// the values equal the stored PF bit for bit, and the non-zeros are written in element order (orientation-major,
// then segment) at the block's position in the pointer array — or, in count mode, counted. The counts are taken
// once per angle set; the pointers of an angle list are their exclusive scan.
#include "dfm_common.h"

namespace {

constexpr int G_THREADS = 256;
constexpr int G_WAVES = G_THREADS / 64;

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
  extern __shared__ unsigned gen_lds[];   // bitmap: 32 rows x n_jb words; values: G_WAVES x 32 x 32 floats
  unsigned* bm = gen_lds;
  float* vtmp = reinterpret_cast<float*>(gen_lds + 32 * g.n_jb);
  const int kbi = blockIdx.x % g.n_kb;
  const int a = blockIdx.x / g.n_kb;
  if (a >= g.n_ang) return;
  const int r = g.omegas[a];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int CP = g.C * g.P;
  for (int i = tid; i < 32 * g.n_jb; i += G_THREADS) bm[i] = 0u;
  __syncthreads();

  // ---- candidates
  const int items = 32 * g.A * g.S;
  for (int i = tid; i < items; i += G_THREADS) {
    const int kl = i & 31;
    const int rest = i >> 5;
    const int ax = rest % g.A, s = rest / g.A;
    const int kloc = kbi * 32 + kl;
    if (kloc >= g.kb) continue;
    const int kg = g.k0 + kloc;
    int p = 0;
    while (p + 1 < g.P && g.axis_start[p + 1] <= ax) ++p;
    const double tau = 1.0 - g.cut / g.inv_sigma2[kg] - 1e-4;
    const double* n = g.poles + ((int64_t)kg * g.A + ax) * 3;
    const double* f = g.frames + ((int64_t)r * g.S + s) * 9;
    const double x = n[0] * f[0] + n[1] * f[1] + n[2] * f[2];
    const double y = n[0] * f[3] + n[1] * f[4] + n[2] * f[5];
    const double z = n[0] * f[6] + n[1] * f[7] + n[2] * f[8];
    const double Aa = g.cos_th[p] * sqrt(x * x + y * y), B = g.sin_th[p] * z;
    if (tau - fabs(B) >= Aa) continue;                       // neither n nor -n reaches the cone's cap
    const double phi = atan2(y, x);
    for (int sg = 0; sg < 2; ++sg) {
      const double ratio = (tau + (sg ? -B : B)) / fmax(Aa, 1e-300);
      if (ratio >= 1.0) continue;
      const double w = ratio <= -1.0 ? 3.2 : acos(ratio) + 1e-9;
      const double center = phi + (sg ? 3.141592653589793 : 0.0);
      for (int sh = -1; sh <= 1; ++sh) {                    // eta is an angle: the window and its turns
        const double cc = center + sh * 6.283185307179586;
        // bin c has its nodes at eta0 + (c + x) deta, xlo <= x <= xhi
        int m_lo = (int)ceil((cc - w - g.eta0) / g.deta - g.xhi);
        int m_hi = (int)floor((cc + w - g.eta0) / g.deta - g.xlo);
        m_lo = max(m_lo, 0);
        m_hi = min(m_hi, g.C - 1);
        for (int c = m_lo; c <= m_hi; ++c) {
          const int j = c * g.P + p;
          atomicOr(&bm[kl * g.n_jb + (j >> 5)], 1u << (j & 31));
        }
      }
    }
  }
  __syncthreads();

  // ---- evaluate: a wave per block, lane = (row kl = lane % 32, half h = lane / 32: the row's odd / even bits)
  const int kl = lane & 31, h = lane >> 5;
  const int kloc = kbi * 32 + kl;
  const bool krow = kloc < g.kb;
  const int kg = g.k0 + kloc;
  float* vt = vtmp + (wv * 32 + kl) * 32;
  for (int jb = wv; jb < g.n_jb; jb += G_WAVES) {
    const unsigned word = krow ? bm[kl * g.n_jb + jb] : 0u;
    unsigned mine = word & (h ? 0xAAAAAAAAu : 0x55555555u);
    unsigned nz = 0u;
    while (mine) {
      const int b = __builtin_ctz(mine);
      mine &= mine - 1u;
      const int j = jb * 32 + b;
      float v = 0.f;
      if (j < CP) {
        v = dfm::pf_entry(g.coords, g.poles, g.axis_start, g.axis_count, g.inv_sigma2, g.normf, g.qw, r, kg, j / g.P,
                     j % g.P, g.C, g.P, g.A, g.S, g.Se, g.cut, g.quad_weighted);
      }
      if (v != 0.f) {
        nz |= 1u << b;
        vt[b] = v;
      }
    }
    const unsigned rownz = nz | __shfl_xor(nz, 32, 64);       // the row's non-zeros (both halves)
    const int cnt = __popc(rownz);
    // exclusive prefix of the row counts over rows 0..31 (the same in both halves)
    int incl = cnt;
    for (int off = 1; off < 32; off <<= 1) {
      const int t = __shfl_up(incl, off, 32);
      if (kl >= off) incl += t;
    }
    const int total = __shfl(incl, 31, 64);
    const int64_t blk = ((int64_t)a * g.n_kb + kbi) * g.n_jb + jb;
    if (g.mode == 0) {
      if (lane == 0) g.counts[blk] = total;
      continue;
    }
    const int base = g.ptr[blk] + incl - cnt;
    unsigned m = nz;
    while (m) {
      const int b = __builtin_ctz(m);
      m &= m - 1u;
      const int o = base + __popc(rownz & ((1u << b) - 1u));
      const float v = vt[b];
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
  const size_t lds = (size_t)32 * g.n_jb * sizeof(unsigned) + (size_t)G_WAVES * 32 * 32 * sizeof(float);
  DFM_CHECK_ARG(lds <= 163840, "dfm_pfgen: %d segment blocks exceed the LDS", g.n_jb);
  const int64_t n_wg = (int64_t)n_ang * g.n_kb;
  DFM_CHECK_ARG(n_wg < ((int64_t)1 << 31), "dfm_pfgen: too many workgroups");
  hipLaunchKernelGGL(pf_gen_kernel, dim3((unsigned)n_wg), dim3(G_THREADS), lds, (hipStream_t)stream, g);
  DFM_LAUNCH_CHECK("dfm_pfgen");
  return DFM_OK;
}

}  // extern "C"
