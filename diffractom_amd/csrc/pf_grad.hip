// Gradients of <r, A c> with respect to the PF parameters (differentiable/gradients.py:20-189 PFGradient, kernels
// pf_grad.cl:42-187): the rotation (3) and kernel width of every orientation, the eta offset, the rotation-axis tilt
// (2) and the Bragg angle of every ring; optionally the Gauss-Newton block J_k^T diag(curv) J_k (4 x 4) of every
// orientation's rotation and width.
//
// With S = Radon(c) (the caller's sinogram of the angle list, orientation-fastest) and G[a, k, j] = sum_y r[a, y, j]
// S[a, y, k] (dL / dPF), the gradient is the sum over the PF's non-zeros of G times the entry's derivative:
//   PF[a, k, j] = N_k sum_nodes w sum_{axes h of ring p} m_h g(n_kh . v),  g(d) = e1 + e2,
//   e1 = exp(-(1 - |d|) / sigma_k^2), e2 = exp(-(1 + |d|) / sigma_k^2) (each dropped at the cut),
//   d g / d d = sign(d) (e1 - e2) / sigma^2, d g / d(1 / sigma^2) = -(1 - |d|) e1 - (1 + |d|) e2,
//   d(n . v) / d rot = n x v (a left increment of U_k), d(n . v) / d eta_offset = n . dv_eta, and the same with the
//   probed directions' derivatives for the tilt and the Bragg angles (host-computed, float64, shaped as v).
//
// A workgroup owns one angle and 32 orientations, as the PF generation: the candidate segments are marked from the
// eta windows (pf_cand.h), the orientations' sinogram columns S[a, :, k] staged once in LDS, and for every segment
// block with candidates the data rows r[a, :, j] (and curv) of its 32 segments; one thread per candidate forms G by
// an My-long dot (float64) and the entry's derivative terms (float64, as pf_entry), accumulated per orientation in
// LDS (float64 atomics) and written per (angle, orientation); the per-workgroup sums of the global parameters go to
// a partial array. dfm_pf_grad_reduce sums both over angles / workgroups in a fixed order.
#include "pf_cand.h"

namespace {

constexpr int GR_THREADS = 256;

struct GradArgs {
  dfm::CandGeo cg;
  const double* coords;      // probed directions (R, S, C, Se, P, 3)
  const float* axis_count;   // (A,)
  const double* normf;       // (K,)
  const float* qw;           // (S * Se) quadrature weights or null (the mean of the nodes)
  const double* dv_eta;      // d v / d eta_offset, shaped as coords (or null)
  const double* dv_t0;       // d v / d tilt_0
  const double* dv_t1;       // d v / d tilt_1
  const double* dv_tth;      // d v / d (2 theta_p), ring p's slice
  const float* sino;         // (n_ang, My, sks): S of the angle list, orientation-fastest
  const float* r;            // (R, My, CP): dL / d data
  const float* curv;         // (R, My, CP) or null
  const int* omegas;         // the angle list (n_ang)
  double* part_k;            // (n_ang, kb, nqk)
  double* part_g;            // (n_wg, 3 + P): eta, tilt 0, tilt 1, two theta per ring
  int n_ang, n_kb, n_jb, Se, My, CP, sks, nqk, hess;
};

__global__ void __launch_bounds__(GR_THREADS) pf_grad_kernel(GradArgs g) {
  extern __shared__ unsigned gr_lds[];
  const dfm::CandGeo& cg = g.cg;
  const int W = 32 * g.n_jb;
  char* base = reinterpret_cast<char*>(gr_lds);
  const int off_pre = W * 4, off_sb = off_pre + (W + 1) * 4, off_rb = off_sb + g.My * 32 * 4;
  const int off_cb = off_rb + g.My * 32 * 4;
  const int off_acc = (off_cb + (g.hess ? g.My * 32 * 4 : 0) + 7) & ~7;
  const int off_accg = off_acc + 32 * g.nqk * 8, off_ws = off_accg + (3 + cg.P) * 8;
  unsigned* bm = gr_lds;                                        // W words
  int* pre = reinterpret_cast<int*>(base + off_pre);            // W + 1 prefix counts
  float* sb = reinterpret_cast<float*>(base + off_sb);          // My x 32: S[a, y, k]
  float* rb = reinterpret_cast<float*>(base + off_rb);          // My x 32: r[a, y, j]
  float* cb = reinterpret_cast<float*>(base + off_cb);          // My x 32: curv (hessian)
  double* acc_k = reinterpret_cast<double*>(base + off_acc);    // 32 x nqk
  double* acc_g = reinterpret_cast<double*>(base + off_accg);   // 3 + P
  int* wsum = reinterpret_cast<int*>(base + off_ws);            // 4 wave totals

  const int kbi = blockIdx.x % g.n_kb;
  const int a = blockIdx.x / g.n_kb;
  if (a >= g.n_ang) return;
  const int r = g.omegas[a];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int i = tid; i < W; i += GR_THREADS) bm[i] = 0u;
  for (int i = tid; i < 32 * g.nqk; i += GR_THREADS) acc_k[i] = 0.0;
  for (int i = tid; i < 3 + cg.P; i += GR_THREADS) acc_g[i] = 0.0;
  for (int i = tid; i < g.My * 32; i += GR_THREADS) {
    const int y = i >> 5, kl = i & 31, kloc = kbi * 32 + kl;
    sb[i] = kloc < cg.kb ? g.sino[((int64_t)a * g.My + y) * g.sks + kloc] : 0.f;
  }
  __syncthreads();
  dfm::mark_candidates(cg, r, kbi, bm, tid, GR_THREADS);
  __syncthreads();

  // ---- exclusive prefix of the words' bit counts (each thread a run of words)
  const int per = (W + GR_THREADS - 1) / GR_THREADS;
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
  int run = incl - mine;
  for (int q = 0; q < wv; ++q) run += wsum[q];
  for (int w = w0; w < w1; ++w) {
    pre[w] = run;
    run += __popc(bm[w]);
  }
  if (w1 == W && w0 < W) pre[W] = run;
  if (W == 0 && tid == 0) pre[0] = 0;
  __syncthreads();

  const int CPn = g.CP, P = cg.P, A = cg.A, S = cg.S, Se = g.Se;
  const double inv_nodes = 1.0 / (double)(S * Se);
  for (int jb = 0; jb < g.n_jb; ++jb) {
    const int c_lo = pre[jb * 32], c_hi = pre[jb * 32 + 32];
    if (c_lo == c_hi) continue;                                // uniform
    for (int i = tid; i < g.My * 32; i += GR_THREADS) {
      const int y = i >> 5, jl = i & 31, j = jb * 32 + jl;
      const int64_t o = ((int64_t)r * g.My + y) * CPn + j;
      rb[i] = j < CPn ? g.r[o] : 0.f;
      if (g.hess) cb[i] = j < CPn ? g.curv[o] : 0.f;
    }
    __syncthreads();
    for (int ci = c_lo + tid; ci < c_hi; ci += GR_THREADS) {
      // the candidate's word (the last of the block's 32 with pre[w] <= ci) and bit (the rank-th set bit)
      int lo = jb * 32, hi = jb * 32 + 31;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= ci) lo = mid; else hi = mid - 1;
      }
      const int w = lo, kl = w & 31;
      unsigned bits = bm[w];
      for (int rk = ci - pre[w]; rk > 0; --rk) bits &= bits - 1u;
      const int jl = __builtin_ctz(bits), j = jb * 32 + jl;
      const int c = j / P, p = j - c * P;
      const int kg = cg.k0 + kbi * 32 + kl;
      double G = 0.0, C2 = 0.0;
      for (int y = 0; y < g.My; ++y) {
        const double sv = sb[y * 32 + kl];
        G += sv * (double)rb[y * 32 + jl];
        if (g.hess) C2 += sv * sv * (double)cb[y * 32 + jl];
      }
      const double is2 = cg.inv_sigma2[kg], nk = g.normf[kg];
      const double sig_fac1 = -2.0 * sqrt(is2), sig_fac3 = -2.0 * is2 * sqrt(is2);   // -2 / sigma, -2 / sigma^3
      double J0 = 0.0, J1 = 0.0, J2 = 0.0, J3 = 0.0, ge = 0.0, gt0 = 0.0, gt1 = 0.0, gtt = 0.0;
      const int a0 = cg.axis_start[p], a1 = cg.axis_start[p + 1];
      const double* pk = cg.poles + (int64_t)kg * A * 3;
      for (int s = 0; s < S; ++s) {
        for (int se = 0; se < Se; ++se) {
          const int64_t vo = (((((int64_t)r * S + s) * cg.C + c) * Se + se) * P + p) * 3;
          const double vx = g.coords[vo], vy = g.coords[vo + 1], vz = g.coords[vo + 2];
          const double wn = (g.qw ? (double)g.qw[s * Se + se] : inv_nodes) * nk;
          for (int ax = a0; ax < a1; ++ax) {
            const double nx = pk[ax * 3], ny = pk[ax * 3 + 1], nz = pk[ax * 3 + 2];
            const double d = nx * vx + ny * vy + nz * vz, ad = fabs(d);
            const double x1 = (1.0 - ad) * is2, x2 = (1.0 + ad) * is2;
            const double e1 = x1 < cg.cut ? exp(-x1) : 0.0, e2 = x2 < cg.cut ? exp(-x2) : 0.0;
            if (e1 == 0.0 && e2 == 0.0) continue;
            const double m = (double)g.axis_count[ax] * wn;
            const double dgd = (d < 0.0 ? -is2 : is2) * (e1 - e2) * m;
            const double gv = (e1 + e2) * m, dgs = (-(1.0 - ad) * e1 - (1.0 + ad) * e2) * m;
            J0 += dgd * (ny * vz - nz * vy);                  // n x v
            J1 += dgd * (nz * vx - nx * vz);
            J2 += dgd * (nx * vy - ny * vx);
            J3 += sig_fac1 * gv + sig_fac3 * dgs;
            if (g.dv_eta) ge += dgd * (nx * g.dv_eta[vo] + ny * g.dv_eta[vo + 1] + nz * g.dv_eta[vo + 2]);
            if (g.dv_t0) {
              gt0 += dgd * (nx * g.dv_t0[vo] + ny * g.dv_t0[vo + 1] + nz * g.dv_t0[vo + 2]);
              gt1 += dgd * (nx * g.dv_t1[vo] + ny * g.dv_t1[vo + 1] + nz * g.dv_t1[vo + 2]);
            }
            if (g.dv_tth) gtt += dgd * (nx * g.dv_tth[vo] + ny * g.dv_tth[vo + 1] + nz * g.dv_tth[vo + 2]);
          }
        }
      }
      double* ak = acc_k + kl * g.nqk;
      atomicAdd(ak + 0, G * J0);
      atomicAdd(ak + 1, G * J1);
      atomicAdd(ak + 2, G * J2);
      atomicAdd(ak + 3, G * J3);
      if (g.hess) {
        const double Jv[4] = {J0, J1, J2, J3};
        int q = 4;
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = u; v < 4; ++v) atomicAdd(ak + q++, C2 * Jv[u] * Jv[v]);
      }
      if (g.dv_eta) atomicAdd(acc_g + 0, G * ge);
      if (g.dv_t0) {
        atomicAdd(acc_g + 1, G * gt0);
        atomicAdd(acc_g + 2, G * gt1);
      }
      if (g.dv_tth) atomicAdd(acc_g + 3 + p, G * gtt);
    }
    __syncthreads();                                           // rb / cb are staged again for the next block
  }
  for (int i = tid; i < 32 * g.nqk; i += GR_THREADS) {
    const int kl = i / g.nqk, kloc = kbi * 32 + kl;
    if (kloc < cg.kb) g.part_k[((int64_t)a * cg.kb + kloc) * g.nqk + (i - kl * g.nqk)] = acc_k[i];
  }
  for (int i = tid; i < 3 + P; i += GR_THREADS) g.part_g[(int64_t)blockIdx.x * (3 + P) + i] = acc_g[i];
}

// out_k[k, q] = sum over the angle list (in order) of part_k[a, k, q]; out_g[i] = sum over the workgroups of
// part_g[w, i] (each thread a strided share in order, then a fixed tree): both deterministic.
__global__ void __launch_bounds__(256) grad_rows_kernel(const double* __restrict__ part, double* __restrict__ out,
                                                        int n_ang, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int a = 0; a < n_ang; ++a) s += part[(int64_t)a * n + i];
    out[i] = s;
  }
}

__global__ void __launch_bounds__(256) grad_sum_kernel(const double* __restrict__ part, double* __restrict__ out,
                                                       int64_t n_rows, int width) {
  __shared__ double red[256];
  const int i = blockIdx.x;
  double s = 0.0;
  for (int64_t w = threadIdx.x; w < n_rows; w += 256) s += part[w * width + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = red[0];
}

// out[a] = sum over y < my, k < kn of u[a, y, k] v[a, y, k] (float64; a workgroup per angle, fixed tree)
__global__ void __launch_bounds__(256) angle_dot_kernel(const float* __restrict__ u, const float* __restrict__ v,
                                                        double* __restrict__ out, int my, int sks, int kn) {
  __shared__ double red[256];
  const int a = blockIdx.x;
  double s = 0.0;
  const int64_t n = (int64_t)my * kn;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const int y = (int)(i / kn), k = (int)(i - (int64_t)y * kn);
    const int64_t o = ((int64_t)a * my + y) * sks + k;
    s += (double)u[o] * (double)v[o];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[a] = red[0];
}

}  // namespace

extern "C" {

int dfm_pf_grad(const double* coords, const double* poles, const int32_t* axis_start, const float* axis_count,
                const double* inv_sigma2, const double* norm, const float* qw, const double* frames,
                const double* sin_th, const double* cos_th, const double* dv_eta, const double* dv_t0,
                const double* dv_t1, const double* dv_tth, const float* sino, int32_t sks, const float* r,
                const float* curv, const int32_t* omegas, int32_t n_ang, int32_t k0, int32_t kb, int32_t My,
                int32_t C, int32_t P, int32_t A, int32_t S, int32_t Se, float pf_cut, double eta0, double deta,
                double xlo, double xhi, double* part_k, double* part_g, dfm_stream_t stream) {
  DFM_CHECK_ARG(coords && poles && axis_start && axis_count && inv_sigma2 && norm && frames && sin_th && cos_th &&
                    sino && r && omegas && part_k && part_g && (!dv_t0 == !dv_t1),
                "dfm_pf_grad: null pointer");
  DFM_CHECK_ARG(n_ang >= 0 && kb > 0 && k0 >= 0 && My > 0 && C > 0 && P > 0 && A >= P && S > 0 && Se > 0 &&
                    sks >= kb && deta > 0.0 && xlo <= xhi && C * P <= 65536,
                "dfm_pf_grad: bad sizes");
  if (n_ang == 0) return DFM_OK;
  GradArgs g{};
  g.cg = dfm::CandGeo{poles, frames, sin_th, cos_th, axis_start, inv_sigma2, k0, kb, C, P, A, S, (double)pf_cut,
                      eta0, deta, xlo, xhi};
  g.coords = coords;
  g.axis_count = axis_count;
  g.normf = norm;
  g.qw = qw;
  g.dv_eta = dv_eta;
  g.dv_t0 = dv_t0;
  g.dv_t1 = dv_t1;
  g.dv_tth = dv_tth;
  g.sino = sino;
  g.r = r;
  g.curv = curv;
  g.omegas = omegas;
  g.part_k = part_k;
  g.part_g = part_g;
  g.n_ang = n_ang;
  g.n_kb = (kb + 31) / 32;
  g.n_jb = (C * P + 31) / 32;
  g.Se = Se;
  g.My = My;
  g.CP = C * P;
  g.sks = sks;
  g.hess = curv != nullptr;
  g.nqk = g.hess ? 14 : 4;
  const int W = 32 * g.n_jb;
  // the kernel's layout: bitmap, prefix counts, S / r (/ curv) columns, then (8-byte aligned) the accumulators
  const size_t head = (size_t)(2 * W + 1) * 4 + (size_t)(g.hess ? 3 : 2) * My * 32 * 4;
  const size_t lds = ((head + 7) & ~(size_t)7) + (size_t)(32 * g.nqk + 3 + P) * 8 + 4 * sizeof(int);
  DFM_CHECK_ARG(lds <= 163840, "dfm_pf_grad: %d segment blocks and %d translations exceed the LDS", g.n_jb, My);
  const int64_t n_wg = (int64_t)n_ang * g.n_kb;
  DFM_CHECK_ARG(n_wg < ((int64_t)1 << 31), "dfm_pf_grad: too many workgroups");
  hipLaunchKernelGGL(pf_grad_kernel, dim3((unsigned)n_wg), dim3(GR_THREADS), lds, (hipStream_t)stream, g);
  DFM_LAUNCH_CHECK("dfm_pf_grad");
  return DFM_OK;
}

int dfm_pf_grad_reduce(const double* part_k, const double* part_g, int32_t n_ang, int32_t kb, int32_t nqk,
                       int64_t n_wg, int32_t n_glob, double* out_k, double* out_g, dfm_stream_t stream) {
  DFM_CHECK_ARG(part_k && part_g && out_k && out_g && n_ang > 0 && kb > 0 && nqk > 0 && n_wg > 0 && n_glob > 0,
                "dfm_pf_grad_reduce: bad arguments");
  const int64_t n = (int64_t)kb * nqk;
  hipLaunchKernelGGL(grad_rows_kernel, dim3(dfm::grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     part_k, out_k, n_ang, n);
  hipLaunchKernelGGL(grad_sum_kernel, dim3(n_glob), dim3(256), 0, (hipStream_t)stream, part_g, out_g, n_wg, n_glob);
  DFM_LAUNCH_CHECK("dfm_pf_grad_reduce");
  return DFM_OK;
}

int dfm_angle_dots(const float* u, const float* v, int32_t n_ang, int32_t my, int32_t sks, int32_t kn, double* out,
                   dfm_stream_t stream) {
  DFM_CHECK_ARG(u && v && out && n_ang >= 0 && my > 0 && kn >= 0 && sks >= kn, "dfm_angle_dots: bad arguments");
  if (n_ang == 0) return DFM_OK;
  hipLaunchKernelGGL(angle_dot_kernel, dim3(n_ang), dim3(256), 0, (hipStream_t)stream, u, v, out, my, sks, kn);
  DFM_LAUNCH_CHECK("dfm_angle_dots");
  return DFM_OK;
}

}  // extern "C"
