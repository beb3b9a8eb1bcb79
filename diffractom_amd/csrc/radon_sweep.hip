// Forward projection by line sweep (parallel_radon.py:92-124: the trapezoid / hat footprint weights, radon_kernels.cl
// radon_forward_k):
//
//   sino[a, s, k] = sum over pixels with |t_a(p) - s| < hw_a of scale_a * min(hw_a - |t_a(p) - s|, top_a) img[p, k]
//   t_a(x, y) = X_a x + Y_a y + T0_a
//
// The table-driven projection gathers every (pixel, orientation) value from L2 once per angle and tap (44 GB per
// subset at 99x99 pixels and 29713 orientations), which bounds it by the vector-memory address rate. Here a workgroup
// owns 16 orientations (4 quads per lane) and a group of angles of one sweep class, and walks the image line by line
// (rows for angles with |X| >= |Y|, columns otherwise, so that every ray meets a line at one to three pixels): RB lines
// of the 16-orientation image slice are staged in LDS (double buffered, the next block's loads in flight during the
// current block's sums), and each thread, one ray, finds its pixels on each staged line from the geometry (float64,
// as the host tables) and adds them from LDS. The image is read from L2 once per angle group instead of once per
// angle and tap, the gathers become LDS reads, and no table is read.
#include "dfm_common.h"

namespace {

constexpr int N_XCD = 8;
constexpr int SW_THREADS = 1024;

__device__ __forceinline__ bool xcd_split(int per_blk, int n_blk, int& blk, int& local) {
  const int i = blockIdx.x;
  const int xcd = i % N_XCD;
  const int j = i / N_XCD;
  blk = (j / per_blk) * N_XCD + xcd;
  local = j % per_blk;
  return blk < n_blk;
}

struct SweepGeo {
  double X, Y, T0, hw, top, scale;
};

// CQ orientation quads per lane (4 CQ orientations per workgroup); LDS pixel stride CQ + 1 float4, so rays at
// neighbouring pixels start on distinct banks
template <int RB, int CQ, int TH>
__global__ void __launch_bounds__(TH) radon_forward_sweep_kernel(
    const float4* __restrict__ img, float4* __restrict__ sino, const SweepGeo* __restrict__ geo,
    const int* __restrict__ grp_ang, const int* __restrict__ grp_cls, int n_grp, int na_grp, int ns, int nx, int ny,
    int kstride4, int k4_off, int k4_n, int skstride4) {
  constexpr int SW_CQ = CQ, SW_PS = CQ + 1;
  extern __shared__ float4 limg[];   // 2 buffers x RB lines x len pixels x SW_PS float4
  const int n_cb = (k4_n + SW_CQ - 1) / SW_CQ;
  int cb, g;
  if (!xcd_split(n_grp, n_cb, cb, g)) return;                 // uniform over the workgroup
  const int cls = grp_cls[g];
  const int len = cls ? ny : nx, nlines = cls ? nx : ny;
  const int buf = RB * len * SW_PS;
  const int t = threadIdx.x;
  const int al = t / ns, s = t - al * ns;
  const int a = al < na_grp ? grp_ang[g * na_grp + al] : -1;
  const bool live = a >= 0;
  // along-line coefficient (|coef| >= |cross| by the class), the line's offset, the footprint
  double coef = 1.0, cross = 0.0, off0 = 0.0, hw = 0.0, top = 0.0, scale = 0.0;
  if (live) {
    const SweepGeo gg = geo[a];
    coef = cls ? gg.Y : gg.X;
    cross = cls ? gg.X : gg.Y;
    off0 = gg.T0 - (double)s;
    hw = gg.hw;
    top = gg.top;
    scale = gg.scale;
  }
  const double inv = 1.0 / coef, hinv = fabs(hw * inv);
  float4 acc[SW_CQ];
#pragma unroll
  for (int qq = 0; qq < SW_CQ; ++qq) acc[qq] = make_float4(0.f, 0.f, 0.f, 0.f);
  const int q0 = cb * SW_CQ;

  // staging element i = t + 1024 u of a block: quad i % SW_CQ of pixel (line l, position p), i / SW_CQ = l len + p;
  // RB * len * SW_CQ <= RB * CQ * 256 for len <= 256
  constexpr int NV = (RB * CQ * 256 + TH - 1) / TH;
  float4 v[NV];
  auto stage_load = [&](int L0) {
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      const int i = t + TH * u;
      const int q = i & (SW_CQ - 1), pp = i / SW_CQ;
      const int l = pp / len, pos = pp - l * len;
      v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (l < RB && L0 + l < nlines && q0 + q < k4_n) {
        const int64_t p = cls ? (int64_t)pos * nx + (L0 + l) : (int64_t)(L0 + l) * nx + pos;
        v[u] = img[p * kstride4 + k4_off + q0 + q];
      }
    }
  };
  auto stage_store = [&](float4* lb) {
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      const int i = t + TH * u;
      const int q = i & (SW_CQ - 1), pp = i / SW_CQ;
      if (pp < RB * len) lb[pp * SW_PS + q] = v[u];
    }
  };

  stage_load(0);
  for (int L0 = 0, bi = 0; L0 < nlines; L0 += RB, ++bi) {
    float4* lb = limg + (bi & 1) * buf;
    // buffer bi & 1 was last read for block bi - 2, which every wave finished before the barrier of block bi - 1
    stage_store(lb);
    __syncthreads();
    if (L0 + RB < nlines) stage_load(L0 + RB);
    if (live) {
      const double cb0 = cross * (double)L0 + off0;
      for (int l = 0; l < RB && L0 + l < nlines; ++l) {
        // positions p with |coef p + c| < hw on line L0 + l: p in (-c / coef -+ hw / |coef|)
        const double c = cb0 + cross * (double)l;
        const double mid = -c * inv;
        const int p0 = max((int)ceil(mid - hinv), 0), p1 = min((int)floor(mid + hinv), len - 1);
        const float4* row = lb + l * len * SW_PS;
        double d = coef * (double)p0 + c;                      // the footprint offset, stepped by coef
        for (int p = p0; p <= p1; ++p, d += coef) {
          const double k = hw - fabs(d);
          if (k > 0.0) {
            const float w = (float)(scale * fmin(k, top));
            const float4* px = row + p * SW_PS;
#pragma unroll
            for (int qq = 0; qq < SW_CQ; ++qq) {
              const float4 x = px[qq];
              acc[qq].x += w * x.x;
              acc[qq].y += w * x.y;
              acc[qq].z += w * x.z;
              acc[qq].w += w * x.w;
            }
          }
        }
      }
    }
  }
  if (live) {
    // plain stores: a thread's CQ quads are one 64-128 B run of its ray's row, the lanes' runs lie a row apart, so
    // each store instruction touches 64 lines; L2 merges the runs into whole lines (non-temporal partial-line
    // stores went to HBM as 4x the bytes)
    const int64_t r = (int64_t)a * ns + s;
#pragma unroll
    for (int qq = 0; qq < SW_CQ; ++qq)
      if (q0 + qq < k4_n) sino[r * skstride4 + q0 + qq] = acc[qq];
  }
}

}  // namespace

extern "C" {

int dfm_radon_sweep_lines(int32_t nx, int32_t ny) {
  const int len = nx > ny ? nx : ny;
  if (len <= 0 || len > 256) return 0;
  for (int rb = 8; rb >= 2; rb /= 2)
    if ((size_t)2 * rb * len * 5 * sizeof(float4) <= 163840) return rb;
  return 0;
}

int dfm_radon_forward_sweep(const float* img_k, float* sino, const double* geo, const int32_t* grp_ang,
                            const int32_t* grp_cls, int32_t n_grp, int32_t na_grp, int32_t ns, int32_t nx, int32_t ny,
                            int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride, int32_t lines,
                            dfm_stream_t stream) {
  DFM_CHECK_ARG(img_k && sino && geo && grp_ang && grp_cls, "dfm_radon_forward_sweep: null pointer");
  DFM_CHECK_ARG(n_grp >= 0 && na_grp >= 1 && ns >= 1 && nx >= 1 && ny >= 1 &&
                    nx <= 256 && ny <= 256 && kstride % 4 == 0 && k_off % 4 == 0 && k_n % 4 == 0 && k_off >= 0 &&
                    k_n >= 0 && k_off + k_n <= kstride && skstride % 4 == 0 && skstride >= k_n,
                "dfm_radon_forward_sweep: bad sizes (na_grp * ns <= 1024, image <= 256 x 256)");
  // variant = lines per LDS block (1, 2, 4, 8) + 16 w: 16 (w 0), 32 (w 1) or 64 (w 2) orientations per workgroup,
  // + 64 for workgroups of 512 threads (na_grp * ns <= 512; 32 or 64 orientations); instantiated: 16 x 2/4/8 and
  // 32 x 1/2/4 at 1024 threads, 32 x 2 and 64 x 1/2 at 512 (the others spill registers)
  const int len = nx > ny ? nx : ny;
  const int rb = lines % 16, w = (lines / 16) % 4, half = lines / 64;
  const int cq = 4 << (w > 2 ? 0 : w);
  const int th = half ? 512 : SW_THREADS;
  const bool inst = half ? ((w == 2 && rb <= 2) || (w == 1 && rb == 2)) : ((w == 0 && rb >= 2) || (w == 1 && rb <= 4));
  DFM_CHECK_ARG((rb == 1 || rb == 2 || rb == 4 || rb == 8) && w <= 2 && half <= 1 && inst && na_grp * ns <= th &&
                    (size_t)2 * rb * len * (cq + 1) * sizeof(float4) <= 163840,
                "dfm_radon_forward_sweep: variant %d: %d lines of %d pixels x %d orientations, %d threads not "
                "available", lines, rb, len, 4 * cq, th);
  if (n_grp == 0 || k_n == 0) return DFM_OK;
  const int k4n = k_n / 4;
  const int n_cb = (k4n + cq - 1) / cq;
  const unsigned grid = (unsigned)((n_cb + N_XCD - 1) / N_XCD * N_XCD * n_grp);
  const size_t lds = (size_t)2 * rb * len * (cq + 1) * sizeof(float4);
  const hipStream_t st = (hipStream_t)stream;
#define DFM_SW(RB, CQ, TH)                                                                                        \
  hipLaunchKernelGGL((radon_forward_sweep_kernel<RB, CQ, TH>), dim3(grid), dim3(TH), lds, st,                     \
                     (const float4*)img_k, (float4*)sino, (const SweepGeo*)geo, grp_ang, grp_cls, n_grp, na_grp, ns,  \
                     nx, ny, kstride / 4, k_off / 4, k4n, skstride / 4)
  if (half) {
    if (cq == 8) DFM_SW(2, 8, 512);
    else if (rb == 1) DFM_SW(1, 16, 512);
    else DFM_SW(2, 16, 512);
  } else if (cq == 4) {
    if (rb == 2) DFM_SW(2, 4, 1024);
    else if (rb == 4) DFM_SW(4, 4, 1024);
    else DFM_SW(8, 4, 1024);
  } else {
    if (rb == 1) DFM_SW(1, 8, 1024);
    else if (rb == 2) DFM_SW(2, 8, 1024);
    else DFM_SW(4, 8, 1024);
  }
#undef DFM_SW
  DFM_LAUNCH_CHECK("dfm_radon_forward_sweep");
  return DFM_OK;
}

}  // extern "C"
