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

// ------------------------------------------------------------------------------------------ ray-pair sweep
// The same line sweep with two neighbouring rays (bins 2 sp and 2 sp + 1 of one angle) per thread. On a line the two
// rays' footprints lie within 4 consecutive pixels p0 .. p0 + 3, the first ray's (the one further left on the line)
// within p0 .. p0 + 2 and the other's within p0 + 1 .. p0 + 3, when 2 hw / |coef| + 1 / |coef| < 4 (any square
// image with the trapezoid footprint); so each of the 4 pixels is read from LDS once for both rays (4 reads of 32
// orientations instead of 6) and every candidate's weight is computed from the geometry (0 outside the footprint):
// no per-lane loops, no divergent branches. The line geometry is float64 (per line, not per pixel), the six weights
// float32 with one v_med3 clamp each. Pixels are 8 float4 (32 orientations) apart in LDS and a lane reads quad
// (qq + lane) % 8 at step qq, so lanes 0-7 (and 8-15) of a read hit 8 distinct 16-byte bank groups whatever pixels
// they read (the DUP layout below also keeps the two halves apart); two zero pixels pad each staged line on both
// sides, so the candidates need no bounds checks. Staging addresses are computed once per thread (the block's line
// offset is added per block).
constexpr int PR_TH = 512;

// PIPE: the next block's image loads are issued before the current block's sums and held in registers across them
// (about 155 VGPRs: one 512-thread workgroup per CU); otherwise each block's loads are issued, waited for and stored
// at its start, their addresses formed there from the thread index (at most 128 VGPRs: two workgroups per CU, the
// other one's sums hiding the loads)
//
// DUP: a b128 LDS read serves 16 lanes per cycle from 64 banks (16 groups of 16 bytes); with pixels 8 quads apart a
// lane's group is 8 (pixel parity) + its quad, so lanes i and i + 8 (the same rotation) collide whenever their
// pixels have equal parity: 1.8 cycles per 16 lanes on average at al_c1. In the DUP layout every staged pixel is
// held twice, in the first half of its 16-quad block and in the second half of the block before it; lanes 0-7 of
// each 16 read the first copies and lanes 8-15 the second (groups 0-7 and 8-15), so no two lanes of a cycle meet,
// for twice the LDS and twice the staging stores.
template <int RB, int NV, bool PIPE, bool DUP>
__global__ void __launch_bounds__(PR_TH, PIPE ? 1 : 4) radon_forward_pairs_sweep_kernel(
    const float4* __restrict__ img, float4* __restrict__ sino, const SweepGeo* __restrict__ geo,
    const int* __restrict__ grp_ang, const int* __restrict__ grp_cls, int n_grp, int na_grp, int ns, int nx, int ny,
    int kstride4, int k4_off, int k4_n, int skstride4) {
  constexpr int CQ = 8;
  extern __shared__ float4 limg[];   // 2 buffers x RB lines x (len + 4) pixels x CQ float4 (DUP: (len + 5) x 2 CQ)
  constexpr int PS = DUP ? 2 * CQ : CQ;                      // float4 per pixel slot
  const int n_cb = (k4_n + CQ - 1) / CQ;
  int cb, g;
  if (!xcd_split(n_grp, n_cb, cb, g)) return;                 // uniform over the workgroup
  const int cls = grp_cls[g];
  const int len = cls ? ny : nx, nlines = cls ? nx : ny;
  const int row = (len + 4 + (DUP ? 1 : 0)) * PS;          // float4 per staged line
  const int buf = RB * row;
  const int t = threadIdx.x, lane = t & 63;
  const int npair = (ns + 1) >> 1;
  const int al = t / npair, sp = t - al * npair;
  const int a = al < na_grp ? grp_ang[g * na_grp + al] : -1;
  const bool live = a >= 0;
  const int q0 = cb * CQ;

  // ---- zero pads (pixels -2, -1, len, len + 1 of every line of both buffers, both copies); never overwritten. Slot
  // S (pixel S - 2) lies at S PS + q, or with DUP at (S + 1) PS + q and S PS + CQ + q
  for (int i = t; i < 2 * RB * 4 * CQ; i += PR_TH) {
    const int q = i % CQ, pp = (i / CQ) % 4, l = i / (4 * CQ);       // l over 2 * RB lines
    const int px = pp < 2 ? pp : len + pp;                            // slots 0, 1, len + 2, len + 3
    if (DUP) {
      limg[l * row + (px + 1) * PS + q] = make_float4(0.f, 0.f, 0.f, 0.f);
      limg[l * row + px * PS + CQ + q] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      limg[l * row + px * CQ + q] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }

  // ---- staging slots of this thread: element i = t + PR_TH u of a block = (line l, position pos, quad q)
  int lds_u[NV];
  int64_t gof_u[NV];
  bool ok_u[NV];
  const int64_t line_step = cls ? (int64_t)kstride4 : (int64_t)nx * kstride4;
  auto slots = [&](int tt) {
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      const int i = tt + PR_TH * u;
      const int q = i & (CQ - 1), pp = i / CQ;
      const int l = pp / len, pos = pp - l * len;
      ok_u[u] = l < RB && q0 + q < k4_n;
      lds_u[u] = l * row + (pos + 2 + (DUP ? 1 : 0)) * PS + q;    // (DUP: the first copy; the second CQ before)
      const int64_t pix = cls ? (int64_t)pos * nx + l : (int64_t)l * nx + pos;
      gof_u[u] = pix * kstride4 + k4_off + q0 + q;
    }
  };
  if (PIPE) slots(t);
  float4 v[NV];
  auto stage_load = [&](int L0) {
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      const int i = t + PR_TH * u;
      if (ok_u[u] && L0 + i / (CQ * len) < nlines) v[u] = img[gof_u[u] + (int64_t)L0 * line_step];
    }
  };
  auto stage_store = [&](float4* lb) {
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      const int i = t + PR_TH * u;
      if (i < RB * len * CQ) {
        lb[lds_u[u]] = v[u];
        if (DUP) lb[lds_u[u] - CQ] = v[u];
      }
    }
  };

  // ---- ray-pair geometry: mid(l) of the left ray = m1 l + mL0 (float64), footprint hinv, float32 weight terms
  double m1 = 0.0, mL0 = 0.0, hinv = 0.0, ainv = 1.0;
  float hs = 0.f, ts = 0.f, ds = 0.f;
  int bin_l = 0, bin_r = 0;
  if (live) {
    const SweepGeo gg = geo[a];
    const double coef = cls ? gg.Y : gg.X, cross = cls ? gg.X : gg.Y;
    const double inv = 1.0 / coef;
    const int s = 2 * sp;
    m1 = -cross * inv;
    mL0 = -(gg.T0 - (double)s) * inv + (inv < 0.0 ? inv : 0.0);   // the smaller of the two rays' mids
    ainv = fabs(inv);
    hinv = gg.hw * ainv;
    hs = (float)(gg.hw * gg.scale);
    ts = (float)(gg.top * gg.scale);
    ds = (float)(coef * gg.scale);
    bin_l = inv > 0.0 ? s : s + 1;
    bin_r = inv > 0.0 ? s + 1 : s;
  }
  const float c1 = ds, c2 = 2.f * ds, cr1 = (float)((1.0 - ainv) * ds), cr2 = (float)((2.0 - ainv) * ds),
              cr3 = (float)((3.0 - ainv) * ds);
  const int rot = (lane & 7);                               // this lane's quad rotation
  const int hoff = DUP ? (lane & 8) : 0;                     // DUP: lanes 8-15 of 16 read the second copies
  float4 accL[CQ], accR[CQ];
#pragma unroll
  for (int qq = 0; qq < CQ; ++qq) accL[qq] = accR[qq] = make_float4(0.f, 0.f, 0.f, 0.f);

  if (PIPE) stage_load(0);
  for (int L0 = 0, bi = 0; L0 < nlines; L0 += RB, ++bi) {
    float4* lb = limg + (bi & 1) * buf;
    // buffer bi & 1 was last read for block bi - 2, which every wave finished before the barrier of block bi - 1
    if (!PIPE) {
      // the slots from a thread index passed through an empty asm: formed here, not hoisted to stay live
      int tt = t;
      asm volatile("" : "+v"(tt));
      slots(tt);
      stage_load(L0);
    }
    stage_store(lb);
    __syncthreads();
    if (PIPE && L0 + RB < nlines) stage_load(L0 + RB);
    if (live) {
#pragma unroll
      for (int l = 0; l < RB; ++l) {
        if (L0 + l >= nlines) break;
        const double midL = fma(m1, (double)(L0 + l), mL0);
        double p0 = ceil(midL - hinv);
        p0 = fmin(fmax(p0, -2.0), (double)(len - 2));
        const float uu = (float)(p0 - midL);
        // the six weights: left ray candidates p0, p0 + 1, p0 + 2; right ray p0 + 1 .. p0 + 3
        const float wl0 = __builtin_amdgcn_fmed3f(hs - fabsf(ds * uu), 0.f, ts);
        const float wl1 = __builtin_amdgcn_fmed3f(hs - fabsf(fmaf(ds, uu, c1)), 0.f, ts);
        const float wl2 = __builtin_amdgcn_fmed3f(hs - fabsf(fmaf(ds, uu, c2)), 0.f, ts);
        const float wr1 = __builtin_amdgcn_fmed3f(hs - fabsf(fmaf(ds, uu, cr1)), 0.f, ts);
        const float wr2 = __builtin_amdgcn_fmed3f(hs - fabsf(fmaf(ds, uu, cr2)), 0.f, ts);
        const float wr3 = __builtin_amdgcn_fmed3f(hs - fabsf(fmaf(ds, uu, cr3)), 0.f, ts);
        const float4* base = lb + l * row + ((int)p0 + 2 + (DUP ? 1 : 0)) * PS - hoff;
#pragma unroll
        for (int qq = 0; qq < CQ; ++qq) {
          const float4* px = base + ((qq + rot) & (CQ - 1));
          const float4 x0 = px[0], x1 = px[PS], x2 = px[2 * PS], x3 = px[3 * PS];
          // FMA chains into the accumulators (acc += w0 x0 + w1 x1 + w2 x2 as written would cost a multiply and an
          // add besides the two FMAs: 4 VALU operations per 3 taps instead of 3)
#define DFM_PS3(acc, c, wa, xa, wb, xb, wc, xc) acc.c = fmaf(wc, xc.c, fmaf(wb, xb.c, fmaf(wa, xa.c, acc.c)));
          DFM_PS3(accL[qq], x, wl0, x0, wl1, x1, wl2, x2) DFM_PS3(accL[qq], y, wl0, x0, wl1, x1, wl2, x2)
          DFM_PS3(accL[qq], z, wl0, x0, wl1, x1, wl2, x2) DFM_PS3(accL[qq], w, wl0, x0, wl1, x1, wl2, x2)
          DFM_PS3(accR[qq], x, wr1, x1, wr2, x2, wr3, x3) DFM_PS3(accR[qq], y, wr1, x1, wr2, x2, wr3, x3)
          DFM_PS3(accR[qq], z, wr1, x1, wr2, x2, wr3, x3) DFM_PS3(accR[qq], w, wr1, x1, wr2, x2, wr3, x3)
#undef DFM_PS3
        }
      }
    }
  }
  // ---- stores through a wave-private LDS transpose: lane l holds its two rays' 8 quads (128 B of a sinogram row
  // each); after the transpose, store instruction j has lanes 8 i .. 8 i + 7 write the 8 quads of lane 8 j + i's
  // row, i.e. 8 whole 128-byte lines per instruction (with per-lane 16-byte stores the L2 wrote the rows back in
  // 32-byte pieces: twice the sinogram's bytes)
  __syncthreads();                                           // every wave is done with the staged lines
  float4* wl = limg + (t >> 6) * (64 * CQ);
  const int src = lane >> 3, qd = lane & (CQ - 1);
#pragma unroll
  for (int side = 0; side < 2; ++side) {
#pragma unroll
    for (int qq = 0; qq < CQ; ++qq) wl[lane * CQ + ((qq + rot) & (CQ - 1))] = side ? accR[qq] : accL[qq];
    __builtin_amdgcn_wave_barrier();
    const int bin = side ? bin_r : bin_l;
#pragma unroll
    for (int j = 0; j < CQ; ++j) {
      const int sl = 8 * j + src;                             // the lane whose row this lane stores a quad of
      const float4 val = wl[sl * CQ + qd];
      const int a_s = __shfl(a, sl, 64), bin_s = __shfl(bin, sl, 64);
      if (a_s >= 0 && bin_s < ns && q0 + qd < k4_n) sino[((int64_t)a_s * ns + bin_s) * skstride4 + q0 + qd] = val;
    }
    __builtin_amdgcn_wave_barrier();
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

int dfm_radon_forward_pairs_sweep(const float* img_k, float* sino, const double* geo, const int32_t* grp_ang,
                                  const int32_t* grp_cls, int32_t n_grp, int32_t na_grp, int32_t ns, int32_t nx,
                                  int32_t ny, int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride,
                                  int32_t lines, dfm_stream_t stream) {
  DFM_CHECK_ARG(img_k && sino && geo && grp_ang && grp_cls, "dfm_radon_forward_pairs_sweep: null pointer");
  const int len = nx > ny ? nx : ny;
  DFM_CHECK_ARG(n_grp >= 0 && na_grp >= 1 && ns >= 1 && nx >= 1 && ny >= 1 && len <= 256 && kstride % 4 == 0 &&
                    k_off % 4 == 0 && k_n % 4 == 0 && k_off >= 0 && k_n >= 0 && k_off + k_n <= kstride &&
                    skstride % 4 == 0 && skstride >= k_n && na_grp * ((ns + 1) / 2) <= PR_TH,
                "dfm_radon_forward_pairs_sweep: bad sizes (na_grp * ceil(ns / 2) <= 512, image <= 256 x 256)");
  // lines: lines per staged block (1, 2, 4), + 8 for the form that loads each block at its start (no loads held
  // across the sums: two workgroups per CU), or + 16 for the duplicated-pixel layout (lines 1 or 2; conflict-free LDS
  // reads, twice the LDS)
  const int rb = lines & 7;
  const bool pipe = (lines & 8) == 0, dup = (lines & 16) != 0;
  DFM_CHECK_ARG((rb == 1 || rb == 2 || (rb == 4 && len <= 128)) && (lines & ~31) == 0 &&
                    (pipe || rb == 1 || (rb == 2 && len <= 128)) &&
                    (!dup || (pipe && (rb == 1 || (rb == 2 && len <= 128)))),
                "dfm_radon_forward_pairs_sweep: lines %d not 1, 2 or (lines <= 128 pixels) 4, or 1 + 8 or (lines <= "
                "128 pixels) 2 + 8, or 1 + 16 or (lines <= 128 pixels) 2 + 16", lines);
  // staged lines, or the store transpose (64 lanes x 8 quads per wave), whichever is larger
  size_t lds = dup ? (size_t)2 * rb * (len + 5) * 16 * sizeof(float4)
                   : (size_t)2 * rb * (len + 4) * 8 * sizeof(float4);
  if (lds < (size_t)(PR_TH / 64) * 64 * 8 * sizeof(float4)) lds = (size_t)(PR_TH / 64) * 64 * 8 * sizeof(float4);
  DFM_CHECK_ARG(lds <= 163840, "dfm_radon_forward_pairs_sweep: %d lines of %d pixels exceed the LDS", rb, len);
  if (n_grp == 0 || k_n == 0) return DFM_OK;
  const int k4n = k_n / 4;
  const int n_cb = (k4n + 7) / 8;
  const unsigned grid = (unsigned)((n_cb + N_XCD - 1) / N_XCD * N_XCD * n_grp);
  const hipStream_t st = (hipStream_t)stream;
  // staging elements per thread: lines * len * 8 float4 over 512 threads (len <= 128 or <= 256)
  const bool small = len <= 128;
#define DFM_PS(RB, NV)              \
  if (dup) {                        \
    DFM_PS1(RB, NV, true, true);    \
  } else if (pipe) {                \
    DFM_PS1(RB, NV, true, false);   \
  } else {                          \
    DFM_PS1(RB, NV, false, false);  \
  }
#define DFM_PS1(RB, NV, P, D)                                                                                    \
  hipLaunchKernelGGL((radon_forward_pairs_sweep_kernel<RB, NV, P, D>), dim3(grid), dim3(PR_TH), lds, st,           \
                     (const float4*)img_k, (float4*)sino, (const SweepGeo*)geo, grp_ang, grp_cls, n_grp, na_grp, ns, \
                     nx, ny, kstride / 4, k_off / 4, k4n, skstride / 4)
  if (rb == 1) {
    if (small) { DFM_PS(1, 2) } else { DFM_PS(1, 4) }
  } else if (rb == 2) {
    if (small) { DFM_PS(2, 4) } else { DFM_PS1(2, 8, true, false); }
  } else {
    DFM_PS1(4, 8, true, false);
  }
#undef DFM_PS1
#undef DFM_PS
  DFM_LAUNCH_CHECK("dfm_radon_forward_pairs_sweep");
  return DFM_OK;
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
  // variant = lines per LDS block + 16 w: 16 orientations per workgroup (w 0) with 2, 4 or 8 lines, or 32 (w 1) with
  // 1 or 2 lines (1024 threads each)
  const int len = nx > ny ? nx : ny;
  const int rb = lines % 16, w = lines / 16;
  const int cq = w ? 8 : 4;
  const bool inst = w == 0 ? (rb == 2 || rb == 4 || rb == 8) : (w == 1 && (rb == 1 || rb == 2));
  DFM_CHECK_ARG(inst && na_grp * ns <= SW_THREADS && (size_t)2 * rb * len * (cq + 1) * sizeof(float4) <= 163840,
                "dfm_radon_forward_sweep: variant %d: %d lines of %d pixels x %d orientations not available (16 "
                "orientations x 2, 4, 8 lines or 32 x 1, 2 lines, within the LDS; na_grp * ns <= 1024)",
                lines, rb, len, 4 * cq);
  if (n_grp == 0 || k_n == 0) return DFM_OK;
  const int k4n = k_n / 4;
  const int n_cb = (k4n + cq - 1) / cq;
  const unsigned grid = (unsigned)((n_cb + N_XCD - 1) / N_XCD * N_XCD * n_grp);
  const size_t lds = (size_t)2 * rb * len * (cq + 1) * sizeof(float4);
  const hipStream_t st = (hipStream_t)stream;
#define DFM_SW(RB, CQ)                                                                                            \
  hipLaunchKernelGGL((radon_forward_sweep_kernel<RB, CQ, SW_THREADS>), dim3(grid), dim3(SW_THREADS), lds, st,     \
                     (const float4*)img_k, (float4*)sino, (const SweepGeo*)geo, grp_ang, grp_cls, n_grp, na_grp, ns,  \
                     nx, ny, kstride / 4, k_off / 4, k4n, skstride / 4)
  if (cq == 4) {
    if (rb == 2) DFM_SW(2, 4);
    else if (rb == 4) DFM_SW(4, 4);
    else DFM_SW(8, 4);
  } else {
    if (rb == 1) DFM_SW(1, 8);
    else DFM_SW(2, 8);
  }
#undef DFM_SW
  DFM_LAUNCH_CHECK("dfm_radon_forward_sweep");
  return DFM_OK;
}

}  // extern "C"
