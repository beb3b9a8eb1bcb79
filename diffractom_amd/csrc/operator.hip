// Forward/adjoint kernels of the single-phase operator on gfx950: channel-fastest transposes, the parallel-beam
// Radon projector over many channels, the dense pole-figure evaluation with its CSR conversion, and the CSR PF
// products. Wave64 throughout; channels (orientations) run across the lanes, so every image and sinogram access of
// a wave is one contiguous 1 KiB (float4 per lane) and the ray geometry is uniform across the wave.
#include <stdarg.h>

#include "dfm_common.h"

namespace dfm {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ------------------------------------------------------------------------------------------------ transposes
// 64 x 64 tiles through LDS (+1 padding against bank conflicts), 256 threads.
constexpr int TT = 64;

__global__ void __launch_bounds__(256) gather_k_kernel(const float* __restrict__ coeffs, float* __restrict__ img_k,
                                                       int64_t npix, int64_t k0, int kb, int kstride) {
  __shared__ float tile[TT][TT + 1];
  const int64_t p0 = (int64_t)blockIdx.x * TT;
  const int kk0 = blockIdx.y * TT;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // ty 0..3
  // read coeffs[k0 + kk, p] rows kk (tile rows), pixels along the lanes
  for (int r = ty; r < TT; r += 4) {
    const int kk = kk0 + r;
    const int64_t p = p0 + tx;
    float v = 0.0f;
    if (kk < kb && p < npix) v = coeffs[(k0 + kk) * npix + p];
    tile[r][tx] = v;
  }
  __syncthreads();
  // write img_k[p, kk] with the channels along the lanes (zeros for kb <= kk < kstride)
  for (int r = ty; r < TT; r += 4) {
    const int64_t p = p0 + r;
    const int kk = kk0 + tx;
    if (p < npix && kk < kstride) img_k[p * kstride + kk] = tile[tx][r];
  }
}

__global__ void __launch_bounds__(256) scatter_k_kernel(const float* __restrict__ img_k, float* __restrict__ coeffs,
                                                        int64_t npix, int64_t k0, int kb, int kstride) {
  __shared__ float tile[TT][TT + 1];
  const int64_t p0 = (int64_t)blockIdx.x * TT;
  const int kk0 = blockIdx.y * TT;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < TT; r += 4) {
    const int64_t p = p0 + r;
    const int kk = kk0 + tx;
    float v = 0.0f;
    if (p < npix && kk < kb) v = img_k[p * kstride + kk];
    tile[tx][r] = v;
  }
  __syncthreads();
  for (int r = ty; r < TT; r += 4) {
    const int kk = kk0 + r;
    const int64_t p = p0 + tx;
    if (kk < kb && p < npix) coeffs[(k0 + kk) * npix + p] = tile[r][tx];
  }
}

// ------------------------------------------------------------------------------------------------ Radon
struct AngleGeo {
  float X, Y, T0, wscale, hw, top, pad0, pad1;
};

// XCD-aware work split. A launch of n_blk channel blocks (16 float4 quads = 64 orientations each) x per_blk
// workgroups per block is a 1-D grid whose workgroup i runs on XCD i % 8 (the dispatcher's round robin). Block c is
// mapped to XCD c % 8 and its per_blk workgroups run there back to back, so the block's working set (its image
// columns, 2.5 MB at 99 x 99 pixels, or its sinogram slab) stays in that XCD's 4 MB L2 while every angle / pixel
// of the block is processed, instead of being re-read from HBM once per angle.
constexpr int N_XCD = 8;
constexpr int QB = 16;   // float4 quads per channel block

__device__ __forceinline__ bool xcd_split(int per_blk, int n_blk, int& blk, int& local) {
  const int i = blockIdx.x;
  const int xcd = i % N_XCD;
  const int j = i / N_XCD;
  blk = (j / per_blk) * N_XCD + xcd;
  local = j % per_blk;
  return blk < n_blk;
}

inline unsigned xcd_grid(int per_blk, int n_blk) {
  const int rounds = (n_blk + N_XCD - 1) / N_XCD;
  return (unsigned)rounds * N_XCD * per_blk;
}

// Forward: a thread per (channel quad, NB consecutive bins) of one angle; a workgroup covers 16 quads x 16 bin groups
// (16 NB bins) of one angle. For every image row along the axis with the larger |coefficient| only the pixels whose
// footprint reaches the thread's bins are visited.
template <int NB>
__global__ void __launch_bounds__(256) radon_forward_kernel(const float4* __restrict__ img, float4* __restrict__ sino,
                                                            const AngleGeo* __restrict__ geo, int nx, int ny, int ns,
                                                            int na, int kstride4, int k4_off, int k4_n, int skstride4,
                                                            float scale) {
  const int sblocks = (ns + 16 * NB - 1) / (16 * NB);
  const int n_blk = (k4_n + QB - 1) / QB;
  int blk, local;
  if (!xcd_split(na * sblocks, n_blk, blk, local)) return;
  const int a = local / sblocks;
  const int kq = blk * QB + (threadIdx.x & (QB - 1));
  const int s0 = ((local % sblocks) * 16 + (threadIdx.x >> 4)) * NB;
  if (kq >= k4_n || s0 >= ns) return;
  const int k = kq + k4_off;
  const AngleGeo g = geo[a];
  const bool rows_y = fabsf(g.X) >= fabsf(g.Y);
  const float A = rows_y ? g.X : g.Y;
  const float B = rows_y ? g.Y : g.X;
  const int nu = rows_y ? nx : ny;
  const int nv = rows_y ? ny : nx;
  const int64_t du = rows_y ? (int64_t)kstride4 : (int64_t)nx * kstride4;
  const int64_t dv = rows_y ? (int64_t)nx * kstride4 : (int64_t)kstride4;
  const float inv = 1.0f / A;
  const float hw = g.hw, top = g.top;
  float4 acc[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) acc[j] = make_float4(0.f, 0.f, 0.f, 0.f);
  const float4* col = img + k;
  for (int v = 0; v < nv; ++v) {
    const float c = B * (float)v + g.T0;
    const float e1 = ((float)s0 - hw - c) * inv;
    const float e2 = ((float)(s0 + NB - 1) + hw - c) * inv;
    const int u0 = max((int)floorf(fminf(e1, e2)), 0);
    const int u1 = min((int)ceilf(fmaxf(e1, e2)), nu - 1);
    const float4* row = col + (int64_t)v * dv;
    for (int u = u0; u <= u1; ++u) {
      const float4 val = row[(int64_t)u * du];
      const float t = A * (float)u + c - (float)s0;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const float w = fminf(fmaxf(hw - fabsf(t - (float)j), 0.0f), top);
        acc[j].x += w * val.x;
        acc[j].y += w * val.y;
        acc[j].z += w * val.z;
        acc[j].w += w * val.w;
      }
    }
  }
  const float f = scale * g.wscale;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (s0 + j < ns) {
      float4 o;
      o.x = f * acc[j].x;
      o.y = f * acc[j].y;
      o.z = f * acc[j].z;
      o.w = f * acc[j].w;
      sino[((int64_t)a * ns + s0 + j) * skstride4 + kq] = o;
    }
  }
}

// Backward: a thread per (channel quad, pixel); a workgroup covers 16 quads x 16 consecutive pixels, which share
// most of their sinogram bins at every angle. MODE 0 writes the image, MODE 1 applies the SPDHG primal update to
// (x, z) with the step T.
template <int NTAP, int MODE>
__global__ void __launch_bounds__(256) radon_backward_kernel(const float4* __restrict__ sino, float4* __restrict__ out,
                                                             const AngleGeo* __restrict__ geo, int nx, int ny, int ns,
                                                             int na, int kstride4, int k4_off, int k4_n, int skstride4,
                                                             float4* __restrict__ x, float4* __restrict__ z,
                                                             const float4* __restrict__ T, float inv_p, float lam) {
  const int64_t npix = (int64_t)nx * ny;
  const int pblocks = (int)((npix + 15) / 16);
  const int n_blk = (k4_n + QB - 1) / QB;
  int blk, local;
  if (!xcd_split(pblocks, n_blk, blk, local)) return;
  const int kq = blk * QB + (threadIdx.x & (QB - 1));
  const int64_t pix = (int64_t)local * 16 + (threadIdx.x >> 4);
  if (kq >= k4_n || pix >= npix) return;
  const int k = kq + k4_off;
  const float px = (float)(pix % nx), py = (float)(pix / nx);
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int a = 0; a < na; ++a) {
    const AngleGeo g = geo[a];
    const float t = g.X * px + g.Y * py + g.T0;
    const int sf = (int)floorf(t - g.hw) + 1;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4* srow = sino + (int64_t)a * ns * skstride4 + kq;
#pragma unroll
    for (int j = 0; j < NTAP; ++j) {
      const int s = sf + j;
      if (s >= 0 && s < ns) {
        const float w = fminf(fmaxf(g.hw - fabsf(t - (float)s), 0.0f), g.top);
        const float4 sv = srow[(int64_t)s * skstride4];
        v.x += w * sv.x;
        v.y += w * sv.y;
        v.z += w * sv.z;
        v.w += w * sv.w;
      }
    }
    acc.x += g.wscale * v.x;
    acc.y += g.wscale * v.y;
    acc.z += g.wscale * v.z;
    acc.w += g.wscale * v.w;
  }
  const int64_t o = pix * kstride4 + k;
  if (MODE == 0) {
    st_stream(out + o, acc);
  } else {
    float4 zz = ld_stream(z + o), xx = ld_stream(x + o);
    const float4 tt = ld_stream(T + o);
    float4 zn, xn;
#define DFM_UPD(c)                                                                  \
  zn.c = zz.c + acc.c;                                                              \
  xn.c = tt.c > 0.0f ? fmaxf(0.0f, xx.c - tt.c * ((zn.c + inv_p * acc.c) + lam)) : 0.0f;
    DFM_UPD(x) DFM_UPD(y) DFM_UPD(z) DFM_UPD(w)
#undef DFM_UPD
    st_stream(z + o, zn);
    st_stream(x + o, xn);
  }
}

// ------------------------------------------------------------------------------------------------ Radon, tables
// The projector's geometry depends on the angles and the pixel grid only, not on the orientations, so it is computed
// once per angle set (on the host, in float64) and the kernels spend their instructions on the orientations:
//   forward table:  per (angle, bin) the pixels whose footprint reaches the bin and their weights (a CSR of the
//                   projection matrix, the scale and 1/A folded in);
//   pixel table:    per (angle, pixel) its two bins and the backprojection weights (angle weight and 1/A folded in);
//   tile bases:     per (angle, 4 x 4 pixel tile) the first of the TB bins the tile reaches.

// Forward from the table: a thread per (channel quad, bin); a workgroup covers QBT quads x 256 / QBT consecutive bins
// of one angle; the channel blocks are split over the XCDs as above. The bins of a workgroup are consecutive
// rows of the CSR, i.e. one contiguous entry range, loaded once with coalesced loads (FT_CHUNK entries per pass) and
// then read from LDS by all channel quads of a bin (a broadcast): the vector memory path carries the image gathers
// only. UNR entries are fetched before they are used, so a wave keeps UNR independent image loads in flight.
constexpr int FT_CHUNK = 2048;

template <int UNR, int QBT>
__global__ void __launch_bounds__(256) radon_forward_table_lds_kernel(const float4* __restrict__ img,
                                                                      float4* __restrict__ sino,
                                                                      const int* __restrict__ row_ptr,
                                                                      const int* __restrict__ ent_pix,
                                                                      const float* __restrict__ ent_w, int ns, int na,
                                                                      int kstride4, int k4_off, int k4_n,
                                                                      int skstride4) {
  __shared__ int lp[FT_CHUNK];
  __shared__ float lw[FT_CHUNK];
  constexpr int BINS = 256 / QBT;
  const int sblocks = (ns + BINS - 1) / BINS;
  const int n_blk = (k4_n + QBT - 1) / QBT;
  int blk, local;
  if (!xcd_split(na * sblocks, n_blk, blk, local)) return;       // uniform over the workgroup
  const int a = local / sblocks;
  const int s0 = (local % sblocks) * BINS;
  const int s = s0 + threadIdx.x / QBT;
  const int kq = blk * QBT + (threadIdx.x & (QBT - 1));
  const bool live = kq < k4_n && s < ns;
  const int g0 = row_ptr[a * ns + s0];
  const int g1 = row_ptr[a * ns + min(s0 + BINS, ns)];
  const int r = a * ns + min(s, ns - 1);
  const int e0 = live ? row_ptr[r] : 0;
  const int e1 = live ? row_ptr[r + 1] : 0;
  const float4* col = img + (live ? kq : 0) + k4_off;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int c0 = g0; c0 < g1; c0 += FT_CHUNK) {
    const int c1 = min(c0 + FT_CHUNK, g1);
    __syncthreads();
    for (int i = threadIdx.x; i < c1 - c0; i += 256) {
      lp[i] = ent_pix[c0 + i];
      lw[i] = ent_w[c0 + i];
    }
    __syncthreads();
    const int en = min(e1, c1);
    int e = max(e0, c0);
    for (; e + UNR <= en; e += UNR) {
      int p[UNR];
      float w[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        p[u] = lp[e + u - c0];
        w[u] = lw[e + u - c0];
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
    for (; e < en; ++e) {
      const float w = lw[e - c0];
      const float4 v = col[(int64_t)lp[e - c0] * kstride4];
      acc.x += w * v.x;
      acc.y += w * v.y;
      acc.z += w * v.z;
      acc.w += w * v.w;
    }
  }
  if (live) st_stream(sino + (int64_t)r * skstride4 + kq, acc);
}

// Backward from the tables, one 4 x 4 pixel tile x 16 channel quads per workgroup: the TB bins of every angle the
// tile reaches are staged in LDS (TB x 16 quads x 16 B per angle, angles in chunks of ACH), then every thread (pixel,
// quad) reads its two taps from LDS. MODE 0 writes the image, MODE 1 applies the SPDHG primal update.
constexpr int TB = 8;     // bins per (angle, tile)

template <int MODE, int ACH>
__global__ void __launch_bounds__(256) radon_backward_tile_kernel(const float4* __restrict__ sino,
                                                                  float4* __restrict__ out,
                                                                  const int4* __restrict__ ptab,
                                                                  const int* __restrict__ tbase, int nx, int ny,
                                                                  int ns, int na, int kstride4, int k4_off, int k4_n,
                                                                  int skstride4, float4* __restrict__ x,
                                                                  float4* __restrict__ z, const float4* __restrict__ T,
                                                                  float inv_p, float lam) {
  __shared__ float4 lds[ACH * TB * QB];   // ACH angles x TB bins x 16 quads
  __shared__ int4 lpt[ACH * 16];          // the pixel table of the tile's 16 pixels for the ACH angles
  __shared__ int lbase[ACH];
  const int tnx = (nx + 3) / 4, tny = (ny + 3) / 4;
  const int ntiles = tnx * tny;
  const int n_blk = (k4_n + QB - 1) / QB;
  int blk, tile;
  if (!xcd_split(ntiles, n_blk, blk, tile)) return;
  const int q = threadIdx.x & (QB - 1);
  const int li = threadIdx.x >> 4;                       // pixel of the tile, row-major 4 x 4
  const int px = (tile % tnx) * 4 + (li & 3), py = (tile / tnx) * 4 + (li >> 2);
  const bool inside = px < nx && py < ny;
  const int kq = blk * QB + q;
  const bool live = inside && kq < k4_n;
  const int64_t pix = (int64_t)py * nx + px;
  const int64_t npix = (int64_t)nx * ny;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  // the update's operands are fetched first, so their HBM latency overlaps the sinogram work
  const int64_t o = pix * kstride4 + kq + k4_off;
  float4 zz = acc, xx = acc, tt = acc;
  if (MODE == 1 && live) {
    zz = ld_stream(z + o);
    xx = ld_stream(x + o);
    tt = ld_stream(T + o);
  }
  for (int a0 = 0; a0 < na; a0 += ACH) {
    const int nan = min(ACH, na - a0);
    __syncthreads();
    // the chunk's tile bases and pixel-table rows, once per workgroup (not once per channel quad)
    if (threadIdx.x < nan) lbase[threadIdx.x] = tbase[(int64_t)(a0 + threadIdx.x) * ntiles + tile];
    for (int i = threadIdx.x; i < nan * 16; i += 256) {
      const int al = i >> 4, l = i & 15;
      const int ppx = (tile % tnx) * 4 + (l & 3), ppy = (tile / tnx) * 4 + (l >> 2);
      lpt[i] = (ppx < nx && ppy < ny) ? ptab[(int64_t)(a0 + al) * npix + (int64_t)ppy * nx + ppx]
                                      : make_int4(0, 0, 0, 0);
    }
    __syncthreads();
    // stage: element i = (angle al, bin b, quad qq); all of the chunk's loads issued before the first LDS store
    constexpr int PER = ACH * TB * QB / 256;
    static_assert(ACH * TB * QB % 256 == 0, "whole staging rounds");
    float4 v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = threadIdx.x + 256 * u;
      const int qq = i & (QB - 1);
      const int b = (i >> 4) % TB;
      const int al = i / (TB * QB);
      v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (al < nan) {
        const int s = lbase[al] + b;
        const int kk = blk * QB + qq;
        if (s >= 0 && s < ns && kk < k4_n) v[u] = sino[((int64_t)(a0 + al) * ns + s) * skstride4 + kk];
      }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) lds[threadIdx.x + 256 * u] = v[u];
    __syncthreads();
    if (live) {
#pragma unroll 4
      for (int al = 0; al < nan; ++al) {
        const int4 e = lpt[al * 16 + li];
        const int base = lbase[al];
        const float w0 = __int_as_float(e.z), w1 = __int_as_float(e.w);
        const float4 s0 = lds[(al * TB + (e.x - base)) * QB + q];
        const float4 s1 = lds[(al * TB + (e.y - base)) * QB + q];
        acc.x = fmaf(w1, s1.x, fmaf(w0, s0.x, acc.x));
        acc.y = fmaf(w1, s1.y, fmaf(w0, s0.y, acc.y));
        acc.z = fmaf(w1, s1.z, fmaf(w0, s0.z, acc.z));
        acc.w = fmaf(w1, s1.w, fmaf(w0, s0.w, acc.w));
      }
    }
  }
  if (!live) return;
  if (MODE == 0) {
    st_stream(out + o, acc);
  } else {
    float4 zn, xn;
#define DFM_UPD(c)                                                                  \
  zn.c = zz.c + acc.c;                                                              \
  xn.c = tt.c > 0.0f ? fmaxf(0.0f, xx.c - tt.c * ((zn.c + inv_p * acc.c) + lam)) : 0.0f;
    DFM_UPD(x) DFM_UPD(y) DFM_UPD(z) DFM_UPD(w)
#undef DFM_UPD
    st_stream(z + o, zn);
    st_stream(x + o, xn);
  }
}

// The same backprojection with a workgroup covering one TS x TS tile for G consecutive channel blocks of QW quads: the
// tile's bases and pixel-table rows (channel-independent) are staged once per TCH angles and reused by every block, so
// the per-block work is the sinogram staging (all of a pass's loads in flight) and the pixel sums. Blocks
// g*G .. g*G+G-1 run on XCD g % 8, so their sinogram slabs stay in its L2. The blocks of a workgroup run one after
// another (QW = 16 quads: 64 channels per block).
constexpr int TCH = 40;   // angles whose table rows are staged at once (a subset of 36 at C2 in one go)

template <int MODE, int ACH, int G, int TS, int QW>
__global__ void __launch_bounds__(TS * TS * QW, TS == 8 ? 8 : 1) radon_backward_group_kernel(
    const float4* __restrict__ sino, float4* __restrict__ out, const double* __restrict__ bgeo,
    const int* __restrict__ tbase, int nx, int ny, int ns, int na, int kstride4, int k4_off, int k4_n, int skstride4,
    float4* __restrict__ x, float4* __restrict__ z, const float4* __restrict__ T, float inv_p, float lam) {
  constexpr int NPX = TS * TS, NTH = NPX * QW;    // QW float4 quads of every pixel per channel block
  constexpr int TBB = TS == 4 ? TB : 16;   // bins a tile reaches at one angle (host-checked)
  __shared__ float4 lds[ACH * TBB * QW];  // ACH angles x TBB bins x QW quads
  // pixel-table rows of the tile's pixels for TCH angles, as staged: the byte offsets of the two bins' rows within
  // the angle's LDS slab ((bin - first bin of the tile) x QW quads x 16 B < 2^16, packed in one int) and the two
  // weights (12 bytes per entry instead of 16: 8 angles' entries in flight take 24 VGPRs, not 32)
  __shared__ int lofs[TCH * NPX];
  __shared__ float2 lw[TCH * NPX];
  __shared__ int lbase[TCH];
  static_assert(ACH * TBB * QW % NTH == 0 && TCH % ACH == 0, "whole staging rounds");
  const int tnx = (nx + TS - 1) / TS, tny = (ny + TS - 1) / TS;
  const int ntiles = tnx * tny;
  const int n_cb = (k4_n + QW - 1) / QW;
  const int n_grp = (n_cb + G - 1) / G;
  int grp, tile;
  if (!xcd_split(ntiles, n_grp, grp, tile)) return;
  const int q = threadIdx.x & (QW - 1);
  const int li = threadIdx.x / QW;                       // pixel of the tile, row-major TS x TS
  const int px = (tile % tnx) * TS + li % TS, py = (tile / tnx) * TS + li / TS;
  const bool inside = px < nx && py < ny;
  const int pix = py * nx + px;                          // < 2^31 pixels (a 32-bit value: one VGPR less live)
  const bool tables_once = na <= TCH;
  // stage the table rows of angles t0 .. t0 + tn - 1, computed from the angles' float64 geometry (bgeo: X, Y, T0,
  // hw, top, w / A per angle) by the rule of the host's pixel table (radon_tables: taps floor(t - hw) + 1 and the
  // next bin, weights (w / A) clamp(hw - |t - s|, 0, top) where the bin is on the detector, bins clamped to it), so
  // no table is read from memory; the bins become byte offsets from the tile's first bin, clamped into the staged
  // range (a bin outside it has weight 0)
  auto stage_tables = [&](int t0, int tn) {
    if (threadIdx.x < tn) lbase[threadIdx.x] = tbase[(int64_t)(t0 + threadIdx.x) * ntiles + tile];
    // (the loop's first index passes an empty asm: its geometry address is formed here, not hoisted out of the
    // channel-block and angle loops to stay live, and spilled, across the whole kernel)
    int i0 = threadIdx.x;
    asm volatile("" : "+v"(i0));
    for (int i = i0; i < tn * NPX; i += NTH) {
      const int al = i / NPX, l = i % NPX;
      const int ppx = (tile % tnx) * TS + l % TS, ppy = (tile / tnx) * TS + l / TS;
      int ofs = 0;
      float2 ww = make_float2(0.f, 0.f);
      if (ppx < nx && ppy < ny) {
        const double* g = bgeo + (int64_t)(t0 + al) * 6;
        const double t = g[0] * (double)ppx + g[1] * (double)ppy + g[2];
        const double hw = g[3], tp = g[4], sc = g[5];
        const int s0 = (int)floor(t - hw) + 1, s1 = s0 + 1;
        const double k0 = fmin(fmax(hw - fabs(t - (double)s0), 0.0), tp);
        const double k1 = fmin(fmax(hw - fabs(t - (double)s1), 0.0), tp);
        ww.x = (s0 >= 0 && s0 < ns && k0 > 0.0) ? (float)(sc * k0) : 0.0f;
        ww.y = (s1 >= 0 && s1 < ns && k1 > 0.0) ? (float)(sc * k1) : 0.0f;
        const int b0 = tbase[(int64_t)(t0 + al) * ntiles + tile];
        const int o0 = min(max(min(max(s0, 0), ns - 1) - b0, 0), TBB - 1);
        const int o1 = min(max(min(max(s1, 0), ns - 1) - b0, 0), TBB - 1);
        ofs = (o0 * QW * (int)sizeof(float4)) | ((o1 * QW * (int)sizeof(float4)) << 16);
      }
      lofs[i] = ofs;
      lw[i] = ww;
    }
  };
  if (tables_once) stage_tables(0, na);
  for (int gi = 0; gi < G; ++gi) {
    const int cb = grp * G + gi;
    if (cb >= n_cb) break;                                // uniform over the workgroup
    const int kq = cb * QW + q;
    const bool live = inside && kq < k4_n;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int t0 = 0; t0 < na; t0 += TCH) {
      const int tn = min(TCH, na - t0);
      if (!tables_once) {
        __syncthreads();
        stage_tables(t0, tn);
      }
      // PIPE: software-pipelined passes, the sinogram loads of pass a0 + ACH in flight while pass a0 is summed (for
      // the 512 / 1024-thread forms; in the 256-thread ones the 8 more VGPRs cost more in occupancy than they gain)
      constexpr bool PIPE = NTH >= 512;
      constexpr int PER = ACH * TBB * QW / NTH;
      float4 v[PER];
      auto load_pass = [&](int a0) {
        const int nan = min(ACH, tn - a0);
#pragma unroll
        for (int u = 0; u < PER; ++u) {
          const int i = threadIdx.x + NTH * u;
          const int qq = i & (QW - 1);
          const int b = (i / QW) % TBB;
          const int al = i / (TBB * QW);
          v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
          if (al < nan) {
            const int s = lbase[a0 + al] + b;
            const int kk = cb * QW + qq;
            if (s >= 0 && s < ns && kk < k4_n) v[u] = sino[((int64_t)(t0 + a0 + al) * ns + s) * skstride4 + kk];
          }
        }
      };
      if (PIPE) {
        __syncthreads();                                   // the chunk's bases staged
        load_pass(0);
      }
      for (int a0 = 0; a0 < tn; a0 += ACH) {
        const int nan = min(ACH, tn - a0);
        __syncthreads();                                   // every wave done with the previous pass's LDS
        if (!PIPE) load_pass(a0);
#pragma unroll
        for (int u = 0; u < PER; ++u) lds[threadIdx.x + NTH * u] = v[u];
        __syncthreads();
        if (PIPE && a0 + ACH < tn) load_pass(a0 + ACH);
        if (live) {
          // per angle: one table read, two LDS addresses (byte offset + this lane's quad), 8 FMAs; a full pass
          // is unrolled without branches so that its table reads are all issued first
          const char* lq = reinterpret_cast<const char*>(lds + q);
          auto angle = [&](int al) {
            const unsigned ofs = (unsigned)lofs[(a0 + al) * NPX + li];
            const float2 ww = lw[(a0 + al) * NPX + li];
            const float w0 = ww.x, w1 = ww.y;
            const float4 s0 = reinterpret_cast<const float4*>(lq + (ofs & 0xFFFFu))[al * TBB * QW];
            const float4 s1 = reinterpret_cast<const float4*>(lq + (ofs >> 16))[al * TBB * QW];
            acc.x = fmaf(w1, s1.x, fmaf(w0, s0.x, acc.x));
            acc.y = fmaf(w1, s1.y, fmaf(w0, s0.y, acc.y));
            acc.z = fmaf(w1, s1.z, fmaf(w0, s0.z, acc.z));
            acc.w = fmaf(w1, s1.w, fmaf(w0, s0.w, acc.w));
          };
          if (nan == ACH) {
#pragma unroll
            for (int al = 0; al < ACH; ++al) angle(al);
          } else {
            for (int al = 0; al < nan; ++al) angle(al);
          }
        }
      }
    }
    if (live) {
      // the element offset is formed here from the 32-bit pixel index: the empty asm keeps the compiler from hoisting
      // the 64-bit product out of the loops, where it would stay live across the accumulation (2 VGPRs of the 64)
      int pe = pix;
      asm volatile("" : "+v"(pe));
      const int64_t o = (int64_t)pe * kstride4 + kq + k4_off;
      // x, z, T (MODE 2: the EM sensitivity; 3: the PDHG step) are read after the accumulation: held across it they
      // would cost 12 VGPRs of the 64 (8 waves per SIMD), and the other waves hide the latency
      float4 zz = make_float4(0.f, 0.f, 0.f, 0.f), xx = zz, tt = zz;
      if (MODE >= 1) {
        if (MODE == 1) zz = ld_stream(z + o);
        xx = ld_stream(x + o);
        tt = ld_stream(T + o);
      }
      if (MODE == 0) {
        st_stream(out + o, acc);
      } else if (MODE == 1) {
        float4 zn, xn;
#define DFM_UPD(c)                                                                  \
  zn.c = zz.c + acc.c;                                                              \
  xn.c = tt.c > 0.0f ? fmaxf(0.0f, xx.c - tt.c * ((zn.c + inv_p * acc.c) + lam)) : 0.0f;
        DFM_UPD(x) DFM_UPD(y) DFM_UPD(z) DFM_UPD(w)
#undef DFM_UPD
        st_stream(z + o, zn);
        st_stream(x + o, xn);
      } else if (MODE == 2) {
        // the EM update (osem_poisson.py:69-73): x = x A^T r / (sens + lam) where sens > 0
        float4 xn;
        xn.x = tt.x > 0.0f ? xx.x * acc.x / (tt.x + lam) : xx.x;
        xn.y = tt.y > 0.0f ? xx.y * acc.y / (tt.y + lam) : xx.y;
        xn.z = tt.z > 0.0f ? xx.z * acc.z / (tt.z + lam) : xx.z;
        xn.w = tt.w > 0.0f ? xx.w * acc.w / (tt.w + lam) : xx.w;
        st_stream(x + o, xn);
      } else {
        // the PDHG primal step and extrapolation (pdhg_poisson.py:69-80 primal_update): x' = max(0, x - T (A^T y +
        // lam)) where T > 0, else 0; x_bar = 2 x' - x (into z)
        float4 xn, xb;
#define DFM_PD(c)                                                        \
  xn.c = tt.c > 0.0f ? fmaxf(0.0f, xx.c - tt.c * (acc.c + lam)) : 0.0f;  \
  xb.c = 2.0f * xn.c - xx.c;
        DFM_PD(x) DFM_PD(y) DFM_PD(z) DFM_PD(w)
#undef DFM_PD
        st_stream(z + o, xb);
        st_stream(x + o, xn);
      }
    }
  }
}

// ------------------------------------------------------------------------------------------------ dense PF
// One thread per entry (r, k, c, p) of the batch; the sum over quadrature nodes and the ring's pole axes. Directions
// and poles are float64 and so is the arithmetic: near a pole 1 - |n.v| is a difference of nearly equal numbers
// and the kernel amplifies its error by 1/sigma^2 (3300 at sigma = 1 deg), so float32 inputs alone would cost ~1e-4
// of relative accuracy. Setting up the PF is not on the iteration path, and gfx950 runs float64 at full rate.
__global__ void __launch_bounds__(256) pf_eval_kernel(const double* __restrict__ coords, const double* __restrict__ poles,
                                                      const int* __restrict__ axis_start,
                                                      const float* __restrict__ axis_count,
                                                      const double* __restrict__ inv_sigma2,
                                                      const double* __restrict__ normf, const float* __restrict__ qw,
                                                      float* __restrict__ out, int R, int kmax, int C, int P, int A,
                                                      int S, int Se, int k0, int kb, float pf_cut, int quad_weighted) {
  const int64_t total = (int64_t)R * kmax * C * P;
  for (int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)(gid % P);
    int64_t t = gid / P;
    const int c = (int)(t % C);
    t /= C;
    const int k = (int)(t % kmax);
    const int r = (int)(t / kmax);
    out[gid] = k >= kb ? 0.0f
                       : pf_entry(coords, poles, axis_start, axis_count, inv_sigma2, normf, qw, r, k0 + k, c, p, C, P,
                                  A, S, Se, (double)pf_cut, quad_weighted);
  }
}

// Forward-CSR rows (r, j): one block per angle, threads over j, loop over k (coalesced reads of the (r, k, :) rows).
__global__ void __launch_bounds__(256) pf_count_fwd_kernel(const float* __restrict__ pf, int* __restrict__ counts,
                                                           int kmax, int CP, int kb) {
  const int r = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= CP) return;
  const float* base = pf + (int64_t)r * kmax * CP + j;
  int n = 0;
  for (int k = 0; k < kb; ++k) n += base[(int64_t)k * CP] != 0.0f;
  counts[(int64_t)r * CP + j] = n;
}

__global__ void __launch_bounds__(256) pf_fill_fwd_kernel(const float* __restrict__ pf, const int* __restrict__ row_ptr,
                                                          uint16_t* __restrict__ col_k, float* __restrict__ val,
                                                          int kmax, int CP, int kb) {
  const int r = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= CP) return;
  const float* base = pf + (int64_t)r * kmax * CP + j;
  int o = row_ptr[(int64_t)r * CP + j];
  for (int k = 0; k < kb; ++k) {
    const float v = base[(int64_t)k * CP];
    if (v != 0.0f) {
      col_k[o] = (uint16_t)k;
      val[o] = v;
      ++o;
    }
  }
}

// Adjoint-CSR rows (r, k): one wave per row; the lanes scan the row's CP entries in chunks of 64 and compact them
// with a ballot prefix count, keeping j ascending.
__global__ void __launch_bounds__(256) pf_count_adj_kernel(const float* __restrict__ pf, int* __restrict__ counts,
                                                           int R, int kmax, int CP, int kb) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= (int64_t)R * kb) return;
  const int r = (int)(row / kb), k = (int)(row % kb);
  const float* base = pf + ((int64_t)r * kmax + k) * CP;
  int n = 0;
  for (int j0 = 0; j0 < CP; j0 += 64) {
    const int j = j0 + lane;
    const bool nz = j < CP && base[j] != 0.0f;
    n += __popcll(__ballot(nz));
  }
  if (lane == 0) counts[row] = n;
}

__global__ void __launch_bounds__(256) pf_fill_adj_kernel(const float* __restrict__ pf, const int* __restrict__ row_ptr,
                                                          uint16_t* __restrict__ col_j, float* __restrict__ val,
                                                          int R, int kmax, int CP, int kb) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= (int64_t)R * kb) return;
  const int r = (int)(row / kb), k = (int)(row % kb);
  const float* base = pf + ((int64_t)r * kmax + k) * CP;
  int o = row_ptr[row];
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int j0 = 0; j0 < CP; j0 += 64) {
    const int j = j0 + lane;
    const float v = j < CP ? base[j] : 0.0f;
    const unsigned long long m = __ballot(v != 0.0f);
    if (v != 0.0f) {
      const int pos = o + __popcll(m & below);
      col_j[pos] = (uint16_t)j;
      val[pos] = v;
    }
    o += __popcll(m);
  }
}

// ------------------------------------------------------------------------------------------------ PF products
constexpr int SPMM_TY = 16;          // detector positions per block (registers per thread)
constexpr int SPMM_LDS_FLOATS = 12288;  // 48 KiB of staged sinogram / data rows

// Forward: a thread per CSR row (angle, j); the sinogram rows of SPMM_TY detector positions staged in LDS chunk by
// chunk of orientations (the CSR rows list k ascending, so each row advances through the chunks in order).
// With ksplit > 1 the orientations are split into ksplit ranges of kpart (grid z = angle * ksplit + split), each
// row's first entry of its range found by bisection (columns ascend), and the partial sums written to part
// (ksplit, n_angles, my, cp) for a fixed-order reduction: ksplit times the waves of one launch.
__global__ void __launch_bounds__(256) spmm_forward_kernel(const float* __restrict__ sino, const int* __restrict__ row_ptr,
                                                           const uint16_t* __restrict__ col_k,
                                                           const float* __restrict__ val, float* __restrict__ data,
                                                           const int* __restrict__ omegas, int my, int cp, int kstride,
                                                           int kb, int kc, int compact, int accumulate, int ksplit,
                                                           int kpart, float* __restrict__ part, int n_angles) {
  extern __shared__ float lsino[];
  const int lid = threadIdx.x;
  const int j = blockIdx.x * blockDim.x + lid;
  const int y0 = blockIdx.y * SPMM_TY;
  const int rs = blockIdx.z / ksplit;
  const int ks = blockIdx.z % ksplit;
  const int r = omegas ? omegas[rs] : rs;
  const int ny = min(SPMM_TY, my - y0);
  const int rr = compact ? rs : r;
  const int klo = ks * kpart, khi = min(kb, klo + kpart);
  int i = 0, e = 0;
  if (j < cp) {
    i = row_ptr[(int64_t)rr * cp + j];
    e = row_ptr[(int64_t)rr * cp + j + 1];
    if (klo > 0) {                                   // first entry with column >= klo
      int lo = i, hi = e;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int)col_k[mid] < klo) lo = mid + 1;
        else hi = mid;
      }
      i = lo;
    }
  }
  float acc[SPMM_TY];
#pragma unroll
  for (int t = 0; t < SPMM_TY; ++t) acc[t] = 0.0f;
  const float* sbase = sino + ((int64_t)rs * my + y0) * kstride;
  for (int k0 = klo; k0 < khi; k0 += kc) {
    const int kn = min(kc, khi - k0);
    __syncthreads();
    for (int t = 0; t < ny; ++t)
      for (int kk = lid; kk < kn; kk += blockDim.x) lsino[t * kc + kk] = sbase[(int64_t)t * kstride + k0 + kk];
    __syncthreads();
    const int k1 = k0 + kn;
    for (; i < e; ++i) {
      const int k = col_k[i];
      if (k >= k1) break;
      const float v = val[i];
      const float* ls = lsino + (k - k0);
#pragma unroll
      for (int t = 0; t < SPMM_TY; ++t) acc[t] += v * ls[t * kc];
    }
  }
  if (j < cp && ksplit > 1) {
    float* dst = part + (((int64_t)ks * n_angles + rs) * my + y0) * cp + j;
    for (int t = 0; t < ny; ++t) dst[(int64_t)t * cp] = acc[t];
  } else if (j < cp) {
    float* dst = data + ((int64_t)r * my + y0) * cp + j;
    if (accumulate) {
      for (int t = 0; t < ny; ++t) dst[(int64_t)t * cp] += acc[t];
    } else {
      for (int t = 0; t < ny; ++t) dst[(int64_t)t * cp] = acc[t];
    }
  }
}

// Adjoint: a thread per CSR row (angle, k); data rows staged in LDS by chunks of segments j.
__global__ void __launch_bounds__(256) spmm_adjoint_kernel(const float* __restrict__ data, const int* __restrict__ row_ptr,
                                                           const uint16_t* __restrict__ col_j,
                                                           const float* __restrict__ val, float* __restrict__ sino,
                                                           const int* __restrict__ omegas, int my, int cp, int kstride,
                                                           int kb, int jc, float alpha, int compact) {
  extern __shared__ float ldata[];
  const int lid = threadIdx.x;
  const int k = blockIdx.x * blockDim.x + lid;
  const int y0 = blockIdx.y * SPMM_TY;
  const int rs = blockIdx.z;
  const int r = omegas ? omegas[rs] : rs;
  const int ny = min(SPMM_TY, my - y0);
  const int rr = compact ? rs : r;
  int i = 0, e = 0;
  if (k < kb) {
    i = row_ptr[(int64_t)rr * kb + k];
    e = row_ptr[(int64_t)rr * kb + k + 1];
  }
  float acc[SPMM_TY];
#pragma unroll
  for (int t = 0; t < SPMM_TY; ++t) acc[t] = 0.0f;
  const float* dbase = data + ((int64_t)r * my + y0) * cp;
  for (int j0 = 0; j0 < cp; j0 += jc) {
    const int jn = min(jc, cp - j0);
    __syncthreads();
    for (int t = 0; t < ny; ++t)
      for (int jj = lid; jj < jn; jj += blockDim.x) ldata[t * jc + jj] = dbase[(int64_t)t * cp + j0 + jj];
    __syncthreads();
    const int j1 = j0 + jn;
    for (; i < e; ++i) {
      const int jj = col_j[i];
      if (jj >= j1) break;
      const float v = val[i];
      const float* ld = ldata + (jj - j0);
#pragma unroll
      for (int t = 0; t < SPMM_TY; ++t) acc[t] += v * ld[t * jc];
    }
  }
  if (k < kstride) {
    float* dst = sino + ((int64_t)rs * my + y0) * kstride + k;
    for (int t = 0; t < ny; ++t) dst[(int64_t)t * kstride] = (k < kb) ? alpha * acc[t] : 0.0f;
  }
}

__global__ void zero_rows_kernel(float* __restrict__ data, const int* __restrict__ omegas, int64_t row_len,
                                 int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = i / row_len;
    data[(int64_t)omegas[s] * row_len + (i - s * row_len)] = 0.0f;
  }
}

static int check_radon(const char* who, int nx, int ny, int ns, int na, int kstride, int k_off, int k_n,
                       int skstride, int ntap) {
  DFM_CHECK_ARG(nx > 0 && ny > 0 && ns > 0 && na >= 0, "%s: bad image/detector sizes", who);
  DFM_CHECK_ARG(kstride > 0 && kstride % 4 == 0 && k_off >= 0 && k_off % 4 == 0 && k_n >= 0 && k_n % 4 == 0 &&
                    k_off + k_n <= kstride,
                "%s: channel range must be multiples of 4 inside kstride (k_off=%d k_n=%d kstride=%d)", who, k_off,
                k_n, kstride);
  DFM_CHECK_ARG(ntap >= 2 && ntap <= 4, "%s: ntap must be 2..4 (got %d)", who, ntap);
  DFM_CHECK_ARG(skstride % 4 == 0 && skstride >= k_n, "%s: sinogram stride %d must be a multiple of 4 >= k_n", who,
                skstride);
  DFM_CHECK_ARG((int64_t)nx * ny * (kstride / 4) < (int64_t)1 << 40, "%s: array too large", who);
  return DFM_OK;
}

template <int MODE>
static int launch_backward(const char* who, const float* sino, float* img_k, const float* geo, int nx, int ny, int ns,
                           int na, int kstride, int k_off, int k_n, int skstride, int ntap, float* x, float* z,
                           const float* T, float inv_p, float lam, hipStream_t stream) {
  const int k4n = k_n / 4;
  const int64_t npix = (int64_t)nx * ny;
  const dim3 grid(xcd_grid((int)((npix + 15) / 16), (k4n + QB - 1) / QB));
#define DFM_BWD(NT)                                                                                               \
  hipLaunchKernelGGL((radon_backward_kernel<NT, MODE>), grid, dim3(256), 0, stream, (const float4*)sino,          \
                     (float4*)img_k, (const AngleGeo*)geo, nx, ny, ns, na, kstride / 4, k_off / 4, k4n, skstride / 4, (float4*)x, \
                     (float4*)z, (const float4*)T, inv_p, lam)
  if (ntap == 2)
    DFM_BWD(2);
  else if (ntap == 3)
    DFM_BWD(3);
  else
    DFM_BWD(4);
#undef DFM_BWD
  DFM_LAUNCH_CHECK(who);
  return DFM_OK;
}

}  // namespace dfm

using namespace dfm;

// ================================================================================================ C ABI
extern "C" {

int dfm_version(void) { return 100; }

const char* dfm_last_error(void) { return dfm::g_err; }

int dfm_gather_k(const float* coeffs, float* img_k, int64_t npix, int64_t k0, int32_t kb, int32_t kstride,
                 dfm_stream_t stream) {
  DFM_CHECK_ARG(coeffs && img_k, "dfm_gather_k: null pointer");
  DFM_CHECK_ARG(npix > 0 && k0 >= 0 && kb >= 0 && kstride >= kb && kstride % 4 == 0,
                "dfm_gather_k: bad sizes npix=%lld kb=%d kstride=%d", (long long)npix, kb, kstride);
  dim3 grid((unsigned)((npix + TT - 1) / TT), (unsigned)((kstride + TT - 1) / TT));
  hipLaunchKernelGGL(gather_k_kernel, grid, dim3(256), 0, (hipStream_t)stream, coeffs, img_k, npix, k0, kb, kstride);
  DFM_LAUNCH_CHECK("dfm_gather_k");
  return DFM_OK;
}

int dfm_scatter_k(const float* img_k, float* coeffs, int64_t npix, int64_t k0, int32_t kb, int32_t kstride,
                  dfm_stream_t stream) {
  DFM_CHECK_ARG(coeffs && img_k, "dfm_scatter_k: null pointer");
  DFM_CHECK_ARG(npix > 0 && k0 >= 0 && kb >= 0 && kstride >= kb && kstride % 4 == 0,
                "dfm_scatter_k: bad sizes npix=%lld kb=%d kstride=%d", (long long)npix, kb, kstride);
  if (kb == 0) return DFM_OK;
  dim3 grid((unsigned)((npix + TT - 1) / TT), (unsigned)((kb + TT - 1) / TT));
  hipLaunchKernelGGL(scatter_k_kernel, grid, dim3(256), 0, (hipStream_t)stream, img_k, coeffs, npix, k0, kb, kstride);
  DFM_LAUNCH_CHECK("dfm_scatter_k");
  return DFM_OK;
}

int dfm_radon_forward(const float* img_k, float* sino, const float* geo, int32_t nx, int32_t ny, int32_t ns,
                      int32_t n_angles, int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride, float scale,
                      int32_t ntap, dfm_stream_t stream) {
  DFM_CHECK_ARG(img_k && sino && geo, "dfm_radon_forward: null pointer");
  int st = check_radon("dfm_radon_forward", nx, ny, ns, n_angles, kstride, k_off, k_n, skstride, ntap);
  if (st) return st;
  if (n_angles == 0 || k_n == 0) return DFM_OK;
  constexpr int NB = 2;
  const int k4n = k_n / 4;
  const int sblocks = (ns + 16 * NB - 1) / (16 * NB);
  const unsigned grid = xcd_grid(n_angles * sblocks, (k4n + QB - 1) / QB);
  hipLaunchKernelGGL(radon_forward_kernel<NB>, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const float4*)img_k,
                     (float4*)sino, (const AngleGeo*)geo, nx, ny, ns, n_angles, kstride / 4, k_off / 4, k4n,
                     skstride / 4, scale);
  DFM_LAUNCH_CHECK("dfm_radon_forward");
  return DFM_OK;
}

int dfm_radon_backward(const float* sino, float* img_k, const float* geo, int32_t nx, int32_t ny, int32_t ns,
                       int32_t n_angles, int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride, int32_t ntap,
                       dfm_stream_t stream) {
  DFM_CHECK_ARG(img_k && geo && (sino || n_angles == 0), "dfm_radon_backward: null pointer");
  int st = check_radon("dfm_radon_backward", nx, ny, ns, n_angles, kstride, k_off, k_n, skstride, ntap);
  if (st) return st;
  if (k_n == 0) return DFM_OK;
  return launch_backward<0>("dfm_radon_backward", sino, img_k, geo, nx, ny, ns, n_angles, kstride, k_off, k_n,
                            skstride, ntap, nullptr, nullptr, nullptr, 0.f, 0.f, (hipStream_t)stream);
}

int dfm_radon_backward_spdhg(const float* sino, const float* geo, int32_t nx, int32_t ny, int32_t ns, int32_t n_angles,
                             int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride, int32_t ntap, float* x,
                             float* z, const float* T, float inv_p, float lam, dfm_stream_t stream) {
  DFM_CHECK_ARG(x && z && T && geo && (sino || n_angles == 0), "dfm_radon_backward_spdhg: null pointer");
  int st = check_radon("dfm_radon_backward_spdhg", nx, ny, ns, n_angles, kstride, k_off, k_n, skstride, ntap);
  if (st) return st;
  if (k_n == 0) return DFM_OK;
  return launch_backward<1>("dfm_radon_backward_spdhg", sino, nullptr, geo, nx, ny, ns, n_angles, kstride, k_off, k_n,
                            skstride, ntap, x, z, T, inv_p, lam, (hipStream_t)stream);
}

int dfm_radon_forward_table(const float* img_k, float* sino, const int32_t* row_ptr, const int32_t* ent_pix,
                            const float* ent_w, int32_t ns, int32_t n_angles, int32_t kstride, int32_t k_off,
                            int32_t k_n, int32_t skstride, int32_t variant, dfm_stream_t stream) {
  DFM_CHECK_ARG(img_k && sino && row_ptr && ent_pix && ent_w, "dfm_radon_forward_table: null pointer");
  DFM_CHECK_ARG(ns > 0 && n_angles >= 0 && kstride % 4 == 0 && k_off % 4 == 0 && k_n % 4 == 0 && k_off >= 0 &&
                    k_n >= 0 && k_off + k_n <= kstride && skstride % 4 == 0 && skstride >= k_n,
                "dfm_radon_forward_table: bad sizes");
  // variant = unr + 256 + 512 q: table entries fetched ahead (4 or 8), staged in LDS (the + 256), channel blocks of
  // 64 (q 0) or 128 (q 3) channels per workgroup (the image slice an XCD streams for all angles at once)
  const int unr = variant % 256, staged = (variant / 256) % 2, q = variant / 512;
  DFM_CHECK_ARG((unr == 4 || unr == 8) && staged == 1 && (q == 0 || q == 3),
                "dfm_radon_forward_table: variant %d: entries ahead 4 or 8, + 256 (LDS-staged entries), + 512 * q "
                "for 64 (q 0) or 128 (q 3) channels per block", variant);
  if (n_angles == 0 || k_n == 0) return DFM_OK;
  const int k4n = k_n / 4;
  const int qb = q == 3 ? 2 * QB : QB;
  const unsigned grid = xcd_grid(n_angles * ((ns + 256 / qb - 1) / (256 / qb)), (k4n + qb - 1) / qb);
#define DFM_FT(U, Q)                                                                                              \
  hipLaunchKernelGGL((radon_forward_table_lds_kernel<U, Q>), dim3(grid), dim3(256), 0, (hipStream_t)stream,      \
                     (const float4*)img_k, (float4*)sino, row_ptr, ent_pix, ent_w, ns, n_angles, kstride / 4,      \
                     k_off / 4, k4n, skstride / 4)
  if (unr == 4) {
    if (qb == 16) DFM_FT(4, 16); else DFM_FT(4, 32);
  } else {
    if (qb == 16) DFM_FT(8, 16); else DFM_FT(8, 32);
  }
#undef DFM_FT
  DFM_LAUNCH_CHECK("dfm_radon_forward_table");
  return DFM_OK;
}

int dfm_radon_backward_table(const float* sino, float* img_k, const int32_t* ptab, const double* bgeo,
                             const int32_t* tbase, int32_t nx, int32_t ny, int32_t ns, int32_t n_angles,
                             int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride, float* x, float* z,
                             const float* T, float inv_p, float lam, int32_t mode, int32_t variant,
                             dfm_stream_t stream) {
  DFM_CHECK_ARG(sino && ptab && bgeo && tbase &&
                    (mode == 0 ? img_k != nullptr : mode == 2 ? (x && T) : (x && z && T)),
                "dfm_radon_backward_table: null pointer");
  DFM_CHECK_ARG(nx > 0 && ny > 0 && ns > 0 && n_angles >= 0 && kstride % 4 == 0 && k_off % 4 == 0 &&
                    k_n % 4 == 0 && k_off >= 0 && k_n >= 0 && k_off + k_n <= kstride && skstride % 4 == 0 &&
                    skstride >= k_n && mode >= 0 && mode <= 3 && (lam >= 0.0f || mode != 2),
                "dfm_radon_backward_table: bad sizes");
  // variant = 4: a workgroup per 4x4 tile and channel block (4 angles per pass; modes 0 and 1), or ACH + 64 G with
  // ACH 4 or 8 and G 2, 4 or 8: a workgroup per 4x4 tile and G channel blocks, the tile's tables staged once;
  // + 1024: 8x8 tiles (tbase then holds the 8x8 tiles' first bins), ACH 4 or 8, G 1, 2 or 4
  const int ach = variant % 64, grp = (variant / 64) % 16, t8 = (variant / 1024) % 4;
  DFM_CHECK_ARG(variant < 4096 &&
                    ((t8 == 0 && grp == 0 && ach == 4) ||
                     (t8 == 0 && (grp == 2 || grp == 4 || grp == 8) && (ach == 4 || ach == 8)) ||
                     (t8 == 1 && (grp == 1 || grp == 2 || grp == 4) && (ach == 4 || ach == 8))),
                "dfm_radon_backward_table: variant %d: 4, or 4 / 8 + 64 G for G = 2, 4, 8 channel blocks per "
                "workgroup, or 4 / 8 + 64 G + 1024 (8x8 tiles, G = 1, 2, 4)",
                variant);
  if (k_n == 0) return DFM_OK;
  const int k4n = k_n / 4;
  const int ntiles = ((nx + 3) / 4) * ((ny + 3) / 4);
  if (t8) {
    const int ntiles8 = ((nx + 7) / 8) * ((ny + 7) / 8);
    const int n_grp = ((k4n + QB - 1) / QB + grp - 1) / grp;
    const unsigned ggrid = xcd_grid(ntiles8, n_grp);
#define DFM_B8(M, A, G)                                                                                           \
  hipLaunchKernelGGL((radon_backward_group_kernel<M, A, G, 8, QB>), dim3(ggrid), dim3(1024), 0,                   \
                     (hipStream_t)stream,                                                                         \
                     (const float4*)sino, (float4*)img_k, bgeo, tbase, nx, ny, ns, n_angles,          \
                     kstride / 4, k_off / 4, k4n, skstride / 4, (float4*)x, (float4*)z, (const float4*)T, inv_p, lam)
#define DFM_B8G(M, A)        \
  if (grp == 1) {            \
    DFM_B8(M, A, 1);         \
  } else if (grp == 2) {     \
    DFM_B8(M, A, 2);         \
  } else {                   \
    DFM_B8(M, A, 4);         \
  }
    if (mode == 0) {
      if (ach == 4) { DFM_B8G(0, 4) } else { DFM_B8G(0, 8) }
    } else if (mode == 1) {
      if (ach == 4) { DFM_B8G(1, 4) } else { DFM_B8G(1, 8) }
    } else if (mode == 2) {
      if (ach == 4) { DFM_B8G(2, 4) } else { DFM_B8G(2, 8) }
    } else {
      if (ach == 4) { DFM_B8G(3, 4) } else { DFM_B8G(3, 8) }
    }
#undef DFM_B8G
#undef DFM_B8
    DFM_LAUNCH_CHECK("dfm_radon_backward_table");
    return DFM_OK;
  }
  if (grp > 0) {
    const int n_grp = ((k4n + QB - 1) / QB + grp - 1) / grp;
    const unsigned ggrid = xcd_grid(ntiles, n_grp);
#define DFM_BG(M, A, G)                                                                                           \
  hipLaunchKernelGGL((radon_backward_group_kernel<M, A, G, 4, QB>), dim3(ggrid), dim3(256), 0,                    \
                     (hipStream_t)stream,                                                                         \
                     (const float4*)sino, (float4*)img_k, bgeo, tbase, nx, ny, ns, n_angles,          \
                     kstride / 4, k_off / 4, k4n, skstride / 4, (float4*)x, (float4*)z, (const float4*)T, inv_p, lam)
#define DFM_BGA(M, G)            \
  if (ach == 4) {                \
    DFM_BG(M, 4, G);             \
  } else {                       \
    DFM_BG(M, 8, G);             \
  }
#define DFM_BGG(M)               \
  if (grp == 2) {                \
    DFM_BGA(M, 2)                \
  } else if (grp == 4) {         \
    DFM_BGA(M, 4)                \
  } else {                       \
    DFM_BGA(M, 8)                \
  }
    if (mode == 0) {
      DFM_BGG(0)
    } else if (mode == 1) {
      DFM_BGG(1)
    } else if (mode == 2) {
      DFM_BGG(2)
    } else {
      DFM_BGG(3)
    }
#undef DFM_BGG
#undef DFM_BGA
#undef DFM_BG
    DFM_LAUNCH_CHECK("dfm_radon_backward_table");
    return DFM_OK;
  }
  if (mode >= 2) {
    dfm::set_error("dfm_radon_backward_table: mode %d (the EM / PDHG update) needs a grouped variant (+ 64 G, "
                   "+ 1024), not %d", mode, variant);
    return DFM_ERR_UNSUPPORTED;
  }
  const unsigned grid = xcd_grid(ntiles, (k4n + QB - 1) / QB);
#define DFM_BT(M, A)                                                                                              \
  hipLaunchKernelGGL((radon_backward_tile_kernel<M, A>), dim3(grid), dim3(256), 0, (hipStream_t)stream,          \
                     (const float4*)sino, (float4*)img_k, (const int4*)ptab, tbase, nx, ny, ns, n_angles,          \
                     kstride / 4, k_off / 4, k4n, skstride / 4, (float4*)x, (float4*)z, (const float4*)T, inv_p, lam)
  if (mode == 0) {
    DFM_BT(0, 4);
  } else {
    DFM_BT(1, 4);
  }
#undef DFM_BT
  DFM_LAUNCH_CHECK("dfm_radon_backward_table");
  return DFM_OK;
}

int dfm_pf_eval_dense(const double* coords, const double* poles, const int32_t* axis_start, const float* axis_count,
                      const double* inv_sigma2, const double* norm, const float* qw, float* out, int32_t R,
                      int32_t kmax, int32_t C, int32_t P, int32_t A, int32_t S, int32_t Se, int32_t k0, int32_t kb,
                      float pf_cut, int32_t quad_weighted, dfm_stream_t stream) {
  DFM_CHECK_ARG(coords && poles && axis_start && axis_count && inv_sigma2 && norm && out && (qw || !quad_weighted),
                "dfm_pf_eval_dense: null pointer");
  DFM_CHECK_ARG(R > 0 && kmax > 0 && C > 0 && P > 0 && A >= P && S > 0 && Se > 0 && k0 >= 0 && kb >= 0 &&
                    kb <= kmax,
                "dfm_pf_eval_dense: bad sizes");
  const int64_t total = (int64_t)R * kmax * C * P;
  hipLaunchKernelGGL(pf_eval_kernel, dim3(grid_stride_blocks(total, 256, 1u << 20)), dim3(256), 0, (hipStream_t)stream,
                     coords, poles, axis_start, axis_count, inv_sigma2, norm, qw, out, R, kmax, C, P, A, S, Se, k0, kb,
                     pf_cut, quad_weighted);
  DFM_LAUNCH_CHECK("dfm_pf_eval_dense");
  return DFM_OK;
}

int dfm_pf_count_fwd(const float* pf, int32_t* counts, int32_t R, int32_t kmax, int32_t CP, int32_t kb,
                     dfm_stream_t stream) {
  DFM_CHECK_ARG(pf && counts && R > 0 && kmax > 0 && CP > 0 && kb <= kmax && kb <= 65536 && CP <= 65536,
                "dfm_pf_count_fwd: bad arguments");
  hipLaunchKernelGGL(pf_count_fwd_kernel, dim3((CP + 255) / 256, R), dim3(256), 0, (hipStream_t)stream, pf, counts,
                     kmax, CP, kb);
  DFM_LAUNCH_CHECK("dfm_pf_count_fwd");
  return DFM_OK;
}

int dfm_pf_fill_fwd(const float* pf, const int32_t* row_ptr, uint16_t* col_k, float* val, int32_t R, int32_t kmax,
                    int32_t CP, int32_t kb, dfm_stream_t stream) {
  DFM_CHECK_ARG(pf && row_ptr && col_k && val && R > 0 && kmax > 0 && CP > 0 && kb <= kmax && kb <= 65536,
                "dfm_pf_fill_fwd: bad arguments");
  hipLaunchKernelGGL(pf_fill_fwd_kernel, dim3((CP + 255) / 256, R), dim3(256), 0, (hipStream_t)stream, pf, row_ptr,
                     col_k, val, kmax, CP, kb);
  DFM_LAUNCH_CHECK("dfm_pf_fill_fwd");
  return DFM_OK;
}

int dfm_pf_count_adj(const float* pf, int32_t* counts, int32_t R, int32_t kmax, int32_t CP, int32_t kb,
                     dfm_stream_t stream) {
  DFM_CHECK_ARG(pf && counts && R > 0 && kmax > 0 && CP > 0 && kb <= kmax && CP <= 65536,
                "dfm_pf_count_adj: bad arguments");
  const int64_t rows = (int64_t)R * kb;
  if (rows == 0) return DFM_OK;
  hipLaunchKernelGGL(pf_count_adj_kernel, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, (hipStream_t)stream, pf,
                     counts, R, kmax, CP, kb);
  DFM_LAUNCH_CHECK("dfm_pf_count_adj");
  return DFM_OK;
}

int dfm_pf_fill_adj(const float* pf, const int32_t* row_ptr, uint16_t* col_j, float* val, int32_t R, int32_t kmax,
                    int32_t CP, int32_t kb, dfm_stream_t stream) {
  DFM_CHECK_ARG(pf && row_ptr && col_j && val && R > 0 && kmax > 0 && CP > 0 && kb <= kmax && CP <= 65536,
                "dfm_pf_fill_adj: bad arguments");
  const int64_t rows = (int64_t)R * kb;
  if (rows == 0) return DFM_OK;
  hipLaunchKernelGGL(pf_fill_adj_kernel, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, (hipStream_t)stream, pf,
                     row_ptr, col_j, val, R, kmax, CP, kb);
  DFM_LAUNCH_CHECK("dfm_pf_fill_adj");
  return DFM_OK;
}

int dfm_spmm_forward(const float* sino, const int32_t* row_ptr, const uint16_t* col_k, const float* val, float* data,
                     const int32_t* omegas, int32_t n_angles, int32_t my, int32_t cp, int32_t kstride, int32_t kb,
                     int32_t compact, int32_t accumulate, dfm_stream_t stream) {
  DFM_CHECK_ARG(sino && row_ptr && col_k && val && data, "dfm_spmm_forward: null pointer");
  DFM_CHECK_ARG(n_angles >= 0 && my > 0 && cp > 0 && kb >= 0 && kb <= kstride && kb <= 65536,
                "dfm_spmm_forward: bad sizes");
  if (n_angles == 0) return DFM_OK;
  const int kc = max(1, min(max(kb, 1), SPMM_LDS_FLOATS / SPMM_TY));
  dim3 grid((unsigned)((cp + 255) / 256), (unsigned)((my + SPMM_TY - 1) / SPMM_TY), (unsigned)n_angles);
  hipLaunchKernelGGL(spmm_forward_kernel, grid, dim3(256), (size_t)SPMM_TY * kc * sizeof(float), (hipStream_t)stream,
                     sino, row_ptr, col_k, val, data, omegas, my, cp, kstride, kb, kc, compact, accumulate, 1, kb,
                     (float*)nullptr, n_angles);
  DFM_LAUNCH_CHECK("dfm_spmm_forward");
  return DFM_OK;
}

int dfm_spmm_forward_split(const float* sino, const int32_t* row_ptr, const uint16_t* col_k, const float* val,
                           float* data, const int32_t* omegas, int32_t n_angles, int32_t my, int32_t cp,
                           int32_t kstride, int32_t kb, int32_t compact, int32_t accumulate, float* part,
                           int32_t ksplit, int32_t kc, dfm_stream_t stream) {
  DFM_CHECK_ARG(sino && row_ptr && col_k && val && data, "dfm_spmm_forward_split: null pointer");
  DFM_CHECK_ARG(n_angles >= 0 && my > 0 && cp > 0 && kb >= 0 && kb <= kstride && kb <= 65536,
                "dfm_spmm_forward_split: bad sizes");
  DFM_CHECK_ARG(ksplit >= 1 && ksplit <= 64 && (ksplit == 1 || part), "dfm_spmm_forward_split: k split %d needs "
                "a workspace of ksplit * n_angles * my * cp floats (ksplit 1 .. 64)", ksplit);
  DFM_CHECK_ARG(kc >= 32 && kc <= SPMM_LDS_FLOATS / SPMM_TY, "dfm_spmm_forward_split: chunk %d not in 32 .. %d", kc,
                SPMM_LDS_FLOATS / SPMM_TY);
  if (n_angles == 0) return DFM_OK;
  const int kpart = (max(kb, 1) + ksplit - 1) / ksplit;
  const int kcc = min(kc, max(kpart, 1));
  dim3 grid((unsigned)((cp + 255) / 256), (unsigned)((my + SPMM_TY - 1) / SPMM_TY), (unsigned)(n_angles * ksplit));
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(spmm_forward_kernel, grid, dim3(256), (size_t)SPMM_TY * kcc * sizeof(float), st, sino, row_ptr,
                     col_k, val, data, omegas, my, cp, kstride, kb, kcc, compact, accumulate, ksplit, kpart, part,
                     n_angles);
  DFM_LAUNCH_CHECK("dfm_spmm_forward_split");
  if (ksplit > 1) return dfm::launch_splitk_reduce(part, data, omegas, n_angles, (int64_t)my * cp, ksplit, accumulate, st);
  return DFM_OK;
}

int dfm_spmm_adjoint(const float* data, const int32_t* row_ptr, const uint16_t* col_j, const float* val, float* sino,
                     const int32_t* omegas, int32_t n_angles, int32_t my, int32_t cp, int32_t kstride, int32_t kb,
                     float alpha, int32_t compact, dfm_stream_t stream) {
  DFM_CHECK_ARG(sino && row_ptr && col_j && val && data, "dfm_spmm_adjoint: null pointer");
  DFM_CHECK_ARG(n_angles >= 0 && my > 0 && cp > 0 && kb >= 0 && kb <= kstride && cp <= 65536,
                "dfm_spmm_adjoint: bad sizes");
  if (n_angles == 0) return DFM_OK;
  const int jc = max(1, min(cp, SPMM_LDS_FLOATS / SPMM_TY));
  dim3 grid((unsigned)((kstride + 255) / 256), (unsigned)((my + SPMM_TY - 1) / SPMM_TY), (unsigned)n_angles);
  hipLaunchKernelGGL(spmm_adjoint_kernel, grid, dim3(256), (size_t)SPMM_TY * jc * sizeof(float), (hipStream_t)stream,
                     data, row_ptr, col_j, val, sino, omegas, my, cp, kstride, kb, jc, alpha, compact);
  DFM_LAUNCH_CHECK("dfm_spmm_adjoint");
  return DFM_OK;
}

int dfm_zero_rows(float* data, const int32_t* omegas, int32_t n_sub, int64_t row_len, dfm_stream_t stream) {
  DFM_CHECK_ARG(data && omegas && n_sub >= 0 && row_len > 0, "dfm_zero_rows: bad arguments");
  const int64_t n = (int64_t)n_sub * row_len;
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(zero_rows_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, data,
                     omegas, row_len, n);
  DFM_LAUNCH_CHECK("dfm_zero_rows");
  return DFM_OK;
}

}  // extern "C"
