// Projector kernels with a channel group's whole image resident in LDS (radon_kernels.cl:44-141 radon_forward_k /
// radon_backward_k; spdhg_poisson.py:95-106 spdhg_update), including the fused projector step of SPDHG.
//
// A workgroup owns CW channels (orientations) of the batch over the whole image (Npix x CW floats: 157 KB at 99 x 99
// pixels and CW = 4, the LDS of a CU holds 160 KB).
// * Backprojection (phase A): the CW channels of the sinogram are staged in LDS as many angles at a time as fit,
//   every thread accumulates its pixels over the angles in registers, then writes the image or applies the SPDHG
//   primal update z += A_m^T D, x = prox(x - T (z + inv_p dz + lam)).
// * Forward projection (phase B): the image (the updated x in the fused step, else read from HBM) is in LDS and a
//   thread per (angle, bin) sums its ray from LDS: x is read from HBM once per projection, never through the L2
//   once per angle.
// The fused step runs A then B in one launch: the backprojection of subset m, the primal update and the forward
// projection of the next subset, with x read and written once per SPDHG step.
#include "dfm_common.h"

namespace {

struct AngleGeo {
  float X, Y, T0, wscale, hw, top, pad0, pad1;
};

constexpr int N_XCD = 8;
constexpr int THREADS = 1024;
constexpr int LDS_FLOATS = 40960;   // 160 KiB, the LDS of one CU

__device__ __forceinline__ int64_t xcd_block() {
  const int64_t nb = gridDim.x, i = blockIdx.x;
  return (i % N_XCD) * (nb / N_XCD) + i / N_XCD;
}

template <int CW>
__device__ __forceinline__ void ldv(const float* p, float v[CW]) {
  if constexpr (CW == 4) {
    const float4 t = *reinterpret_cast<const float4*>(p);
    v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
  } else if constexpr (CW == 2) {
    const float2 t = *reinterpret_cast<const float2*>(p);
    v[0] = t.x; v[1] = t.y;
  } else {
    v[0] = *p;
  }
}

template <int CW>
__device__ __forceinline__ void stv(float* p, const float v[CW]) {
  if constexpr (CW == 4) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
  } else if constexpr (CW == 2) {
    *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
  } else {
    *p = v[0];
  }
}

// MODE: 0 backprojection into img (x), 1 backprojection + SPDHG primal update of (x, z), 2 forward projection of x,
// 3 = 1 followed by the forward projection of the updated x (the fused step).
enum { BWD_IMG = 0, BWD_SPDHG = 1, FWD = 2, FUSED = 3 };

template <int CW, int MODE>
__global__ void __launch_bounds__(THREADS) projector_lds_kernel(
    // sino_b (read) and sino_f (written) are not __restrict__: the fused step (MODE 3) reads the backprojection's
    // sinogram from the same buffer its forward projection then writes
    const float* sino_b, const AngleGeo* __restrict__ geo_b, int na_b, float* __restrict__ x,
    float* __restrict__ z, const float* __restrict__ T, float inv_p, float lam, float* sino_f,
    const AngleGeo* __restrict__ geo_f, int na_f, float scale, int nx, int ny, int ns, int kstride, int k_off,
    int k_n, int skstride) {
  constexpr int PPT = LDS_FLOATS / CW / THREADS;      // pixels per thread at most (npix * CW <= LDS_FLOATS)
  constexpr bool BACK = MODE != FWD;
  constexpr bool FORW = MODE == FWD || MODE == FUSED;
  __shared__ float lds[LDS_FLOATS];
  const int kc = (int)xcd_block() * CW;                // first channel of this workgroup, within the batch
  if (kc >= k_n) return;                               // uniform
  const int tid = threadIdx.x;
  const int npix = nx * ny;

  if (BACK) {
    // ---- phase A: backprojection (+ primal update); the sinogram's channels staged in LDS, ach angles at a time
    const int ach = LDS_FLOATS / (ns * CW);
    float acc[PPT][CW];
#pragma unroll
    for (int j = 0; j < PPT; ++j)
#pragma unroll
      for (int c = 0; c < CW; ++c) acc[j][c] = 0.f;
    for (int a0 = 0; a0 < na_b; a0 += ach) {
      const int na = min(ach, na_b - a0);
      if (a0 > 0) __syncthreads();                     // the previous chunk is no longer read
      for (int i = tid; i < na * ns; i += THREADS) {
        float v[CW];
        ldv<CW>(sino_b + ((int64_t)a0 * ns + i) * skstride + kc, v);
        stv<CW>(lds + i * CW, v);
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        const int p = tid + j * THREADS;
        if (p >= npix) break;
        const float px = (float)(p % nx), py = (float)(p / nx);
        for (int a = 0; a < na; ++a) {
          const AngleGeo g = geo_b[a0 + a];
          const float t = g.X * px + g.Y * py + g.T0;
          const int sf = (int)floorf(t - g.hw) + 1;
          float v[CW];
#pragma unroll
          for (int c = 0; c < CW; ++c) v[c] = 0.f;
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const int s = sf + q;
            if (s >= 0 && s < ns) {
              const float w = fminf(fmaxf(g.hw - fabsf(t - (float)s), 0.0f), g.top);
              float sv[CW];
              ldv<CW>(lds + (a * ns + s) * CW, sv);
#pragma unroll
              for (int c = 0; c < CW; ++c) v[c] += w * sv[c];
            }
          }
#pragma unroll
          for (int c = 0; c < CW; ++c) acc[j][c] += g.wscale * v[c];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const int p = tid + j * THREADS;
      if (p >= npix) break;
      const int64_t o = (int64_t)p * kstride + k_off + kc;
      if (MODE == BWD_IMG) {
        stv<CW>(x + o, acc[j]);
      } else {
        float xx[CW], zz[CW], tt[CW];
        ldv<CW>(x + o, xx);
        ldv<CW>(z + o, zz);
        ldv<CW>(T + o, tt);
#pragma unroll
        for (int c = 0; c < CW; ++c) {
          const float zn = zz[c] + acc[j][c];
          xx[c] = tt[c] > 0.0f ? fmaxf(0.0f, xx[c] - tt[c] * ((zn + inv_p * acc[j][c]) + lam)) : 0.0f;
          zz[c] = zn;
          acc[j][c] = xx[c];                           // the new x, for the forward projection
        }
        stv<CW>(z + o, zz);
        stv<CW>(x + o, xx);
      }
    }
    if (FORW) {
      __syncthreads();                                 // the sinogram in LDS is no longer read
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        const int p = tid + j * THREADS;
        if (p >= npix) break;
        stv<CW>(lds + p * CW, acc[j]);
      }
    }
  } else {
    for (int p = tid; p < npix; p += THREADS) {
      float v[CW];
      ldv<CW>(x + (int64_t)p * kstride + k_off + kc, v);
      stv<CW>(lds + p * CW, v);
    }
  }
  if (!FORW) return;
  __syncthreads();

  // ---- phase B: forward projection, a thread per (angle, bin); consecutive threads take consecutive bins, so the
  // LDS reads of a wave at one image row are consecutive pixels
  for (int i = tid; i < na_f * ns; i += THREADS) {
    const int a = i / ns;
    const int s = i - a * ns;
    const AngleGeo g = geo_f[a];
    const bool rows_y = fabsf(g.X) >= fabsf(g.Y);
    const float A = rows_y ? g.X : g.Y;
    const float B = rows_y ? g.Y : g.X;
    const int nu = rows_y ? nx : ny;
    const int nv = rows_y ? ny : nx;
    const int du = rows_y ? CW : nx * CW;
    const int dv = rows_y ? nx * CW : CW;
    const float inv = 1.0f / A;
    float acc[CW];
#pragma unroll
    for (int c = 0; c < CW; ++c) acc[c] = 0.f;
    for (int v = 0; v < nv; ++v) {
      const float cv = B * (float)v + g.T0;
      const float e1 = ((float)s - g.hw - cv) * inv;
      const float e2 = ((float)s + g.hw - cv) * inv;
      const int u0 = max((int)floorf(fminf(e1, e2)), 0);
      const int u1 = min((int)ceilf(fmaxf(e1, e2)), nu - 1);
      for (int u = u0; u <= u1; ++u) {
        const float t = A * (float)u + cv - (float)s;
        const float w = fminf(fmaxf(g.hw - fabsf(t), 0.0f), g.top);
        float xv[CW];
        ldv<CW>(lds + v * dv + u * du, xv);
#pragma unroll
        for (int c = 0; c < CW; ++c) acc[c] += w * xv[c];
      }
    }
    const float f = scale * g.wscale;
#pragma unroll
    for (int c = 0; c < CW; ++c) acc[c] *= f;
    stv<CW>(sino_f + (int64_t)i * skstride + kc, acc);
  }
}

template <int MODE>
void launch_projector(int cw, dim3 grid, hipStream_t st, const float* sino_b, const AngleGeo* geo_b, int na_b,
                      float* x, float* z, const float* T, float inv_p, float lam, float* sino_f,
                      const AngleGeo* geo_f, int na_f, float scale, int nx, int ny, int ns, int kstride, int k_off,
                      int k_n, int skstride) {
#define DFM_PROJ(CW)                                                                                               \
  hipLaunchKernelGGL((projector_lds_kernel<CW, MODE>), grid, dim3(THREADS), 0, st, sino_b, geo_b, na_b, x, z, T,   \
                     inv_p, lam, sino_f, geo_f, na_f, scale, nx, ny, ns, kstride, k_off, k_n, skstride)
  if (cw == 4) DFM_PROJ(4);
  else if (cw == 2) DFM_PROJ(2);
  else DFM_PROJ(1);
#undef DFM_PROJ
}

}  // namespace

extern "C" {

int dfm_projector_lds_width(int32_t nx, int32_t ny, int32_t ns) {
  const int64_t npix = (int64_t)nx * ny;
  for (int cw = 4; cw >= 1; cw /= 2)
    if (npix * cw <= LDS_FLOATS && (int64_t)ns * cw <= LDS_FLOATS) return cw;
  return 0;
}

int dfm_projector_lds(int32_t mode, const float* sino_b, const void* geo_b, int32_t na_b, float* x, float* z,
                      const float* T, float inv_p, float lam, float* sino_f, const void* geo_f, int32_t na_f,
                      float scale, int32_t nx, int32_t ny, int32_t ns, int32_t kstride, int32_t k_off, int32_t k_n,
                      int32_t skstride, int32_t cw, dfm_stream_t stream) {
  DFM_CHECK_ARG(mode >= BWD_IMG && mode <= FUSED, "dfm_projector_lds: mode %d not 0 .. 3", mode);
  const bool back = mode != FWD, forw = mode == FWD || mode == FUSED;
  DFM_CHECK_ARG(x && (!back || (sino_b && geo_b && na_b > 0)) && (!forw || (sino_f && geo_f && na_f > 0)) &&
                    (mode != BWD_SPDHG && mode != FUSED || (z && T)),
                "dfm_projector_lds: null pointer or no angles for mode %d", mode);
  DFM_CHECK_ARG(nx > 0 && ny > 0 && ns > 0, "dfm_projector_lds: bad sizes");
  DFM_CHECK_ARG(cw == 4 || cw == 2 || cw == 1, "dfm_projector_lds: channel width %d not 4, 2 or 1", cw);
  DFM_CHECK_ARG((int64_t)nx * ny * cw <= LDS_FLOATS && (int64_t)ns * cw <= LDS_FLOATS,
                "dfm_projector_lds: %d x %d pixels (or %d bins) at channel width %d exceed the LDS", nx, ny, ns, cw);
  DFM_CHECK_ARG(k_n > 0 && k_n % cw == 0 && k_off % cw == 0 && kstride % cw == 0 && skstride % cw == 0 &&
                    k_off + k_n <= kstride && k_n <= skstride,
                "dfm_projector_lds: channel range %d + %d, strides %d / %d not aligned to %d", k_off, k_n, kstride,
                skstride, cw);
  int64_t groups = k_n / cw;
  groups = (groups + N_XCD - 1) / N_XCD * N_XCD;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)groups);
  const AngleGeo* gb = (const AngleGeo*)geo_b;
  const AngleGeo* gf = (const AngleGeo*)geo_f;
  switch (mode) {
    case BWD_IMG:
      launch_projector<BWD_IMG>(cw, grid, st, sino_b, gb, na_b, x, z, T, inv_p, lam, sino_f, gf, na_f, scale, nx, ny,
                                ns, kstride, k_off, k_n, skstride);
      break;
    case BWD_SPDHG:
      launch_projector<BWD_SPDHG>(cw, grid, st, sino_b, gb, na_b, x, z, T, inv_p, lam, sino_f, gf, na_f, scale, nx,
                                  ny, ns, kstride, k_off, k_n, skstride);
      break;
    case FWD:
      launch_projector<FWD>(cw, grid, st, sino_b, gb, na_b, x, z, T, inv_p, lam, sino_f, gf, na_f, scale, nx, ny, ns,
                            kstride, k_off, k_n, skstride);
      break;
    default:
      launch_projector<FUSED>(cw, grid, st, sino_b, gb, na_b, x, z, T, inv_p, lam, sino_f, gf, na_f, scale, nx, ny,
                              ns, kstride, k_off, k_n, skstride);
  }
  DFM_LAUNCH_CHECK("dfm_projector_lds");
  return DFM_OK;
}

}  // extern "C"
