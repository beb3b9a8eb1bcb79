// Element-wise kernels and float64 reductions of the SPDHG-Poisson solver on gfx950.
#include <mutex>

#include "dfm_common.h"

namespace dfm {

__device__ __forceinline__ float pval(const float* __restrict__ a, int64_t p, int64_t i, float dflt) {
  return p ? a[i % p] : dflt;
}

// y' = prox_{s F*}(y + s Ax) on the rows of a subset, float64 arithmetic; d: A x in, y' - y out.
__global__ void __launch_bounds__(256) dual_update_sub_kernel(float* __restrict__ y, float* __restrict__ d,
                                                              const float* __restrict__ sig, const float* __restrict__ b,
                                                              const float* __restrict__ w, int64_t pw,
                                                              const float* __restrict__ g, int64_t pg,
                                                              const float* __restrict__ bg, int64_t pbg, float eps,
                                                              const int* __restrict__ omegas, int64_t row_len,
                                                              int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = t / row_len;
    const int64_t i = (int64_t)omegas[s] * row_len + (t - s * row_len);
    const float wi = pval(w, pw, i, 1.0f);
    const float si = sig[i];
    const float yi = y[i];
    float yn = 0.0f;
    if (wi != 0.0f && si != 0.0f) {
      const float gi = pval(g, pg, i, 1.0f);
      const double sc = (double)wi / (double)gi;
      const double r = (double)pval(bg, pbg, i, 0.0f) + (double)eps * gi;
      const double v = (double)yi + (double)si * (double)d[i] + (double)si * r;
      const double dd = v - sc;
      yn = (float)(0.5 * (v + sc - sqrt(dd * dd + 4.0 * (double)si * sc * (double)b[i])));
    }
    d[i] = yn - yi;
    y[i] = yn;
  }
}

__global__ void __launch_bounds__(256) spdhg_update_kernel(float* __restrict__ x, float* __restrict__ z,
                                                           const float* __restrict__ T, const float* __restrict__ dz,
                                                           float inv_p, float lam, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float zi = z[i] + dz[i];
    z[i] = zi;
    const float zb = zi + inv_p * dz[i];
    const float t = T[i];
    x[i] = t > 0.0f ? fmaxf(0.0f, x[i] - t * (zb + lam)) : 0.0f;
  }
}

// (pixel, channel) kernels over channel-fastest (npix, kstride) arrays: blockIdx.y runs over pixels.
__global__ void __launch_bounds__(256) tinit_kernel(float* __restrict__ T, int64_t npix, int kstride) {
  for (int64_t p = blockIdx.y; p < npix; p += gridDim.y)
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kstride; k += gridDim.x * blockDim.x)
      T[p * kstride + k] = INFINITY;
}

__global__ void __launch_bounds__(256) tmin_kernel(float* __restrict__ T, const float* __restrict__ colsum,
                                                   const unsigned char* __restrict__ supp, int64_t npix, int kstride,
                                                   int kvalid, float c) {
  for (int64_t p = blockIdx.y; p < npix; p += gridDim.y) {
    const bool inside = supp ? supp[p] != 0 : true;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kvalid; k += gridDim.x * blockDim.x) {
      const int64_t i = p * kstride + k;
      const float cs = colsum[i];
      if (inside && cs > 0.0f) T[i] = fminf(T[i], c / cs);
    }
  }
}

__global__ void __launch_bounds__(256) tfinal_kernel(float* __restrict__ T, int64_t npix, int kstride, int kvalid) {
  for (int64_t p = blockIdx.y; p < npix; p += gridDim.y)
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kstride; k += gridDim.x * blockDim.x) {
      const int64_t i = p * kstride + k;
      if (k >= kvalid || isinf(T[i])) T[i] = 0.0f;
    }
}

__global__ void __launch_bounds__(256) support_ones_kernel(float* __restrict__ x, const unsigned char* __restrict__ supp,
                                                           int64_t npix, int kstride, int kvalid) {
  for (int64_t p = blockIdx.y; p < npix; p += gridDim.y) {
    const float in = (supp ? supp[p] != 0 : true) ? 1.0f : 0.0f;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kstride; k += gridDim.x * blockDim.x)
      x[p * kstride + k] = k < kvalid ? in : 0.0f;
  }
}

__global__ void __launch_bounds__(256) project_support_kernel(float* __restrict__ x,
                                                              const unsigned char* __restrict__ supp, int64_t npix,
                                                              int kstride, int kvalid) {
  for (int64_t p = blockIdx.y; p < npix; p += gridDim.y) {
    const bool inside = supp ? supp[p] != 0 : true;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kstride; k += gridDim.x * blockDim.x) {
      const int64_t i = p * kstride + k;
      if (!inside || k >= kvalid || x[i] < 0.0f) x[i] = 0.0f;
    }
  }
}

__global__ void __launch_bounds__(256) dual_steps_kernel(float* __restrict__ sig, const float* __restrict__ rowsum,
                                                         const float* __restrict__ w, int64_t pw, float sigma0,
                                                         int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float rs = rowsum[i];
    sig[i] = (pval(w, pw, i, 1.0f) != 0.0f && rs > 0.0f) ? sigma0 / rs : 0.0f;
  }
}

__global__ void __launch_bounds__(256) set_mask_ones_kernel(float* __restrict__ u, const float* __restrict__ w,
                                                            int64_t pw, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    u[i] = pval(w, pw, i, 1.0f) != 0.0f ? 1.0f : 0.0f;
}

__global__ void __launch_bounds__(256) scale_kernel(float* __restrict__ a, float s, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] *= s;
}

// Block reduction of NV doubles (256 threads: wave shuffles, then LDS across the 4 waves).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* __restrict__ out, int stride) {
  __shared__ double red[NV][4];
#pragma unroll
  for (int q = 0; q < NV; ++q)
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) red[q][wv] = v[q];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q * stride + blockIdx.x] = red[q][0] + red[q][1] + red[q][2] + red[q][3];
}

__global__ void __launch_bounds__(256) deviance_kernel(const float* __restrict__ mu, const float* __restrict__ b,
                                                       const float* __restrict__ w, int64_t pw,
                                                       const float* __restrict__ g, int64_t pg,
                                                       const float* __restrict__ bg, int64_t pbg, float eps, int64_t n,
                                                       double* __restrict__ partial) {
  double s[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float wi = pval(w, pw, i, 1.0f);
    if (wi != 0.0f) {
      const float gi = pval(g, pg, i, 1.0f);
      const double bi = (double)b[i];
      const double sc = (double)wi / (double)gi;
      const double m = (double)fmaxf(mu[i], 0.0f) + (double)pval(bg, pbg, i, 0.0f) + (double)eps * gi;
      double v = m - bi;
      if (bi > 0.0) v += bi * log(bi / m);
      s[0] += sc * v;
    }
  }
  block_sum<1>(s, partial, gridDim.x);
}

__global__ void __launch_bounds__(256) mean_counted_kernel(const float* __restrict__ b, const float* __restrict__ w,
                                                           int64_t pw, int64_t n, double* __restrict__ partial) {
  double s[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (pval(w, pw, i, 1.0f) != 0.0f) {
      s[0] += (double)b[i];
      s[1] += 1.0;
    }
  }
  block_sum<2>(s, partial, gridDim.x);
}

__global__ void __launch_bounds__(256) dot_kernel(const float* __restrict__ a, const float* __restrict__ b, int64_t n,
                                                  double* __restrict__ partial) {
  double s[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s[0] += (double)a[i] * (double)b[i];
  block_sum<1>(s, partial, gridDim.x);
}

// ---- total-variation block (spdhg_poisson.py:108-193) on the channel-fastest layout: element i = p * k4 + q of a
// float4 array (pixel p = ix + nx * iy, channel quad q); the right neighbour is i + k4, the lower one i + nx * k4.
// A difference is free when T > 0 at both ends.
__device__ __forceinline__ void tv_core(float tj, float tr, float td, bool vr0, bool vd0, float xj, float xr, float xd,
                                        float ox, float oy, float lam, float ps, float& nx_, float& ny_) {
  const bool vr = vr0 && tj > 0.0f && tr > 0.0f;
  const bool vd = vd0 && tj > 0.0f && td > 0.0f;
  float m = 0.0f;
  if (vr) m = fmaxf(m, tj + tr);
  if (vd) m = fmaxf(m, tj + td);
  float ax = 0.0f, ay = 0.0f;
  if (m > 0.0f) {
    const float s = ps / (4.0f * m);
    if (vr) ax = ox + s * (xr - xj);
    if (vd) ay = oy + s * (xd - xj);
    const float nrm = sqrtf(ax * ax + ay * ay);
    if (nrm > lam) {
      const float c = lam / nrm;
      ax *= c;
      ay *= c;
    }
  }
  nx_ = ax;
  ny_ = ay;
}

// q <- proj_{|q| <= lam}(q + s D x) on the free differences (0 elsewhere); dq = q' - q
__global__ void __launch_bounds__(256) tv_dual_kernel(const float4* __restrict__ x, const float4* __restrict__ T,
                                                      float4* __restrict__ qx, float4* __restrict__ qy,
                                                      float4* __restrict__ dqx, float4* __restrict__ dqy, int nx,
                                                      int ny, int k4, float lam, float ps) {
  const int64_t n = (int64_t)nx * ny * k4;
  const int64_t row = (int64_t)nx * k4;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / k4;
    const int ix = (int)(p % nx), iy = (int)(p / nx);
    const bool vr0 = ix + 1 < nx, vd0 = iy + 1 < ny;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 tj = T[i], tr = vr0 ? T[i + k4] : z4, td = vd0 ? T[i + row] : z4;
    const float4 xj = x[i], xr = vr0 ? x[i + k4] : z4, xd = vd0 ? x[i + row] : z4;
    const float4 ox = qx[i], oy = qy[i];
    float4 nxv, nyv;
    tv_core(tj.x, tr.x, td.x, vr0, vd0, xj.x, xr.x, xd.x, ox.x, oy.x, lam, ps, nxv.x, nyv.x);
    tv_core(tj.y, tr.y, td.y, vr0, vd0, xj.y, xr.y, xd.y, ox.y, oy.y, lam, ps, nxv.y, nyv.y);
    tv_core(tj.z, tr.z, td.z, vr0, vd0, xj.z, xr.z, xd.z, ox.z, oy.z, lam, ps, nxv.z, nyv.z);
    tv_core(tj.w, tr.w, td.w, vr0, vd0, xj.w, xr.w, xd.w, ox.w, oy.w, lam, ps, nxv.w, nyv.w);
    qx[i] = nxv;
    qy[i] = nyv;
    dqx[i] = make_float4(nxv.x - ox.x, nxv.y - ox.y, nxv.z - ox.z, nxv.w - ox.w);
    dqy[i] = make_float4(nyv.x - oy.x, nyv.y - oy.y, nyv.z - oy.z, nyv.w - oy.w);
  }
}

// dz = D^T dq = -dqx - dqy + dqx(left) + dqy(up), then the primal update z += dz, x = prox(x - T (z + inv_p dz + lam))
__global__ void __launch_bounds__(256) tv_adjoint_update_kernel(const float4* __restrict__ dqx,
                                                                const float4* __restrict__ dqy,
                                                                float4* __restrict__ x, float4* __restrict__ z,
                                                                const float4* __restrict__ T, int nx, int ny, int k4,
                                                                float inv_p, float lam) {
  const int64_t n = (int64_t)nx * ny * k4;
  const int64_t row = (int64_t)nx * k4;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / k4;
    const int ix = (int)(p % nx), iy = (int)(p / nx);
    const float4 a = dqx[i], b = dqy[i];
    float4 dz = make_float4(-a.x - b.x, -a.y - b.y, -a.z - b.z, -a.w - b.w);
    if (ix > 0) {
      const float4 l = dqx[i - k4];
      dz.x += l.x; dz.y += l.y; dz.z += l.z; dz.w += l.w;
    }
    if (iy > 0) {
      const float4 u = dqy[i - row];
      dz.x += u.x; dz.y += u.y; dz.z += u.z; dz.w += u.w;
    }
    float4 zz = z[i], xx = x[i];
    const float4 tt = T[i];
#define DFM_TVUPD(c)                                                                           \
  zz.c += dz.c;                                                                                \
  xx.c = tt.c > 0.0f ? fmaxf(0.0f, xx.c - tt.c * ((zz.c + inv_p * dz.c) + lam)) : 0.0f;
    DFM_TVUPD(x) DFM_TVUPD(y) DFM_TVUPD(z) DFM_TVUPD(w)
#undef DFM_TVUPD
    z[i] = zz;
    x[i] = xx;
  }
}

__global__ void __launch_bounds__(256) tv_value_kernel(const float4* __restrict__ x, const float4* __restrict__ T,
                                                       int nx, int ny, int k4, double* __restrict__ partial) {
  const int64_t n = (int64_t)nx * ny * k4;
  const int64_t row = (int64_t)nx * k4;
  double s[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / k4;
    const int ix = (int)(p % nx), iy = (int)(p / nx);
    const bool vr0 = ix + 1 < nx, vd0 = iy + 1 < ny;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 tj = T[i], tr = vr0 ? T[i + k4] : z4, td = vd0 ? T[i + row] : z4;
    const float4 xj = x[i], xr = vr0 ? x[i + k4] : z4, xd = vd0 ? x[i + row] : z4;
#define DFM_TVN(c)                                                                             \
  {                                                                                            \
    const float gx = (vr0 && tj.c > 0.0f && tr.c > 0.0f) ? xr.c - xj.c : 0.0f;                 \
    const float gy = (vd0 && tj.c > 0.0f && td.c > 0.0f) ? xd.c - xj.c : 0.0f;                 \
    s[0] += (double)sqrtf(gx * gx + gy * gy);                                                  \
  }
    DFM_TVN(x) DFM_TVN(y) DFM_TVN(z) DFM_TVN(w)
#undef DFM_TVN
  }
  block_sum<1>(s, partial, gridDim.x);
}

// ---- total-variation block on a flat coefficient vector whose neighbours are given by index arrays (right, down,
// left, up; -1: none), the local basis's patches (spdhg_poisson.py:195-236 tv_*_local): the same arithmetic as the
// channel-fastest kernels above.
__global__ void __launch_bounds__(256) tv_dual_nbr_kernel(const float* __restrict__ x, const float* __restrict__ T,
                                                          float* __restrict__ qx, float* __restrict__ qy,
                                                          float* __restrict__ dqx, float* __restrict__ dqy,
                                                          const int* __restrict__ right, const int* __restrict__ down,
                                                          float lam, float ps, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int r = right[i], d = down[i];
    const bool vr0 = r >= 0, vd0 = d >= 0;
    const float ox = qx[i], oy = qy[i];
    float nx_, ny_;
    tv_core(T[i], vr0 ? T[r] : 0.0f, vd0 ? T[d] : 0.0f, vr0, vd0, x[i], vr0 ? x[r] : 0.0f, vd0 ? x[d] : 0.0f, ox, oy,
            lam, ps, nx_, ny_);
    qx[i] = nx_;
    qy[i] = ny_;
    dqx[i] = nx_ - ox;
    dqy[i] = ny_ - oy;
  }
}

__global__ void __launch_bounds__(256) tv_adjoint_update_nbr_kernel(const float* __restrict__ dqx,
                                                                    const float* __restrict__ dqy,
                                                                    const int* __restrict__ left,
                                                                    const int* __restrict__ up,
                                                                    float* __restrict__ x, float* __restrict__ z,
                                                                    const float* __restrict__ T, float inv_p,
                                                                    float lam, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float dz = -dqx[i] - dqy[i];
    const int l = left[i], u = up[i];
    if (l >= 0) dz += dqx[l];
    if (u >= 0) dz += dqy[u];
    const float zi = z[i] + dz;
    z[i] = zi;
    const float tt = T[i];
    x[i] = tt > 0.0f ? fmaxf(0.0f, x[i] - tt * ((zi + inv_p * dz) + lam)) : 0.0f;
  }
}

__global__ void __launch_bounds__(256) tv_value_nbr_kernel(const float* __restrict__ x, const float* __restrict__ T,
                                                           const int* __restrict__ right,
                                                           const int* __restrict__ down, int64_t n,
                                                           double* __restrict__ partial) {
  double s[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int r = right[i], d = down[i];
    const float tj = T[i], xj = x[i];
    const float gx = (r >= 0 && tj > 0.0f && T[r] > 0.0f) ? x[r] - xj : 0.0f;
    const float gy = (d >= 0 && tj > 0.0f && T[d] > 0.0f) ? x[d] - xj : 0.0f;
    s[0] += (double)sqrtf(gx * gx + gy * gy);
  }
  block_sum<1>(s, partial, gridDim.x);
}

// Smoothed isotropic TV of coefficient images (K, ny, nx) (a differentiable loss term, differentiable/ loss of the
// C4 refinement): value sum_{k,y,x} sqrt(dx^2 + dy^2 + eps) with forward differences (0 at the last column / row),
// and its gradient, in one pass: grad(y, x) = -(dx + dy)/s(y, x) + dx(y, x-1)/s(y, x-1) + dy(y-1, x)/s(y-1, x).
__global__ void __launch_bounds__(256) smooth_tv_kernel(const float* __restrict__ c, float* __restrict__ grad,
                                                        int nx, int ny, int64_t n, float eps,
                                                        double* __restrict__ partial) {
  double s[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % nx);
    const int y = (int)((i / nx) % ny);
    const float v = c[i];
    const bool r = x + 1 < nx, d = y + 1 < ny;
    const float dx = r ? c[i + 1] - v : 0.0f, dy = d ? c[i + nx] - v : 0.0f;
    const float sv = sqrtf(dx * dx + dy * dy + eps);
    s[0] += (double)sv;
    float g = -(dx + dy) / sv;
    if (x > 0) {
      const float l = c[i - 1];
      const float dxl = v - l, dyl = d ? c[i + nx - 1] - l : 0.0f;
      g += dxl / sqrtf(dxl * dxl + dyl * dyl + eps);
    }
    if (y > 0) {
      const float u = c[i - nx];
      const float dyu = v - u, dxu = r ? c[i - nx + 1] - u : 0.0f;
      g += dyu / sqrtf(dxu * dxu + dyu * dyu + eps);
    }
    grad[i] = g;
  }
  block_sum<1>(s, partial, gridDim.x);
}

// FISTA in the SQS metric (optimization/fista_poisson.py:61-282), over n data points or n coefficients (API layout
// (K, Ny, Nx), the support periodic over npix pixels; npix 0: no support).
// value (+ gradient in place of mu): mode 0 the deviance, mode 1 its quadratic approximation at the data
// (weights w / (g max(b, floor g)), halved squares)
__global__ void __launch_bounds__(256) fista_value_grad_kernel(float* __restrict__ mu, const float* __restrict__ b,
                                                               const float* __restrict__ w, int64_t pw,
                                                               const float* __restrict__ g, int64_t pg,
                                                               const float* __restrict__ bg, int64_t pbg, float eps,
                                                               int write_grad, int mode, float floor_counts,
                                                               int64_t n, double* __restrict__ partial) {
  double s[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float wi = pval(w, pw, i, 1.0f);
    float gr = 0.0f;
    if (wi != 0.0f) {
      const float gi = pval(g, pg, i, 1.0f);
      const double bi = (double)b[i];
      const double sc = (double)wi / (double)gi;
      if (mode == 0) {
        const double m = (double)fmaxf(mu[i], 0.0f) + (double)pval(bg, pbg, i, 0.0f) + (double)eps * gi;
        double v = m - bi;
        if (bi > 0.0) v += bi * log(bi / m);
        s[0] += sc * v;
        gr = (float)(sc * (1.0 - bi / m));
      } else {
        const double r = (double)mu[i] + (double)pval(bg, pbg, i, 0.0f) - bi;
        const double c = sc / fmax(bi, (double)floor_counts * gi);
        s[0] += 0.5 * c * r * r;
        gr = (float)(c * r);
      }
    }
    if (write_grad) mu[i] = gr;
  }
  block_sum<1>(s, partial, gridDim.x);
}

// u <- u c with c the Poisson curvature w / g max(b, m) / m^2, m = max(max(mu, 0) + bg, floor g) (mode 1: the
// start's w / (g max(b, floor g))); partial sums of c u^2
__global__ void __launch_bounds__(256) fista_curvature_kernel(float* __restrict__ u, const float* __restrict__ mu,
                                                              const float* __restrict__ b,
                                                              const float* __restrict__ w, int64_t pw,
                                                              const float* __restrict__ g, int64_t pg,
                                                              const float* __restrict__ bg, int64_t pbg,
                                                              float floor_counts, int mode, int64_t n,
                                                              double* __restrict__ partial) {
  double s[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float wi = pval(w, pw, i, 1.0f);
    const float gi = pval(g, pg, i, 1.0f);
    const float r = u[i];
    float c;
    if (mode == 0) {
      const float m = fmaxf(fmaxf(mu[i], 0.0f) + pval(bg, pbg, i, 0.0f), floor_counts * gi);
      c = wi != 0.0f ? wi / gi * fmaxf(b[i], m) / (m * m) : 0.0f;
    } else {
      c = wi != 0.0f ? wi / (gi * fmaxf(b[i], floor_counts * gi)) : 0.0f;
    }
    u[i] = r * c;
    s[0] += (double)c * (double)r * (double)r;
  }
  block_sum<1>(s, partial, gridDim.x);
}

// partial sums: what 0: x over the support; 1: w b and w x (two sums; x a data-sized prediction)
__global__ void __launch_bounds__(256) fista_sums_kernel(const float* __restrict__ x, const uint8_t* __restrict__ supp,
                                                         int64_t npix, const float* __restrict__ b,
                                                         const float* __restrict__ w, int64_t pw, int what,
                                                         int64_t n, double* __restrict__ partial) {
  double s[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (what == 0) {
      if (!npix || supp[i % npix]) s[0] += (double)x[i];
    } else {
      const float wi = pval(w, pw, i, 1.0f);
      s[0] += (double)wi * b[i];
      s[1] += (double)wi * x[i];
    }
  }
  block_sum<2>(s, partial, gridDim.x);
}

// element-wise: op 0 D = 1/d where d > 0 in the support, else 0; op 1 ones on the support; op 2 x = 0 outside the
// support or where x < 0; op 3 xn = max(0, y - th D (grad + lam)) where D > 0, else 0; op 4 out = a + beta (a - c)
__global__ void __launch_bounds__(256) fista_elementwise_kernel(int op, float* __restrict__ out,
                                                                const float* __restrict__ a,
                                                                const float* __restrict__ c,
                                                                const float* __restrict__ D,
                                                                const uint8_t* __restrict__ supp, int64_t npix,
                                                                float th, float lam, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const bool inside = npix ? supp[i % npix] != 0 : true;
    if (op == 0) {
      out[i] = (inside && a[i] > 0.0f) ? 1.0f / a[i] : 0.0f;
    } else if (op == 1) {
      out[i] = inside ? 1.0f : 0.0f;
    } else if (op == 2) {
      if (!inside || out[i] < 0.0f) out[i] = 0.0f;
    } else if (op == 3) {
      out[i] = D[i] > 0.0f ? fmaxf(0.0f, a[i] - th * D[i] * (c[i] + lam)) : 0.0f;
    } else {
      const float ai = a[i];
      out[i] = ai + th * (ai - c[i]);
    }
  }
}

// partial sums of <grad, xn - y>, sum (xn - y)^2 / D over D > 0, and sum xn
__global__ void __launch_bounds__(256) fista_step_sums_kernel(const float* __restrict__ xn,
                                                              const float* __restrict__ y,
                                                              const float* __restrict__ grad,
                                                              const float* __restrict__ D, int64_t n,
                                                              double* __restrict__ partial) {
  double s[3] = {0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double d = (double)xn[i] - (double)y[i];
    s[0] += (double)grad[i] * d;
    if (D[i] > 0.0f) s[1] += d * d / (double)D[i];
    s[2] += (double)xn[i];
  }
  block_sum<3>(s, partial, gridDim.x);
}

static unsigned reduce_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > DFM_MAX_PARTIALS) b = DFM_MAX_PARTIALS;
  return (unsigned)b;
}

// The EM pieces of the Poisson term (optimization/poisson_term.py:39-47 pt_ratio / pt_wg; osem_poisson.py:69-73
// osem_update), float32 arithmetic in the reference's order. A zero weight marks a point outside the fit.
//   ratio: out = w / g * b / mu, mu = max(pred, 0) + bg + eps g        weight over gain: out = w / g
__global__ void __launch_bounds__(256) em_ratio_kernel(float* __restrict__ out, const float* __restrict__ pred,
                                                       const float* __restrict__ b, const float* __restrict__ w,
                                                       int64_t pw, const float* __restrict__ g, int64_t pg,
                                                       const float* __restrict__ bg, int64_t pbg, float eps,
                                                       int64_t n, int ratio) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float wi = pval(w, pw, i, 1.0f);
    const float gi = pval(g, pg, i, 1.0f);
    if (!ratio) {
      out[i] = wi / gi;
      continue;
    }
    const float mu = fmaxf(pred[i], 0.0f) + pval(bg, pbg, i, 0.0f) + eps * gi;
    out[i] = wi == 0.0f ? 0.0f : wi / gi * b[i] / mu;
  }
}

// x = 0 outside the mask (periodic over pm points; pm 0: no mask), else x back / (sens + lam) where sens > 0
__global__ void __launch_bounds__(256) em_update_kernel(float* __restrict__ x, const float* __restrict__ back,
                                                        const float* __restrict__ sens,
                                                        const uint8_t* __restrict__ mask, int64_t pm, float lam,
                                                        int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const bool ok = pm ? mask[i % pm] != 0 : true;
    const float s = sens[i];
    x[i] = !ok ? 0.0f : (s > 0.0f ? x[i] * back[i] / (s + lam) : x[i]);
  }
}

// PDHG (pdhg_poisson.py:69-86 primal_update / average): x' = max(0, x - T (grad + lam)) where T > 0, else 0;
// x_bar = 2 x' - x; and A x_k = (A x_bar_k + A x_(k-1)) / 2
__global__ void __launch_bounds__(256) pdhg_primal_kernel(float* __restrict__ x, float* __restrict__ xb,
                                                          const float* __restrict__ grad,
                                                          const float* __restrict__ T, float lam, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float t = T[i];
    const float xo = x[i];
    const float xn = t > 0.0f ? fmaxf(0.0f, xo - t * (grad[i] + lam)) : 0.0f;
    xb[i] = 2.0f * xn - xo;
    x[i] = xn;
  }
}

__global__ void __launch_bounds__(256) average_kernel(float* __restrict__ ax, const float* __restrict__ axb,
                                                      int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    ax[i] = 0.5f * (axb[i] + ax[i]);
}

static dim3 pix_grid(int64_t npix, int kstride) {
  int64_t gy = npix < 32768 ? npix : 32768;
  return dim3((unsigned)((kstride + 255) / 256), (unsigned)gy);
}

// Scratch of the allocation-free deviance (dfm_deviance_noalloc): the per-workgroup float64 partial sums. One copy
// per GPU, held until the process ends; host code reaches it only through the address hipGetSymbolAddress gives for
// the current device (noalloc_scratch), never through this name.
static __device__ __align__(256) double scratch[DFM_MAX_PARTIALS];

// out[0] = partial[0] + partial[1] + ... + partial[n - 1], added in index order (one workgroup: the partials are
// staged in LDS, one thread adds them), so the sum equals a sequential host sum of the same partials bit for bit.
__global__ void __launch_bounds__(256) ordered_sum_kernel(const double* __restrict__ partial, int n,
                                                          double* __restrict__ out) {
  __shared__ double buf[DFM_MAX_PARTIALS];
  for (int i = threadIdx.x; i < n; i += blockDim.x) buf[i] = partial[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += buf[i];
    out[0] = s;
  }
}

// The device address of `scratch` on the current device, looked up once per device (each GPU holds its own copy
// at its own address).
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mu;
static double* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

static int noalloc_scratch(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    set_error("dfm_deviance_noalloc: no current device (%s)", hipGetErrorString(e));
    return DFM_ERR_LAUNCH;
  }
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (noalloc_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error("dfm_deviance_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                hipGetErrorString(e));
      return DFM_ERR_LAUNCH;
    }
    noalloc_ptr[dev] = static_cast<double*>(p);
  }
  *out = noalloc_ptr[dev];
  return DFM_OK;
}

}  // namespace dfm

using namespace dfm;

extern "C" {

int dfm_dual_update_sub(float* y, float* d, const float* sig, const float* b, const float* w, int64_t pw,
                        const float* g, int64_t pg, const float* bg, int64_t pbg, float eps, const int32_t* omegas,
                        int32_t n_sub, int64_t row_len, dfm_stream_t stream) {
  DFM_CHECK_ARG(y && d && sig && b && omegas, "dfm_dual_update_sub: null pointer");
  DFM_CHECK_ARG((pw == 0 || w) && (pg == 0 || g) && (pbg == 0 || bg) && pw >= 0 && pg >= 0 && pbg >= 0,
                "dfm_dual_update_sub: periodic array missing");
  DFM_CHECK_ARG(n_sub >= 0 && row_len > 0, "dfm_dual_update_sub: bad sizes");
  const int64_t n = (int64_t)n_sub * row_len;
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(dual_update_sub_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, y, d,
                     sig, b, w, pw, g, pg, bg, pbg, eps, omegas, row_len, n);
  DFM_LAUNCH_CHECK("dfm_dual_update_sub");
  return DFM_OK;
}

int dfm_spdhg_update(float* x, float* z, const float* T, const float* dz, float inv_p, float lam, int64_t n,
                     dfm_stream_t stream) {
  DFM_CHECK_ARG(x && z && T && dz && n >= 0, "dfm_spdhg_update: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(spdhg_update_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, x, z, T,
                     dz, inv_p, lam, n);
  DFM_LAUNCH_CHECK("dfm_spdhg_update");
  return DFM_OK;
}

int dfm_tinit(float* T, int64_t npix, int32_t kstride, dfm_stream_t stream) {
  DFM_CHECK_ARG(T && npix > 0 && npix < (1 << 30) && kstride > 0, "dfm_tinit: bad arguments");
  hipLaunchKernelGGL(tinit_kernel, pix_grid(npix, kstride), dim3(256), 0, (hipStream_t)stream, T, npix, kstride);
  DFM_LAUNCH_CHECK("dfm_tinit");
  return DFM_OK;
}

int dfm_tmin(float* T, const float* colsum, const uint8_t* supp, int64_t npix, int32_t kstride, int32_t kvalid,
             float c, dfm_stream_t stream) {
  DFM_CHECK_ARG(T && colsum && npix > 0 && npix < (1 << 30) && kstride > 0 && kvalid <= kstride,
                "dfm_tmin: bad arguments");
  hipLaunchKernelGGL(tmin_kernel, pix_grid(npix, kstride), dim3(256), 0, (hipStream_t)stream, T, colsum, supp, npix,
                     kstride, kvalid, c);
  DFM_LAUNCH_CHECK("dfm_tmin");
  return DFM_OK;
}

int dfm_tfinal(float* T, int64_t npix, int32_t kstride, int32_t kvalid, dfm_stream_t stream) {
  DFM_CHECK_ARG(T && npix > 0 && npix < (1 << 30) && kstride > 0 && kvalid <= kstride, "dfm_tfinal: bad arguments");
  hipLaunchKernelGGL(tfinal_kernel, pix_grid(npix, kstride), dim3(256), 0, (hipStream_t)stream, T, npix, kstride, kvalid);
  DFM_LAUNCH_CHECK("dfm_tfinal");
  return DFM_OK;
}

int dfm_dual_steps(float* sig, const float* rowsum, const float* w, int64_t pw, float sigma0, int64_t n,
                   dfm_stream_t stream) {
  DFM_CHECK_ARG(sig && rowsum && (pw == 0 || w) && n >= 0, "dfm_dual_steps: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(dual_steps_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, sig,
                     rowsum, w, pw, sigma0, n);
  DFM_LAUNCH_CHECK("dfm_dual_steps");
  return DFM_OK;
}

int dfm_set_mask_ones(float* u, const float* w, int64_t pw, int64_t n, dfm_stream_t stream) {
  DFM_CHECK_ARG(u && (pw == 0 || w) && n >= 0, "dfm_set_mask_ones: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(set_mask_ones_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, u, w,
                     pw, n);
  DFM_LAUNCH_CHECK("dfm_set_mask_ones");
  return DFM_OK;
}

int dfm_support_ones_k(float* x, const uint8_t* supp, int64_t npix, int32_t kstride, int32_t kvalid,
                       dfm_stream_t stream) {
  DFM_CHECK_ARG(x && npix > 0 && npix < (1 << 30) && kstride > 0 && kvalid <= kstride,
                "dfm_support_ones_k: bad arguments");
  hipLaunchKernelGGL(support_ones_kernel, pix_grid(npix, kstride), dim3(256), 0, (hipStream_t)stream, x, supp, npix,
                     kstride, kvalid);
  DFM_LAUNCH_CHECK("dfm_support_ones_k");
  return DFM_OK;
}

int dfm_project_support_k(float* x, const uint8_t* supp, int64_t npix, int32_t kstride, int32_t kvalid,
                          dfm_stream_t stream) {
  DFM_CHECK_ARG(x && npix > 0 && npix < (1 << 30) && kstride > 0 && kvalid <= kstride,
                "dfm_project_support_k: bad arguments");
  hipLaunchKernelGGL(project_support_kernel, pix_grid(npix, kstride), dim3(256), 0, (hipStream_t)stream, x, supp,
                     npix, kstride, kvalid);
  DFM_LAUNCH_CHECK("dfm_project_support_k");
  return DFM_OK;
}

int dfm_deviance_partial(const float* mu, const float* b, const float* w, int64_t pw, const float* g, int64_t pg,
                         const float* bg, int64_t pbg, float eps, int64_t n, double* partial, int32_t* n_partial,
                         dfm_stream_t stream) {
  DFM_CHECK_ARG(mu && b && partial && n_partial && (pw == 0 || w) && (pg == 0 || g) && (pbg == 0 || bg) && n >= 0,
                "dfm_deviance_partial: bad arguments");
  const unsigned nb = reduce_blocks(n);
  *n_partial = (int32_t)nb;
  hipLaunchKernelGGL(deviance_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, mu, b, w, pw, g, pg, bg, pbg, eps,
                     n, partial);
  DFM_LAUNCH_CHECK("dfm_deviance_partial");
  return DFM_OK;
}

int dfm_deviance_noalloc(const float* mu, const float* b, const float* w, int64_t pw, const float* g, int64_t pg,
                         const float* bg, int64_t pbg, float eps, int64_t n, double* out, dfm_stream_t stream) {
  DFM_CHECK_ARG(mu && b && out && (pw == 0 || w) && (pg == 0 || g) && (pbg == 0 || bg) && n >= 0,
                "dfm_deviance_noalloc: bad arguments");
  const unsigned nb = reduce_blocks(n);
  DFM_CHECK_ARG(nb <= (unsigned)DFM_MAX_PARTIALS, "dfm_deviance_noalloc: %u partial sums exceed the scratch of %d",
                nb, DFM_MAX_PARTIALS);
  double* part = nullptr;
  const int rc = noalloc_scratch(&part);
  if (rc != DFM_OK) return rc;
  // every partial 0 .. nb - 1 is written by deviance_kernel before ordered_sum_kernel reads it (same stream), so
  // nothing left in the scratch by an earlier call is read
  hipLaunchKernelGGL(deviance_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, mu, b, w, pw, g, pg, bg, pbg, eps,
                     n, part);
  DFM_LAUNCH_CHECK("dfm_deviance_noalloc");
  hipLaunchKernelGGL(ordered_sum_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, part, (int)nb, out);
  DFM_LAUNCH_CHECK("dfm_deviance_noalloc");
  return DFM_OK;
}

int dfm_mean_counted_partial(const float* b, const float* w, int64_t pw, int64_t n, double* partial,
                             int32_t* n_partial, dfm_stream_t stream) {
  DFM_CHECK_ARG(b && partial && n_partial && (pw == 0 || w) && n >= 0, "dfm_mean_counted_partial: bad arguments");
  const unsigned nb = reduce_blocks(n);
  *n_partial = (int32_t)nb;
  hipLaunchKernelGGL(mean_counted_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, b, w, pw, n, partial);
  DFM_LAUNCH_CHECK("dfm_mean_counted_partial");
  return DFM_OK;
}

int dfm_poisson_em_ratio(const float* pred, const float* b, const float* w, int64_t pw, const float* g, int64_t pg,
                         const float* bg, int64_t pbg, float eps, int64_t n, float* out, dfm_stream_t stream) {
  DFM_CHECK_ARG(pred && b && out && (pw == 0 || w) && (pg == 0 || g) && (pbg == 0 || bg) && n >= 0,
                "dfm_poisson_em_ratio: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(em_ratio_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, out, pred,
                     b, w, pw, g, pg, bg, pbg, eps, n, 1);
  DFM_LAUNCH_CHECK("dfm_poisson_em_ratio");
  return DFM_OK;
}

int dfm_poisson_weight_over_gain(const float* w, int64_t pw, const float* g, int64_t pg, int64_t n, float* out,
                                 dfm_stream_t stream) {
  DFM_CHECK_ARG(out && (pw == 0 || w) && (pg == 0 || g) && n >= 0, "dfm_poisson_weight_over_gain: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(em_ratio_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, out,
                     nullptr, nullptr, w, pw, g, pg, nullptr, 0, 0.0f, n, 0);
  DFM_LAUNCH_CHECK("dfm_poisson_weight_over_gain");
  return DFM_OK;
}

int dfm_em_update(float* x, const float* back, const float* sens, const uint8_t* mask, int64_t pm, float lam,
                  int64_t n, dfm_stream_t stream) {
  DFM_CHECK_ARG(x && back && sens && (pm == 0 || mask) && pm >= 0 && n >= 0 && lam >= 0.0f,
                "dfm_em_update: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(em_update_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, x, back,
                     sens, mask, pm, lam, n);
  DFM_LAUNCH_CHECK("dfm_em_update");
  return DFM_OK;
}

int dfm_pdhg_primal_update(float* x, float* xb, const float* grad, const float* T, float lam, int64_t n,
                           dfm_stream_t stream) {
  DFM_CHECK_ARG(x && xb && grad && T && n >= 0, "dfm_pdhg_primal_update: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(pdhg_primal_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, x, xb,
                     grad, T, lam, n);
  DFM_LAUNCH_CHECK("dfm_pdhg_primal_update");
  return DFM_OK;
}

int dfm_average(float* ax, const float* axb, int64_t n, dfm_stream_t stream) {
  DFM_CHECK_ARG(ax && axb && n >= 0, "dfm_average: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(average_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, ax, axb, n);
  DFM_LAUNCH_CHECK("dfm_average");
  return DFM_OK;
}

int dfm_fista_value_grad_partial(float* mu, const float* b, const float* w, int64_t pw, const float* g, int64_t pg,
                                 const float* bg, int64_t pbg, float eps, int32_t write_grad, int32_t mode,
                                 float floor_counts, int64_t n, double* partial, int32_t* n_partial,
                                 dfm_stream_t stream) {
  DFM_CHECK_ARG(mu && b && partial && n_partial && (pw == 0 || w) && (pg == 0 || g) && (pbg == 0 || bg) && n >= 0 &&
                    (mode == 0 || mode == 1),
                "dfm_fista_value_grad_partial: bad arguments");
  const unsigned nb = reduce_blocks(n);
  *n_partial = (int32_t)nb;
  hipLaunchKernelGGL(fista_value_grad_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, mu, b, w, pw, g, pg, bg,
                     pbg, eps, write_grad, mode, floor_counts, n, partial);
  DFM_LAUNCH_CHECK("dfm_fista_value_grad_partial");
  return DFM_OK;
}

int dfm_fista_curvature_partial(float* u, const float* mu, const float* b, const float* w, int64_t pw, const float* g,
                                int64_t pg, const float* bg, int64_t pbg, float floor_counts, int32_t mode, int64_t n,
                                double* partial, int32_t* n_partial, dfm_stream_t stream) {
  DFM_CHECK_ARG(u && mu && b && partial && n_partial && (pw == 0 || w) && (pg == 0 || g) && (pbg == 0 || bg) &&
                    n >= 0 && (mode == 0 || mode == 1),
                "dfm_fista_curvature_partial: bad arguments");
  const unsigned nb = reduce_blocks(n);
  *n_partial = (int32_t)nb;
  hipLaunchKernelGGL(fista_curvature_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, u, mu, b, w, pw, g, pg, bg,
                     pbg, floor_counts, mode, n, partial);
  DFM_LAUNCH_CHECK("dfm_fista_curvature_partial");
  return DFM_OK;
}

int dfm_fista_sums_partial(const float* x, const uint8_t* supp, int64_t npix, const float* b, const float* w,
                           int64_t pw, int32_t what, int64_t n, double* partial, int32_t* n_partial,
                           dfm_stream_t stream) {
  DFM_CHECK_ARG(x && partial && n_partial && (npix == 0 || supp) && npix >= 0 && n >= 0 &&
                    (what == 0 || (what == 1 && b && (pw == 0 || w))),
                "dfm_fista_sums_partial: bad arguments");
  const unsigned nb = reduce_blocks(n);
  *n_partial = (int32_t)nb;
  hipLaunchKernelGGL(fista_sums_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, x, supp, npix, b, w, pw, what,
                     n, partial);
  DFM_LAUNCH_CHECK("dfm_fista_sums_partial");
  return DFM_OK;
}

int dfm_fista_elementwise(int32_t op, float* out, const float* a, const float* c, const float* D,
                          const uint8_t* supp, int64_t npix, float th, float lam, int64_t n, dfm_stream_t stream) {
  DFM_CHECK_ARG(out && op >= 0 && op <= 4 && (npix == 0 || supp) && npix >= 0 && n >= 0 &&
                    (op != 0 || a) && (op != 3 || (a && c && D)) && (op != 4 || (a && c)),
                "dfm_fista_elementwise: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(fista_elementwise_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     op, out, a, c, D, supp, npix, th, lam, n);
  DFM_LAUNCH_CHECK("dfm_fista_elementwise");
  return DFM_OK;
}

int dfm_fista_step_sums_partial(const float* xn, const float* y, const float* grad, const float* D, int64_t n,
                                double* partial, int32_t* n_partial, dfm_stream_t stream) {
  DFM_CHECK_ARG(xn && y && grad && D && partial && n_partial && n >= 0, "dfm_fista_step_sums_partial: bad arguments");
  const unsigned nb = reduce_blocks(n);
  *n_partial = (int32_t)nb;
  hipLaunchKernelGGL(fista_step_sums_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, xn, y, grad, D, n, partial);
  DFM_LAUNCH_CHECK("dfm_fista_step_sums_partial");
  return DFM_OK;
}

int dfm_dot_partial(const float* a, const float* b, int64_t n, double* partial, int32_t* n_partial,
                    dfm_stream_t stream) {
  DFM_CHECK_ARG(a && b && partial && n_partial && n >= 0, "dfm_dot_partial: bad arguments");
  const unsigned nb = reduce_blocks(n);
  *n_partial = (int32_t)nb;
  hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, a, b, n, partial);
  DFM_LAUNCH_CHECK("dfm_dot_partial");
  return DFM_OK;
}

int dfm_tv_dual_k(const float* x, const float* T, float* qx, float* qy, float* dqx, float* dqy, int32_t nx,
                  int32_t ny, int32_t kstride, float lam, float ps, dfm_stream_t stream) {
  DFM_CHECK_ARG(x && T && qx && qy && dqx && dqy, "dfm_tv_dual_k: null pointer");
  DFM_CHECK_ARG(nx > 0 && ny > 0 && kstride > 0 && kstride % 4 == 0, "dfm_tv_dual_k: bad sizes");
  const int64_t n = (int64_t)nx * ny * (kstride / 4);
  hipLaunchKernelGGL(tv_dual_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     (const float4*)x, (const float4*)T, (float4*)qx, (float4*)qy, (float4*)dqx, (float4*)dqy, nx, ny,
                     kstride / 4, lam, ps);
  DFM_LAUNCH_CHECK("dfm_tv_dual_k");
  return DFM_OK;
}

int dfm_tv_adjoint_update_k(const float* dqx, const float* dqy, float* x, float* z, const float* T, int32_t nx,
                            int32_t ny, int32_t kstride, float inv_p, float lam, dfm_stream_t stream) {
  DFM_CHECK_ARG(x && T && z && dqx && dqy, "dfm_tv_adjoint_update_k: null pointer");
  DFM_CHECK_ARG(nx > 0 && ny > 0 && kstride > 0 && kstride % 4 == 0, "dfm_tv_adjoint_update_k: bad sizes");
  const int64_t n = (int64_t)nx * ny * (kstride / 4);
  hipLaunchKernelGGL(tv_adjoint_update_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     (const float4*)dqx, (const float4*)dqy, (float4*)x, (float4*)z, (const float4*)T, nx, ny,
                     kstride / 4, inv_p, lam);
  DFM_LAUNCH_CHECK("dfm_tv_adjoint_update_k");
  return DFM_OK;
}

int dfm_tv_dual_nbr(const float* x, const float* T, float* qx, float* qy, float* dqx, float* dqy, const int32_t* right,
                    const int32_t* down, float lam, float ps, int64_t n, dfm_stream_t stream) {
  DFM_CHECK_ARG(x && T && qx && qy && dqx && dqy && right && down && n >= 0, "dfm_tv_dual_nbr: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(tv_dual_nbr_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, x, T,
                     qx, qy, dqx, dqy, right, down, lam, ps, n);
  DFM_LAUNCH_CHECK("dfm_tv_dual_nbr");
  return DFM_OK;
}

int dfm_tv_adjoint_update_nbr(const float* dqx, const float* dqy, const int32_t* left, const int32_t* up, float* x,
                              float* z, const float* T, float inv_p, float lam, int64_t n, dfm_stream_t stream) {
  DFM_CHECK_ARG(dqx && dqy && left && up && x && z && T && n >= 0, "dfm_tv_adjoint_update_nbr: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(tv_adjoint_update_nbr_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, dqx, dqy, left, up, x, z, T, inv_p, lam, n);
  DFM_LAUNCH_CHECK("dfm_tv_adjoint_update_nbr");
  return DFM_OK;
}

int dfm_tv_value_nbr_partial(const float* x, const float* T, const int32_t* right, const int32_t* down, int64_t n,
                             double* partial, int32_t* n_partial, dfm_stream_t stream) {
  DFM_CHECK_ARG(x && T && right && down && partial && n_partial && n > 0, "dfm_tv_value_nbr_partial: bad arguments");
  const unsigned nb = reduce_blocks(n);
  *n_partial = (int32_t)nb;
  hipLaunchKernelGGL(tv_value_nbr_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, x, T, right, down, n, partial);
  DFM_LAUNCH_CHECK("dfm_tv_value_nbr_partial");
  return DFM_OK;
}

int dfm_tv_value_partial(const float* x, const float* T, int32_t nx, int32_t ny, int32_t kstride, double* partial,
                         int32_t* n_partial, dfm_stream_t stream) {
  DFM_CHECK_ARG(x && T && partial && n_partial, "dfm_tv_value_partial: null pointer");
  DFM_CHECK_ARG(nx > 0 && ny > 0 && kstride > 0 && kstride % 4 == 0, "dfm_tv_value_partial: bad sizes");
  const int64_t n = (int64_t)nx * ny * (kstride / 4);
  const unsigned nb = reduce_blocks(n);
  *n_partial = (int32_t)nb;
  hipLaunchKernelGGL(tv_value_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, (const float4*)x,
                     (const float4*)T, nx, ny, kstride / 4, partial);
  DFM_LAUNCH_CHECK("dfm_tv_value_partial");
  return DFM_OK;
}

int dfm_smooth_tv(const float* c, float* grad, int32_t K, int32_t ny, int32_t nx, float eps, double* partial,
                  int32_t* n_partial, dfm_stream_t stream) {
  DFM_CHECK_ARG(c && grad && partial && n_partial, "dfm_smooth_tv: null pointer");
  DFM_CHECK_ARG(K >= 0 && ny > 0 && nx > 0 && eps >= 0.0f, "dfm_smooth_tv: bad sizes");
  const int64_t n = (int64_t)K * ny * nx;
  const unsigned nb = reduce_blocks(n);
  *n_partial = (int32_t)nb;
  hipLaunchKernelGGL(smooth_tv_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, c, grad, nx, ny, n, eps, partial);
  DFM_LAUNCH_CHECK("dfm_smooth_tv");
  return DFM_OK;
}

int dfm_scale(float* a, float s, int64_t n, dfm_stream_t stream) {
  DFM_CHECK_ARG(a && n >= 0, "dfm_scale: bad arguments");
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(scale_kernel, dim3(grid_stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, a, s, n);
  DFM_LAUNCH_CHECK("dfm_scale");
  return DFM_OK;
}

}  // extern "C"
