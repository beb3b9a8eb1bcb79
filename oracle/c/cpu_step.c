/* FP32 C/OpenMP restatement of one SPDHG-Poisson subset step (test infrastructure: the host-core baseline of
 * bench.py, never the product). It follows the same algorithm as the GPU path and the reference:
 *   dual update on the rows of subset m          (optimization/spdhg_poisson.py:69-91, float64 arithmetic)
 *   CSR adjoint PF^T y  -> sinogram              (operators/pf_kernels.cl:644-687)
 *   backprojection + primal update               (radon_kernels.cl:106-141, spdhg_poisson.py:95-106)
 *   forward projection of the next subset        (radon_kernels.cl:44-101)
 *   CSR forward PF sino -> data rows             (operators/pf_kernels.cl:597-641)
 * Layouts: images (npix, K) and sinograms (n_angles, ns, K), orientations fastest; data (R, ns, cp).
 * Geometry records per angle: X, Y, T0, wscale, hw, top, 0, 0; K(d) = clamp(hw - |d|, 0, top). */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  float X, Y, T0, ws, hw, top, p0, p1;
} geo_t;

void cpu_set_threads(int n) { omp_set_num_threads(n); }
int cpu_get_threads(void) { return omp_get_max_threads(); }

static inline float kern(float hw, float top, float d) {
  float w = hw - fabsf(d);
  if (w < 0.0f) w = 0.0f;
  if (w > top) w = top;
  return w;
}

void cpu_radon_forward(const float* img, float* sino, const geo_t* geo, int nx, int ny, int ns, int na, int K,
                       float scale) {
#pragma omp parallel for schedule(dynamic)
  for (int a = 0; a < na; ++a) {
    const geo_t g = geo[a];
    float* S = sino + (int64_t)a * ns * K;
    memset(S, 0, sizeof(float) * (size_t)ns * K);
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const float t = g.X * x + g.Y * y + g.T0;
        const int sf = (int)floorf(t - g.hw) + 1;
        const float* src = img + ((int64_t)y * nx + x) * K;
        for (int j = 0; j < 4; ++j) {
          const int s = sf + j;
          if (s < 0 || s >= ns) continue;
          const float w = kern(g.hw, g.top, t - (float)s);
          if (w == 0.0f) continue;
          float* dst = S + (int64_t)s * K;
          for (int k = 0; k < K; ++k) dst[k] += w * src[k];
        }
      }
    const float f = scale * g.ws;
    for (int64_t i = 0; i < (int64_t)ns * K; ++i) S[i] *= f;
  }
}

void cpu_radon_backward_update(const float* sino, const geo_t* geo, int nx, int ny, int ns, int na, int K, float* x,
                               float* z, const float* T, float inv_p, float lam) {
#pragma omp parallel
  {
    float* acc = (float*)malloc(sizeof(float) * (size_t)K);
#pragma omp for schedule(static)
    for (int64_t p = 0; p < (int64_t)nx * ny; ++p) {
      const int px = (int)(p % nx), py = (int)(p / nx);
      memset(acc, 0, sizeof(float) * (size_t)K);
      for (int a = 0; a < na; ++a) {
        const geo_t g = geo[a];
        const float t = g.X * px + g.Y * py + g.T0;
        const int sf = (int)floorf(t - g.hw) + 1;
        for (int j = 0; j < 4; ++j) {
          const int s = sf + j;
          if (s < 0 || s >= ns) continue;
          const float w = g.ws * kern(g.hw, g.top, t - (float)s);
          if (w == 0.0f) continue;
          const float* src = sino + ((int64_t)a * ns + s) * K;
          for (int k = 0; k < K; ++k) acc[k] += w * src[k];
        }
      }
      float* xp = x + p * K;
      float* zp = z + p * K;
      const float* tp = T + p * K;
      for (int k = 0; k < K; ++k) {
        const float zn = zp[k] + acc[k];
        zp[k] = zn;
        xp[k] = tp[k] > 0.0f ? fmaxf(0.0f, xp[k] - tp[k] * ((zn + inv_p * acc[k]) + lam)) : 0.0f;
      }
    }
    free(acc);
  }
}

/* forward CSR rows (r, j) list k; data[omegas[a], y, j] = sum_k val sino[a, y, k] */
void cpu_spmm_forward(const float* sino, const int32_t* row_ptr, const uint16_t* col_k, const float* val,
                      float* data, const int32_t* omegas, int na, int ns, int cp, int K) {
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int a = 0; a < na; ++a)
    for (int y = 0; y < ns; ++y) {
      const int r = omegas[a];
      const float* S = sino + ((int64_t)a * ns + y) * K;
      float* D = data + ((int64_t)r * ns + y) * cp;
      for (int j = 0; j < cp; ++j) {
        float acc = 0.0f;
        for (int i = row_ptr[(int64_t)r * cp + j]; i < row_ptr[(int64_t)r * cp + j + 1]; ++i) acc += val[i] * S[col_k[i]];
        D[j] = acc;
      }
    }
}

/* adjoint CSR rows (r, k), k < kb, list j; sino[a, y, k] = alpha sum_j val data[omegas[a], y, j] for k < kb and 0
 * for kb <= k < K (K the sinogram row stride) */
void cpu_spmm_adjoint(const float* data, const int32_t* row_ptr, const uint16_t* col_j, const float* val,
                      float* sino, const int32_t* omegas, int na, int ns, int cp, int K, int kb, float alpha) {
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int a = 0; a < na; ++a)
    for (int y = 0; y < ns; ++y) {
      const int r = omegas[a];
      const float* D = data + ((int64_t)r * ns + y) * cp;
      float* S = sino + ((int64_t)a * ns + y) * K;
      for (int k = 0; k < kb; ++k) {
        float acc = 0.0f;
        for (int i = row_ptr[(int64_t)r * kb + k]; i < row_ptr[(int64_t)r * kb + k + 1]; ++i)
          acc += val[i] * D[col_j[i]];
        S[k] = alpha * acc;
      }
      for (int k = kb; k < K; ++k) S[k] = 0.0f;
    }
}

/* y' = prox(y + s Ax) on the rows of a subset; d: A x in, y' - y out (no weights/gain/background) */
void cpu_dual_update(float* y, float* d, const float* sig, const float* b, float eps, const int32_t* omegas, int na,
                     int64_t row_len) {
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)na * row_len; ++t) {
    const int64_t s = t / row_len;
    const int64_t i = (int64_t)omegas[s] * row_len + (t - s * row_len);
    const float si = sig[i];
    float yn = 0.0f;
    if (si != 0.0f) {
      const double sc = 1.0, r = (double)eps;
      const double v = (double)y[i] + (double)si * d[i] + (double)si * r;
      const double dd = v - sc;
      yn = (float)(0.5 * (v + sc - sqrt(dd * dd + 4.0 * (double)si * sc * (double)b[i])));
    }
    d[i] = yn - y[i];
    y[i] = yn;
  }
}
