// Candidate segments of the pole figure: for one rotation angle and 32 orientations, the eta bins of every ring in
// which an entry can be non-zero (shared by the PF generation, pf_gen.hip, and the PF parameter gradients,
// pf_grad.hip).
//
// For orientation k, pole axis a and omega node s the pole direction n = U_k h_a is taken into the detector frame of
// the node (x, y, z = n . R^T d0, n . R^T d90, n . R^T p0); the probed directions of ring p form the cone
// v(eta) = cos(th_p) (cos(eta) R^T d0 + sin(eta) R^T d90) - sin(th_p) R^T p0, so n . v = A cos(eta - phi) - B with
// A = cos(th_p) sqrt(x^2 + y^2), B = sin(th_p) z, phi = atan2(y, x). |n . v| > tau = 1 - cut sigma^2 - 1e-4 (a
// superset of the entries whose exponent (1 - |n . v|) / sigma^2 stays below the cut) holds on the eta windows
// phi +- acos((tau + B) / A) (for n) and phi + pi +- acos((tau - B) / A) (for -n); every eta bin with a quadrature
// node in a window is a candidate of ring p.
#pragma once
#include "dfm_common.h"

namespace dfm {

struct CandGeo {
  const double* poles;       // (K, A, 3)
  const double* frames;      // (R, S, 9): R^T d0, R^T d90, R^T p0 of every omega node
  const double* sin_th;      // (P,)
  const double* cos_th;      // (P,)
  const int* axis_start;     // (P + 1,)
  const double* inv_sigma2;  // (K,)
  int k0, kb, C, P, A, S;
  double cut, eta0, deta, xlo, xhi;   // bin c has its nodes at eta0 + (c + x) deta, xlo <= x <= xhi
};

// Sets bit j % 32 of word (j / 32) * 32 + kl of bm for every candidate segment j of orientation kbi * 32 + kl at
// angle r (the caller clears bm and synchronises after). Threads tid = 0 .. nth - 1 share the (orientation, axis,
// node) items.
__device__ inline void mark_candidates(const CandGeo& g, int r, int kbi, unsigned* bm, int tid, int nth) {
  const int items = 32 * g.A * g.S;
  for (int i = tid; i < items; i += nth) {
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
        int m_lo = (int)ceil((cc - w - g.eta0) / g.deta - g.xhi);
        int m_hi = (int)floor((cc + w - g.eta0) / g.deta - g.xlo);
        m_lo = max(m_lo, 0);
        m_hi = min(m_hi, g.C - 1);
        for (int c = m_lo; c <= m_hi; ++c) {
          const int j = c * g.P + p;
          atomicOr(&bm[(j >> 5) * 32 + kl], 1u << (j & 31));
        }
      }
    }
  }
}

}  // namespace dfm
