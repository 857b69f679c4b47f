// SU(3) exponential shared by the link update and the smearing kernels.
#pragma once
#include "staples.h"

namespace sfh {

// exp(iQ) for hermitian traceless Q by the Cayley-Hamilton closed form of
// Morningstar & Peardon (hep-lat/0311018), with the guards and small-Q branch
// of Fermions.cpp:843-902.  Returns a status code.
static __device__ int exp_iq(const M& Q, M& out) {
  const M Q2 = mmul(Q, Q);
  const M Q3 = mmul(Q2, Q);
  const C t3 = mtrace(Q3);
  if (fabs(t3.im) > 1e-10) return SFH_E_EXP_QCUBED;
  double c0 = t3.re / 3.0;
  const double c1 = mtrace(Q2).re / 2.0;
  if (c1 < 0.0) return SFH_E_EXP_NEG_TRACE;
  if (c1 < 1e-12) {
    M r = meye<double>();
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      r.e[k].re += -Q.e[k].im - 0.5 * Q2.e[k].re;
      r.e[k].im += Q.e[k].re - 0.5 * Q2.e[k].im;
    }
    out = r;
    return SFH_OK;
  }
  const bool neg = c0 < 0.0;
  c0 = fabs(c0);
  const double c0max = 2.0 * sqrt(c1 * c1 * c1 / 27.0);
  double ratio = c0max > 0.0 ? c0 / c0max : 1.0;
  ratio = fmin(fmax(ratio, -1.0), 1.0);
  const double theta = acos(ratio);
  const double w = sqrt(c1) * sin(theta / 3.0);
  const double u = sqrt(c1 / 3.0) * cos(theta / 3.0);
  double s2u, c2u, su, cu;
  sincos(2.0 * u, &s2u, &c2u);
  sincos(u, &su, &cu);
  const C e2u = {c2u, s2u}, emu = {cu, -su};
  double xi;
  if (fabs(w) <= 0.05) {
    const double w2 = w * w;
    xi = 1.0 - (w2 / 6.0) * (1.0 - (w2 / 20.0) * (1.0 - w2 / 42.0));
  } else {
    xi = sin(w) / w;
  }
  const double cw = cos(w);
  const C h0 = (u * u - w * w) * e2u + emu * C{8.0 * u * u * cw, 2.0 * u * (3.0 * u * u + w * w) * xi};
  const C h1 = (2.0 * u) * e2u - emu * C{2.0 * u * cw, -(3.0 * u * u - w * w) * xi};
  const C h2 = e2u - emu * C{cw, 3.0 * u * xi};
  const double den = 9.0 * u * u - w * w;
  C f0 = {h0.re / den, h0.im / den}, f1 = {h1.re / den, h1.im / den}, f2 = {h2.re / den, h2.im / den};
  if (neg) {
    f0 = conj(f0);
    f1 = {-f1.re, f1.im};
    f2 = conj(f2);
  }
  M r;
#pragma unroll
  for (int k = 0; k < 9; ++k) r.e[k] = f1 * Q.e[k] + f2 * Q2.e[k];
  r.e[0] = r.e[0] + f0;
  r.e[4] = r.e[4] + f0;
  r.e[8] = r.e[8] + f0;
  out = r;
  return SFH_OK;
}

}  // namespace sfh
