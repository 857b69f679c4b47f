// Seeded, counter-based host generators for links, spinors and momenta.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11) with Box-Muller gives N(0,1) draws addressed by
// (stream, element index, draw), so a field is the same whatever the thread
// count or machine.  Distributions follow the reference:
//   spinors  complex Gaussian, sigma per real part    (Fermions.cpp:207-212)
//   momenta  P = sum_a p_a lambda_a / 2, p_a ~ N(0,1)  (Fermions.cpp:342-360)
//   links    U = exp(i eps H), H distributed as P      (ai_check.cpp:42-63)
// The exponential is the Cayley-Hamilton closed form of Morningstar &
// Peardon (hep-lat/0311018), the same method the reference uses
// (Fermions.cpp:843-902).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <thread>
#include <vector>

namespace sfh {

namespace {

using cd = std::complex<double>;

struct Philox {
  uint32_t k0, k1;
  explicit Philox(unsigned long long seed) : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)) {}
  void operator()(uint32_t c[4]) const {
    uint32_t a = k0, b = k1;
    for (int r = 0; r < 10; ++r) {
      const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
      const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
      const uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0;
      const uint32_t h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
      const uint32_t n0 = h1 ^ c[1] ^ a, n2 = h0 ^ c[3] ^ b;
      c[0] = n0;
      c[1] = l1;
      c[2] = n2;
      c[3] = l0;
      a += 0x9E3779B9u;
      b += 0xBB67AE85u;
    }
  }
};

// 8 normals for element idx of a stream.
void normals8(const Philox& ph, uint32_t stream, uint64_t idx, double out[8]) {
  for (int d = 0; d < 2; ++d) {
    uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), stream, (uint32_t)d};
    ph(c);
    for (int k = 0; k < 2; ++k) {
      // uniforms in (0, 1]
      const double u1 = ((double)c[2 * k] + 1.0) * (1.0 / 4294967296.0);
      const double u2 = ((double)c[2 * k + 1] + 0.5) * (1.0 / 4294967296.0);
      const double rad = std::sqrt(-2.0 * std::log(u1));
      const double ang = 6.283185307179586476925286766559 * u2;
      out[4 * d + 2 * k] = rad * std::cos(ang);
      out[4 * d + 2 * k + 1] = rad * std::sin(ang);
    }
  }
}

// H = sum_a r_a lambda_a / 2 (Gell-Mann basis, Fermions.cpp:321-340).
void algebra(const double r[8], cd H[9]) {
  const double s3 = 1.0 / std::sqrt(3.0);
  H[0] = 0.5 * (r[2] + s3 * r[7]);
  H[4] = 0.5 * (-r[2] + s3 * r[7]);
  H[8] = 0.5 * (-2.0 * s3 * r[7]);
  H[1] = 0.5 * cd(r[0], -r[1]);
  H[3] = 0.5 * cd(r[0], r[1]);
  H[2] = 0.5 * cd(r[3], -r[4]);
  H[6] = 0.5 * cd(r[3], r[4]);
  H[5] = 0.5 * cd(r[5], -r[6]);
  H[7] = 0.5 * cd(r[5], r[6]);
}

void mul3(const cd* a, const cd* b, cd* r) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}

void expi(const cd Q[9], cd out[9]) {
  cd Q2[9], Q3[9];
  mul3(Q, Q, Q2);
  mul3(Q2, Q, Q3);
  double c0 = (Q3[0] + Q3[4] + Q3[8]).real() / 3.0;
  const double c1 = (Q2[0] + Q2[4] + Q2[8]).real() / 2.0;
  if (c1 < 1e-12) {
    for (int k = 0; k < 9; ++k) out[k] = cd(0, 1) * Q[k] - 0.5 * Q2[k];
    out[0] += 1.0; out[4] += 1.0; out[8] += 1.0;
    return;
  }
  const bool neg = c0 < 0;
  c0 = std::abs(c0);
  const double c0max = 2.0 * std::sqrt(c1 * c1 * c1 / 27.0);
  const double th = std::acos(std::clamp(c0 / c0max, -1.0, 1.0));
  const double w = std::sqrt(c1) * std::sin(th / 3.0), u = std::sqrt(c1 / 3.0) * std::cos(th / 3.0);
  const double xi = std::abs(w) <= 0.05 ? 1.0 - (w * w / 6.0) * (1.0 - (w * w / 20.0) * (1.0 - w * w / 42.0))
                                        : std::sin(w) / w;
  const cd e2 = std::exp(cd(0, 2 * u)), em = std::exp(cd(0, -u));
  const double cw = std::cos(w);
  const cd h0 = (u * u - w * w) * e2 + em * (8.0 * u * u * cw + cd(0, 2 * u * (3 * u * u + w * w)) * xi);
  const cd h1 = 2.0 * u * e2 - em * (2.0 * u * cw - cd(0, 3 * u * u - w * w) * xi);
  const cd h2 = e2 - em * (cw + cd(0, 3 * u) * xi);
  const double den = 9 * u * u - w * w;
  cd f0 = h0 / den, f1 = h1 / den, f2 = h2 / den;
  if (neg) { f0 = std::conj(f0); f1 = -std::conj(f1); f2 = std::conj(f2); }
  for (int k = 0; k < 9; ++k) out[k] = f1 * Q[k] + f2 * Q2[k];
  out[0] += f0; out[4] += f0; out[8] += f0;
}

template <class F> void parallel_for(long n, F f) {
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const long per = (n + hw - 1) / hw;
  std::vector<std::thread> ts;
  for (unsigned t = 0; t < hw; ++t) {
    const long a = t * per, b = std::min(n, a + per);
    if (a >= b) break;
    ts.emplace_back([=] { for (long i = a; i < b; ++i) f(i); });
  }
  for (auto& t : ts) t.join();
}

long vol(const int* d) { return (long)d[0] * d[1] * d[2] * d[3]; }

}  // namespace

void gen_links(const int* dims, double eps, unsigned long long seed, double* out) {
  const Philox ph(seed);
  parallel_for(4 * vol(dims), [&](long l) {
    double r[8];
    normals8(ph, 1u, (uint64_t)l, r);
    cd H[9], U[9];
    algebra(r, H);
    for (auto& h : H) h *= eps;
    expi(H, U);
    for (int k = 0; k < 9; ++k) {
      out[18 * l + 2 * k] = U[k].real();
      out[18 * l + 2 * k + 1] = U[k].imag();
    }
  });
}

void gen_spinor(const int* dims, double sigma, unsigned long long seed, double* out) {
  const Philox ph(seed);
  parallel_for(vol(dims), [&](long s) {
    double r[8];
    normals8(ph, 2u, (uint64_t)s, r);
    for (int k = 0; k < 6; ++k) out[6 * s + k] = sigma * r[k];
  });
}

void gen_momenta(const int* dims, unsigned long long seed, double* out) {
  const Philox ph(seed);
  parallel_for(4 * vol(dims), [&](long l) {
    double r[8];
    normals8(ph, 3u, (uint64_t)l, r);
    cd H[9];
    algebra(r, H);
    for (int k = 0; k < 9; ++k) {
      out[18 * l + 2 * k] = H[k].real();
      out[18 * l + 2 * k + 1] = H[k].imag();
    }
  });
}

}  // namespace sfh
