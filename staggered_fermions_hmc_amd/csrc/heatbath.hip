// Cabibbo-Marinari heatbath of the Symanzik gauge action on device-resident
// links (Lattice.cpp:509-757, SU3_Sampling.cpp:17-199).
//
// A sweep updates the links one direction mu at a time, each direction in
// four colour classes (2 n_mu + sum of the other coordinates) & 3.  The
// mu-links in the plaquette and rectangle staples of (n, mu) sit at offsets
// +-mu, +-nu, +-2 nu and +-mu +-nu, which change that colour by 2, 1, 2 and
// 1 or 3 (mod 4): a class is a set of links none of whose staples holds
// another, so one launch updates a whole class, one lane per link, in place.
// Every extent must be a multiple of 4.  Each link update is the reference's
// single_link_heatbath (Lattice.cpp:696-723): staple sums A1, A2 from the
// current links, A = c11 A1 + c12 A2, then the SU(2) subgroups (0,1), (0,2),
// (1,2) in turn, then the link's ReTr(U A1) / 3 and -(beta/3) ReTr(U A).
//
// Uniforms come from Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): draw
// k of link l in sweep s is half (k & 1) of the block {k >> 1, l, s, tag}
// under the 64-bit seed, with l = 4 s_lex + mu in the reference's site order
// (Fermions.cpp:42), so a sweep depends only on the seed and the sweep index.
//
// A class is a diagonal set of sites: in the solver's even-odd order its
// sites never fill a cache line, so the sweeps run on a copy of the links in
// a layout of their own, x3 = 4 q + r with q fastest and r slowest
// (HbIndex).  The lanes of a class launch step along q with the other
// coordinates fixed, and so does every staple link they read: each load of
// a wave is a set of whole 128-byte rows.  A call converts the links into
// that layout, runs its sweeps, and converts them back.
//
// About 73 SU(3) products per link: the staple paths grouped by shared
// factors (66, and 6 more for the plaquette sum the measurement needs).
#include <cstring>

#include "internal.h"
#include "reduce.h"
#include "staples.h"

namespace sfh {

namespace {

constexpr uint32_t kHeatbathTag = 0x68623031u;
constexpr double kPi = 3.14159265358979323846;  // SU3_Sampling.h:10
constexpr int kHbThreads = 128;
// Every rejection loop gives up after this many tries and fails the call: the
// reference's loops (SU3_Sampling.cpp:19-28, :54-67, :176-190) would run for
// ever on non-finite links or when a * beta is so small that no proposal is
// accepted.  An ordinary link needs a few tries.
constexpr int kMaxTries = 1 << 20;

struct Philox {
  uint32_t k0, k1;
  __device__ __forceinline__ void block(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3) const {
    uint32_t a = k0, b = k1;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
      const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
      const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
      const uint32_t n0 = hi1 ^ c1 ^ a, n2 = hi0 ^ c3 ^ b;
      c0 = n0;
      c1 = lo1;
      c2 = n2;
      c3 = lo0;
      a += 0x9E3779B9u;
      b += 0xBB67AE85u;
    }
  }
};

// 53-bit uniform in [0, 1).
__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

struct Draws {
  Philox ph;
  uint32_t link, sweep, q;
  int have;
  double next;
  int fail;  // a rejection loop ran out of tries
  __device__ __forceinline__ double uniform() {
    if (have) {
      have = 0;
      return next;
    }
    uint32_t c0 = q++, c1 = link, c2 = sweep, c3 = kHeatbathTag;
    ph.block(c0, c1, c2, c3);
    next = u53(c2, c3);
    have = 1;
    return u53(c0, c1);
  }
};

// x0 with density sqrt(1 - x0^2) exp(a beta x0) (SU3_Sampling.cpp:17-29).
__device__ double gen_x0(double a, double beta, Draws& rs) {
  for (int tries = 0; tries < kMaxTries; ++tries) {
    const double r1 = 1.0 - rs.uniform();
    const double r2 = 1.0 - rs.uniform();
    const double r3 = 1.0 - rs.uniform();
    const double c = cos(2 * kPi * r2);
    const double lam2 = -1 / (2 * a * beta) * (log(r1) + c * c * log(r3));
    const double x = 1 - lam2;
    const double r = rs.uniform();
    if (r * r <= x && x >= -1 && x <= 1) return 1 - 2 * lam2;
  }
  rs.fail = 1;
  return 1.0;
}

// Quaternion q0 + i q.sigma, stored as (q0, q1, q2, q3); as a matrix
// [[q0 + i q3, q2 + i q1], [-q2 + i q1, q0 - i q3]] (SU3_Sampling.cpp:82-95).
struct Quat {
  double q0, q1, q2, q3;
};

// Heatbath SU(2) element (SU3_Sampling.cpp:82-95, :53-78).
__device__ Quat su2_heat(double a, double beta, Draws& rs) {
  Quat x;
  x.q0 = gen_x0(a, beta, rs);
  double v0 = 1.0, v1 = 0.0, v2 = 0.0, len = 1.0;
  int tries = 0;
  for (; tries < kMaxTries; ++tries) {
    v0 = 2.0 * rs.uniform() - 1.0;
    v1 = 2.0 * rs.uniform() - 1.0;
    v2 = 2.0 * rs.uniform() - 1.0;
    len = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    if (len <= 1.0) break;
  }
  if (tries == kMaxTries) rs.fail = 1;
  const double nrm = sqrt(1 - x.q0 * x.q0);
  x.q1 = v0 / len * nrm;
  x.q2 = v1 / len * nrm;
  x.q3 = v2 / len * nrm;
  return x;
}

// Haar-random SU(2) element (SU3_Sampling.cpp:175-199).
__device__ Quat su2_random(Draws& rs) {
  for (int tries = 0; tries < kMaxTries; ++tries) {
    const double r0 = 2.0 * rs.uniform() - 1.0;
    const double r1 = 2.0 * rs.uniform() - 1.0;
    const double r2 = 2.0 * rs.uniform() - 1.0;
    const double r3 = 2.0 * rs.uniform() - 1.0;
    const double len = sqrt(r0 * r0 + r1 * r1 + r2 * r2 + r3 * r3);
    if (len <= 1.0) return Quat{r0 / len, r1 / len, r2 / len, r3 / len};
  }
  rs.fail = 1;
  return Quat{1.0, 0.0, 0.0, 0.0};
}

// One subgroup update U <- embed(R) U on rows/columns (I, J)
// (Lattice.cpp:570-694): the (I, J) block of U A projected onto the
// quaternions, b0 = Re tr B / 2, b_l = Im tr(B sigma_l) / 2; a = |b|; below
// a = 1e-14 R is Haar-random, else R = X (b / a)^dagger with X the heatbath
// sample at (a, 2 beta / 3).
template <int I, int J>
__device__ __forceinline__ void subgroup(M& u, const M& A, double beta, Draws& rs) {
  if (rs.fail) return;
  C w[4];  // (U A)_{II}, (U A)_{IJ}, (U A)_{JI}, (U A)_{JJ}
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int r = e < 2 ? I : J, c = (e & 1) ? J : I;
    C s = u.e[3 * r] * A.e[c];
    cmac(s, u.e[3 * r + 1], A.e[3 + c]);
    cmac(s, u.e[3 * r + 2], A.e[6 + c]);
    w[e] = s;
  }
  const double b0 = 0.5 * (w[0].re + w[3].re);
  const double b1 = 0.5 * (w[1].im + w[2].im);
  const double b2 = 0.5 * (w[1].re - w[2].re);
  const double b3 = 0.5 * (w[0].im - w[3].im);
  const double a = sqrt(b0 * b0 + b1 * b1 + b2 * b2 + b3 * b3);
  Quat r;
  if (a <= 1e-14) {
    r = su2_random(rs);
  } else {
    const Quat x = su2_heat(a, 2 * beta / 3, rs);
    const double v0 = b0 / a, v1 = b1 / a, v2 = b2 / a, v3 = b3 / a;
    // X V^dagger = (x0 v0 + x.v) + i (v0 x - x0 v + x cross v).sigma
    r.q0 = x.q0 * v0 + x.q1 * v1 + x.q2 * v2 + x.q3 * v3;
    r.q1 = v0 * x.q1 - x.q0 * v1 + (x.q2 * v3 - x.q3 * v2);
    r.q2 = v0 * x.q2 - x.q0 * v2 + (x.q3 * v1 - x.q1 * v3);
    r.q3 = v0 * x.q3 - x.q0 * v3 + (x.q1 * v2 - x.q2 * v1);
  }
  const C r00 = {r.q0, r.q3}, r01 = {r.q2, r.q1}, r10 = {-r.q2, r.q1}, r11 = {r.q0, -r.q3};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const C ui = u.e[3 * I + c], uj = u.e[3 * J + c];
    C ni = r00 * ui;
    cmac(ni, r01, uj);
    C nj = r10 * ui;
    cmac(nj, r11, uj);
    u.e[3 * I + c] = ni;
    u.e[3 * J + c] = nj;
  }
}

// Heatbath layout of the links: x3 = 4 q + r, site index
// (((r L0 + x0) L1 + x1) L2 + x2) Q + q with Q = L3 / 4.
// TWO: SU(3) links kept as two rows (a third less to read per staple link),
// the third rebuilt on every load.
template <bool TWO>
struct HbIndex {
  static constexpr bool kTwoRows = TWO;
  __device__ __forceinline__ static int of(const Geom& g, const Site4& x) {
    const int Q = g.L[3] >> 2;
    return ((((x.x3 & 3) * g.L[0] + x.x0) * g.L[1] + x.x1) * g.L[2] + x.x2) * Q + (x.x3 >> 2);
  }
};

// Site of lane t of colour class (mu, col): q = x3 >> 2 fastest, then x2,
// x1, x0, then r = x3 & 3; the coordinate ns (2, or 1 when mu = 2) steps by
// 4 with the offset that gives the colour.  Adding 4 to x3 keeps the colour,
// so consecutive lanes are consecutive q in one row of the layout.
__device__ __forceinline__ Site4 class_site(const Geom& g, int mu, int col, int t, int& lex) {
  const int ns = mu == 2 ? 1 : 2;
  const int Q = g.L[3] >> 2;
  const int q = t % Q;
  t /= Q;
  int x[4];
#pragma unroll
  for (int d = 2; d >= 0; --d) {
    const int e = d == ns ? g.L[d] >> 2 : g.L[d];
    x[d] = t % e;
    t /= e;
  }
  x[3] = 4 * q + t;
  const int xm = mu == 0 ? x[0] : (mu == 1 ? x[1] : (mu == 2 ? x[2] : x[3]));
  if (ns == 2) {
    x[2] = 4 * x[2] + ((col - x[0] - x[1] - x[3] - xm) & 3);
  } else {
    x[1] = 4 * x[1] + ((col - x[0] - x[2] - x[3] - xm) & 3);
  }
  lex = ((x[0] * g.L[1] + x[1]) * g.L[2] + x[2]) * g.L[3] + x[3];
  return Site4{x[0], x[1], x[2], x[3]};
}

template <class IX>
__global__ __launch_bounds__(kHbThreads) void k_heatbath(Geom g, double2* __restrict__ U, int mu, int col,
                                                         double beta, double c11, double c12, uint32_t k0,
                                                         uint32_t k1, uint32_t sweep, double* __restrict__ part,
                                                         int* __restrict__ status, int nb) {
  // consecutive site ranges on one XCD share its L2 for the staple links
  const int lb = xcd_block(nb);
  const int t = lb * kHbThreads + threadIdx.x;
  __shared__ double sA1[18][kHbThreads];
  double v[2] = {0.0, 0.0};
  if (t < (g.V >> 2)) {
    int lex;
    const Site4 n = class_site(g, mu, col, t, lex);
    const int i = IX::of(g, n);
    // A = c11 A1 + c12 A2 from the staple paths grouped by shared factors;
    // the plaquette sum A1 (for the measurement) accumulates in LDS
    // (lane-interleaved, no bank conflicts) while the walks run
#pragma unroll
    for (int e = 0; e < 18; ++e) sA1[e][threadIdx.x] = 0.0;
    const M A = staples_fused<IX::kTwoRows, HbLinks<IX::kTwoRows>, true>(U, g, n, mu, c11, c12,
                                                                         &sA1[0][threadIdx.x], kHbThreads);
    M u = ldLx<IX>(U, g, i, mu);
    Draws rs{Philox{k0, k1}, 4u * (uint32_t)lex + (uint32_t)mu, sweep, 0u, 0, 0.0, 0};
    subgroup<0, 1>(u, A, beta, rs);
    subgroup<0, 2>(u, A, beta, rs);
    subgroup<1, 2>(u, A, beta, rs);
    if (rs.fail) {
      // the link keeps its value and the call fails
      atomicMax(status, (int)SFH_E_SAMPLER);
    } else {
      if constexpr (IX::kTwoRows) {
        orthonormalize_rows01(u);
        u.e[6] = conj(u.e[1] * u.e[5] - u.e[2] * u.e[4]);
        u.e[7] = conj(u.e[2] * u.e[3] - u.e[0] * u.e[5]);
        u.e[8] = conj(u.e[0] * u.e[4] - u.e[1] * u.e[3]);
      }
      stLx<IX>(U, g, i, mu, u);
      M A1;
#pragma unroll
      for (int e = 0; e < 9; ++e) A1.e[e] = {sA1[2 * e][threadIdx.x], sA1[2 * e + 1][threadIdx.x]};
      v[0] = retr_mul(u, A1) / 3.0;
      v[1] = -(beta / 3.0) * retr_mul(u, A);
    }
  }
  block_sum_n<2>(v);
  if (threadIdx.x == 0) {
    part[2 * lb] = v[0];
    part[2 * lb + 1] = v[1];
  }
}

// Links between the solver's even-odd layout and the heatbath layout, one
// lane per (site, direction); to_hb selects the direction of the copy.
template <class IX>
__global__ __launch_bounds__(256) void k_hb_layout(Geom g, const double2* __restrict__ src,
                                                   double2* __restrict__ dst, int to_hb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (i >= g.V) return;
  const int h = IX::of(g, site_of(g, i));
  if (to_hb) stLx<IX>(dst, g, h, mu, ldL(src, g, i, mu));
  else stL(dst, g, i, mu, ldLx<IX>(src, g, h, mu));
}

// Sum of a sweep's n workgroup partials, in a fixed order.
__global__ __launch_bounds__(256) void k_heatbath_total(const double* __restrict__ part, int n,
                                                        double* __restrict__ out) {
  double v[2] = {0.0, 0.0};
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    v[0] += part[2 * k];
    v[1] += part[2 * k + 1];
  }
  block_sum_n<2>(v);
  if (threadIdx.x == 0) {
    out[0] = v[0];
    out[1] = v[1];
  }
}

}  // namespace

bool heatbath_dims_ok(const Geom& g) {
  for (int d = 0; d < 4; ++d)  // the staple walks pack each coordinate into 8 bits
    if (g.L[d] < 4 || g.L[d] % 4 || g.L[d] > 252) return false;
  return true;
}

template <class IX>
static int heatbath_sweeps(sfh_ctx* c, const Geom& g, double2* U, double2* H, double* part, double* sums, double beta,
                    double c11, double c12, unsigned long long seed, unsigned sweep0, int nsweeps, double* plaq,
                    double* act) {
  const int nb = ((g.V >> 2) + kHbThreads - 1) / kHbThreads;
  const dim3 lgrid((g.V + 255) / 256, 4);
  {
    KTimer kt(c, "heatbath_layout");
    hipLaunchKernelGGL(k_hb_layout<IX>, lgrid, dim3(256), 0, c->stream, g, U, H, 1);
  }
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  int* d_status = &c->d_state[1].status;
  hipMemsetAsync(d_status, 0, sizeof(int), c->stream);
  for (int s = 0; s < nsweeps; ++s) {
    {
      KTimer kt(c, "heatbath");
      for (int mu = 0; mu < 4; ++mu)
        for (int col = 0; col < 4; ++col)
          hipLaunchKernelGGL(k_heatbath<IX>, dim3(nb), dim3(kHbThreads), 0, c->stream, g, H, mu, col, beta, c11, c12,
                             k0, k1, (uint32_t)(sweep0 + s), part + 2 * (size_t)nb * (mu * 4 + col), d_status,
                             nb);
    }
    hipLaunchKernelGGL(k_heatbath_total, dim3(1), dim3(256), 0, c->stream, part, 16 * nb, sums + 2 * s);
    if (int e = hip_check(c, hipGetLastError(), "heatbath launch")) return e;
  }
  {
    KTimer kt(c, "heatbath_layout");
    hipLaunchKernelGGL(k_hb_layout<IX>, lgrid, dim3(256), 0, c->stream, g, H, U, 0);
  }
  if (int e = hip_check(c, hipGetLastError(), "heatbath layout")) return e;
  std::vector<double> h(2 * (size_t)nsweeps);
  hipMemcpyAsync(h.data(), sums, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  int* hs = reinterpret_cast<int*>(c->h_scalar);
  hipMemcpyAsync(hs, d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (int e = hip_check(c, hipStreamSynchronize(c->stream), "heatbath")) return e;
  if (*hs != SFH_OK)
    return set_err(c, *hs, "heatbath: a rejection loop found no sample in 2^20 tries (non-finite links or a "
                           "vanishing a*beta); those links kept their values");
  for (int s = 0; s < nsweeps; ++s) {
    if (plaq) plaq[s] = (h[2 * s] / 4.0) / (6.0 * g.V);
    if (act) act[s] = (h[2 * s + 1] / 4.0) / g.V;
  }
  return SFH_OK;
}


int heatbath(sfh_ctx* c, const Geom& g, double2* U, double beta, double c11, double c12, unsigned long long seed,
             unsigned sweep0, int nsweeps, double* plaq, double* act) {
  if (nsweeps == 0) return SFH_OK;
  const int nb = ((g.V >> 2) + kHbThreads - 1) / kHbThreads;
  double* part = static_cast<double*>(scratch(c, "hb_part", 2 * 16 * (size_t)nb * sizeof(double)));
  double* sums = static_cast<double*>(scratch(c, "hb_sums", 2 * (size_t)nsweeps * sizeof(double)));
  double2* H = static_cast<double2*>(scratch(c, "hb_links", 36 * (size_t)g.V * sizeof(double2)));
  unsigned* dev = static_cast<unsigned*>(scratch(c, "su3_dev", sizeof(unsigned)));
  if (!part || !sums || !H || !dev) return SFH_E_NOMEM;
  // SU(3) links are kept as two rows in the heatbath copy
  hipMemsetAsync(dev, 0, sizeof(unsigned), c->stream);
  su3_deviation(c->stream, g, U, dev);
  unsigned* hd = reinterpret_cast<unsigned*>(c->h_scalar);
  hipMemcpyAsync(hd, dev, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream);
  if (int e = hip_check(c, hipStreamSynchronize(c->stream), "SU(3) check")) return e;
  float d;
  std::memcpy(&d, hd, sizeof d);
  return d < kSu3TolF64 ? heatbath_sweeps<HbIndex<true>>(c, g, U, H, part, sums, beta, c11, c12, seed, sweep0, nsweeps,
                                                      plaq, act)
                     : heatbath_sweeps<HbIndex<false>>(c, g, U, H, part, sums, beta, c11, c12, seed, sweep0, nsweeps,
                                                       plaq, act);
}

}  // namespace sfh
