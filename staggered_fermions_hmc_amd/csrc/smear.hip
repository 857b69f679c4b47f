// HEX (hypercubic stout) smearing of device-resident links
// (Smearing.cpp:76-331): three nested stout levels, each
//   V = exp(proj(alpha/2 Gamma U^dagger)) U,
// proj the traceless anti-hermitian part, Gamma a sum of up and down staples
// (paths x -> x + sigma -> x + sigma + mu -> x + mu and the mirror image).
//   level 1: V1 of direction mu dressed in one plane (mu, sigma) from thin
//            links, alpha3;
//   level 2: V2 of direction mu excluding nu, staples in the two sigma
//            outside {mu, nu} built from level-1 links that exclude the
//            fourth direction rho, alpha2;
//   level 3: staples in every nu != mu from level-2 links that exclude the
//            other direction of their plane, alpha1; replaces U.
// The exponential is exp(A) = exp(iQ) with Q = -iA hermitian: the reference
// passes the anti-hermitian A to its exp(iQ) routine, whose guards reject it
// on any non-trivial links, and leaves the level-3 sum uninitialised
// (Smearing.cpp:243); both are restated as intended.  V1 and V2 are stored
// once per ordered pair of distinct directions (12 links per site each): the
// reference's V1[mu][nu][rho] depends on (nu, rho) only through the one
// direction outside {mu, nu, rho}.
//
// One lane per output link; a level is one launch over (site range, pair),
// the pair fastest and the blocks of one site range on one XCD, so the
// links around a range are read from HBM once and from that XCD's L2 by the
// other outputs of the range.  A range is 64 checkerboard sites on both
// parities, and the ranges follow the hop kernels' tiled column walk
// (range_start), so x0- and x1-neighbour ranges are also processed close in
// time.  FP64 throughout; level 2 (four staples of
// level-1 links per output, 12 outputs per site) dominates.
#include <cstring>

#include "internal.h"
#include "staples.h"
#include "su3exp.h"

namespace sfh {

namespace {

constexpr int kHexThreads = 2 * kRangeHalf;

__device__ __forceinline__ int pair_slot(int a, int b) { return 3 * a + (b < a ? b : b - 1); }

// TWO: the thin links are SU(3), and so are the level links (exp of an
// anti-hermitian matrix times an SU(3) link): level links are stored as two
// rows and every link load reads two rows and rebuilds the third.
__device__ __forceinline__ void third_row(M& m) {
  m.e[6] = conj(m.e[1] * m.e[5] - m.e[2] * m.e[4]);
  m.e[7] = conj(m.e[2] * m.e[3] - m.e[0] * m.e[5]);
  m.e[8] = conj(m.e[0] * m.e[4] - m.e[1] * m.e[3]);
}

// Link of pair slot k at site i of a level array W[(k * n + e) * V + i],
// n = 6 (TWO) or 9 elements per link.
template <bool TWO>
__device__ __forceinline__ M ldW(const double2* __restrict__ W, const Geom& g, int i, int k) {
  constexpr int n = TWO ? 6 : 9;
  M m;
#pragma unroll
  for (int e = 0; e < n; ++e) {
    const double2 v = W[(size_t)(k * n + e) * g.V + i];
    m.e[e] = {v.x, v.y};
  }
  if constexpr (TWO) third_row(m);
  return m;
}
template <bool TWO>
__device__ __forceinline__ void stW(double2* __restrict__ W, const Geom& g, int i, int k, const M& m) {
  constexpr int n = TWO ? 6 : 9;
#pragma unroll
  for (int e = 0; e < n; ++e) W[(size_t)(k * n + e) * g.V + i] = make_double2(m.e[e].re, m.e[e].im);
}
// Rows 0 and 1 only of an SU(3) link of a two-row level array (the SU(3)
// product forms need no third row).
__device__ __forceinline__ M ldW2(const double2* __restrict__ W, const Geom& g, int i, int k) {
  M m;
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const double2 v = W[(size_t)(k * 6 + e) * g.V + i];
    m.e[e] = {v.x, v.y};
  }
  return m;
}
// Rows 0 and 1 only of an SU(3) thin link.
__device__ __forceinline__ M ldT2(const double2* __restrict__ U, const Geom& g, int i, int mu) {
  M m;
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const double2 v = U[(size_t)(mu * 9 + e) * g.V + i];
    m.e[e] = {v.x, v.y};
  }
  return m;
}
// Thin link (all nine elements stored; two rows read when TWO).
template <bool TWO>
__device__ __forceinline__ M ldT(const double2* __restrict__ U, const Geom& g, int i, int mu) {
  if constexpr (TWO) {
    M m;
#pragma unroll
    for (int e = 0; e < 6; ++e) {
      const double2 v = U[(size_t)(mu * 9 + e) * g.V + i];
      m.e[e] = {v.x, v.y};
    }
    third_row(m);
    return m;
  } else {
    return ldL(U, g, i, mu);
  }
}

// Up plus down staple of link (x, mu) in the (mu, sig) plane; L(i, d, full)
// is the link of direction d at device site i for the level at hand, with
// all three rows (full) or, for TWO, rows 0 and 1 only.  The site walks
// carry their index (SiteW).  TWO: every link is SU(3), so the products take
// the column-rebuild forms (rmul_n_su3 / rmul_d_su3), which need the third
// row of no link but the first of each path.
template <bool TWO, class F>
__device__ __forceinline__ M staple_pair(const Geom& g, const SiteW& x, int mu, int sig, F L) {
  SiteW xs = x, xm = x, xd = x;
  xs.move<+1>(g, sig);
  xm.move<+1>(g, mu);
  xd.move<-1>(g, sig);
  SiteW xdm = xd;
  xdm.move<+1>(g, mu);
  if constexpr (TWO) {
    M up = L(x.index(g), sig, true);
    rmul_n_su3(up, L(xs.index(g), mu, false));
    rmul_d_su3(up, L(xm.index(g), sig, false));
    M dn = mdag(L(xd.index(g), sig, true));
    rmul_n_su3(dn, L(xd.index(g), mu, false));
    rmul_n_su3(dn, L(xdm.index(g), sig, false));
    return madd(up, dn);
  } else {
    M up = L(x.index(g), sig, true);
    rmul_n(up, L(xs.index(g), mu, true));
    rmul_d(up, L(xm.index(g), sig, true));
    M dn = mdag(L(xd.index(g), sig, true));
    rmul_n(dn, L(xd.index(g), mu, true));
    rmul_n(dn, L(xdm.index(g), sig, true));
    return madd(up, dn);
  }
}

// exp(proj(alpha/2 G U^dagger)) U (Smearing.cpp:126-129, :143-145, :223-225, :291-293).
__device__ __forceinline__ int stout(const M& G, const M& U, double alpha, M& out) {
  const M W = mmul_nd(G, U);
  const double h = 0.25 * alpha;  // alpha/2 times the 1/2 of the anti-hermitian part
  M Q;                            // Q = -i A, A = alpha/2 proj(W)
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const C d = W.e[3 * a + b] - conj(W.e[3 * b + a]);
      Q.e[3 * a + b] = {h * d.im, -h * d.re};
    }
  // remove the trace (real after the -i)
  const double t = (Q.e[0].re + Q.e[4].re + Q.e[8].re) / 3.0;
  Q.e[0].re -= t;
  Q.e[4].re -= t;
  Q.e[8].re -= t;
  double qn = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) qn += Q.e[k].re * Q.e[k].re + Q.e[k].im * Q.e[k].im;
  if (!isfinite(qn)) return SFH_E_INVALID;  // non-finite links would pass the exponential's guards
  M e;
  const int st = exp_iq(Q, e);
  if (st == SFH_OK) out = mmul(e, U);
  return st;
}

// Output slot k (0 .. nk-1) and device site of this lane (-1: none): blocks
// ordered (site range, slot) with the slot fastest, XCD-aware; the lower half
// of a block takes the even sites of its range, the upper half the odd ones.
__device__ __forceinline__ int hex_lane(const Geom& g, int nk, int nb, int& k) {
  const int lb = xcd_block(nb);
  k = lb % nk;
  return range_site(g, lb / nk, threadIdx.x);
}

template <bool TWO>
__global__ __launch_bounds__(kHexThreads) void k_hex1(Geom g, const double2* __restrict__ U,
                                                      double2* __restrict__ V1, double alpha, int* status, int nb) {
  int k;
  const int i = hex_lane(g, 12, nb, k);
  if (i < 0) return;
  const int mu = k / 3, r = k % 3, sig = r < mu ? r : r + 1;
  const SiteW x = SiteW::of(g, site_of(g, i));
  const M G = staple_pair<TWO>(g, x, mu, sig,
                               [&](int j, int d, bool full) { return full ? ldT<TWO>(U, g, j, d) : ldT2(U, g, j, d); });
  M out;
  const int st = stout(G, ldT<TWO>(U, g, i, mu), alpha, out);
  if (st != SFH_OK) atomicMax(status, st);
  else stW<TWO>(V1, g, i, k, out);
}

template <bool TWO>
__global__ __launch_bounds__(kHexThreads) void k_hex2(Geom g, const double2* __restrict__ U,
                                                      const double2* __restrict__ V1, double2* __restrict__ V2,
                                                      double alpha, int* status, int nb) {
  int k;
  const int i = hex_lane(g, 12, nb, k);
  if (i < 0) return;
  const int mu = k / 3, r = k % 3, nu = r < mu ? r : r + 1;
  const SiteW x = SiteW::of(g, site_of(g, i));
  M G = mzero<double>();
#pragma unroll 1
  for (int sig = 0; sig < 4; ++sig) {
    if (sig == mu || sig == nu) continue;
    const int rho = 6 - mu - nu - sig;
    G = madd(G, staple_pair<TWO>(g, x, mu, sig, [&](int j, int d, bool full) {
               return full ? ldW<TWO>(V1, g, j, pair_slot(d, rho)) : ldW2(V1, g, j, pair_slot(d, rho));
             }));
  }
  M out;
  const int st = stout(G, ldT<TWO>(U, g, i, mu), alpha, out);
  if (st != SFH_OK) atomicMax(status, st);
  else stW<TWO>(V2, g, i, k, out);
}

// Level 3 reads only its own thin link besides level-2 links, so it updates
// U in place.
template <bool TWO>
__global__ __launch_bounds__(kHexThreads) void k_hex3(Geom g, double2* __restrict__ U,
                                                      const double2* __restrict__ V2, double alpha, int* status,
                                                      int nb) {
  int mu;
  const int i = hex_lane(g, 4, nb, mu);
  if (i < 0) return;
  const SiteW x = SiteW::of(g, site_of(g, i));
  M G = mzero<double>();
#pragma unroll 1
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
    G = madd(G, staple_pair<TWO>(g, x, mu, nu, [&](int j, int d, bool full) {
               const int k = pair_slot(d, d == mu ? nu : mu);
               return full ? ldW<TWO>(V2, g, j, k) : ldW2(V2, g, j, k);
             }));
  }
  M out;
  const int st = stout(G, ldT<TWO>(U, g, i, mu), alpha, out);
  if (st != SFH_OK) atomicMax(status, st);
  else stL(U, g, i, mu, out);
}

}  // namespace

template <bool TWO>
static int hex_steps(sfh_ctx* c, const Geom& g, double2* U, double2* V1, double2* V2, const double alpha[3],
                     int nsteps) {
  int* d_status = &c->d_state[1].status;
  hipMemsetAsync(d_status, 0, sizeof(int), c->stream);
  const int nb = (g.Vh + kRangeHalf - 1) / kRangeHalf;  // site ranges
  for (int n = 0; n < nsteps; ++n) {
    {
      KTimer kt(c, "hex1");
      hipLaunchKernelGGL(k_hex1<TWO>, dim3(12 * nb), dim3(kHexThreads), 0, c->stream, g, U, V1, alpha[2], d_status,
                         12 * nb);
    }
    {
      KTimer kt(c, "hex2");
      hipLaunchKernelGGL(k_hex2<TWO>, dim3(12 * nb), dim3(kHexThreads), 0, c->stream, g, U, V1, V2, alpha[1],
                         d_status, 12 * nb);
    }
    {
      KTimer kt(c, "hex3");
      hipLaunchKernelGGL(k_hex3<TWO>, dim3(4 * nb), dim3(kHexThreads), 0, c->stream, g, U, V2, alpha[0], d_status,
                         4 * nb);
    }
    if (int e = hip_check(c, hipGetLastError(), "hex smearing launch")) return e;
  }
  int* hs = reinterpret_cast<int*>(c->h_scalar);
  hipMemcpyAsync(hs, d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (int e = hip_check(c, hipStreamSynchronize(c->stream), "hex smearing")) return e;
  if (*hs != SFH_OK) return set_err(c, *hs, "hex smearing: non-finite links");
  return SFH_OK;
}

int hex_smear(sfh_ctx* c, const Geom& g, double2* U, const double alpha[3], int nsteps) {
  if (nsteps == 0) return SFH_OK;
  if (!staple_dims_ok(g)) return set_err(c, SFH_E_INVALID, "hex smearing: lattice extents above 255");
  double2* V1 = static_cast<double2*>(scratch(c, "hex_v1", 108 * (size_t)g.V * sizeof(double2)));
  double2* V2 = static_cast<double2*>(scratch(c, "hex_v2", 108 * (size_t)g.V * sizeof(double2)));
  unsigned* dev = static_cast<unsigned*>(scratch(c, "su3_dev", sizeof(unsigned)));
  if (!V1 || !V2 || !dev) return SFH_E_NOMEM;
  // SU(3) thin links (one check per call: smearing keeps them SU(3)): two rows
  hipMemsetAsync(dev, 0, sizeof(unsigned), c->stream);
  su3_deviation(c->stream, g, U, dev);
  unsigned* hd = reinterpret_cast<unsigned*>(c->h_scalar);
  hipMemcpyAsync(hd, dev, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream);
  if (int e = hip_check(c, hipStreamSynchronize(c->stream), "SU(3) check")) return e;
  float d;
  std::memcpy(&d, hd, sizeof d);
  return d < kSu3TolF64 ? hex_steps<true>(c, g, U, V1, V2, alpha, nsteps) : hex_steps<false>(c, g, U, V1, V2, alpha, nsteps);
}

}  // namespace sfh
