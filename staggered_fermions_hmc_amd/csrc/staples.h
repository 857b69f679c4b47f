// Closed-path staples of the Symanzik gauge action, shared by the force,
// action and heatbath kernels: a site as four coordinates in registers, link
// loads and stores in the device layout, row-wise SU(3) products, and the
// plaquette and rectangle paths walked from n + mu back to n.
#pragma once
#include "internal.h"

namespace sfh {

using M = mat3<double>;
using C = cpx<double>;

// A site as four coordinates in registers.  Moves pick the coordinate with
// selects: a runtime index into a coordinate array (the direction of a move
// is not a compile-time constant) would put the array in scratch memory.
struct Site4 {
  int x0, x1, x2, x3;
  __device__ __forceinline__ void move(const Geom& g, int d, int s) {
    const int L = d == 0 ? g.L[0] : (d == 1 ? g.L[1] : (d == 2 ? g.L[2] : g.L[3]));
    int v = (d == 0 ? x0 : (d == 1 ? x1 : (d == 2 ? x2 : x3))) + s;
    v = v < 0 ? v + L : (v >= L ? v - L : v);
    x0 = d == 0 ? v : x0;
    x1 = d == 1 ? v : x1;
    x2 = d == 2 ? v : x2;
    x3 = d == 3 ? v : x3;
  }
  __device__ __forceinline__ int index(const Geom& g) const {
    const int p = (x0 + x1 + x2 + x3) & 1;
    return p * g.Vh + ((x0 * g.L[1] + x1) * g.L[2] + x2) * g.L3h + (x3 >> 1);
  }
};

// Where a site's links sit in the links array: the solver's even-odd order,
// all nine elements stored.
struct EoIndex {
  static constexpr bool kTwoRows = false;
  __device__ __forceinline__ static int of(const Geom& g, const Site4& x) { return x.index(g); }
};

// Device site of lane t (0 .. 2 kRangeHalf - 1) of site range r: kRangeHalf
// checkerboard sites from the tiled column walk (range_start) on each parity,
// the lower half of the lanes even, the upper half odd; -1 past the end.
// Neighbouring ranges in x0 and x1 come close together in the walk, so
// kernels that read links around their sites find them in L2.
constexpr int kRangeHalf = 64;
__device__ __forceinline__ int range_site(const Geom& g, int r, int t) {
  const int cb = range_start(g, r, kRangeHalf) + (t & (kRangeHalf - 1));
  if (cb >= g.Vh) return -1;
  return (t >= kRangeHalf ? g.Vh : 0) + cb;
}

__device__ __forceinline__ Site4 site_of(const Geom& g, int i) {
  const int p = i >= g.Vh;
  int n[4];
  cb_coords(g, i - p * g.Vh, p, n);
  return Site4{n[0], n[1], n[2], n[3]};
}

__device__ __forceinline__ M ldL(const double2* __restrict__ U, const Geom& g, int i, int mu) {
  M m;
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const double2 v = U[(size_t)(mu * 9 + e) * g.V + i];
    m.e[e] = {v.x, v.y};
  }
  return m;
}
__device__ __forceinline__ void stL(double2* __restrict__ U, const Geom& g, int i, int mu, const M& m) {
#pragma unroll
  for (int e = 0; e < 9; ++e) U[(size_t)(mu * 9 + e) * g.V + i] = make_double2(m.e[e].re, m.e[e].im);
}

// Links of a layout policy IX: all nine elements at (mu * 9 + e) V + i, or,
// for SU(3) links kept as two rows (IX::kTwoRows), six at (mu * 6 + e) V + i
// with the third row rebuilt as conj(row0 x row1).
template <class IX>
__device__ __forceinline__ M ldLx(const double2* __restrict__ U, const Geom& g, int i, int mu) {
  if constexpr (IX::kTwoRows) {
    M m;
#pragma unroll
    for (int e = 0; e < 6; ++e) {
      const double2 v = U[(size_t)(mu * 6 + e) * g.V + i];
      m.e[e] = {v.x, v.y};
    }
    m.e[6] = conj(m.e[1] * m.e[5] - m.e[2] * m.e[4]);
    m.e[7] = conj(m.e[2] * m.e[3] - m.e[0] * m.e[5]);
    m.e[8] = conj(m.e[0] * m.e[4] - m.e[1] * m.e[3]);
    return m;
  } else {
    return ldL(U, g, i, mu);
  }
}
template <class IX>
__device__ __forceinline__ void stLx(double2* __restrict__ U, const Geom& g, int i, int mu, const M& m) {
  if constexpr (IX::kTwoRows) {
#pragma unroll
    for (int e = 0; e < 6; ++e) U[(size_t)(mu * 6 + e) * g.V + i] = make_double2(m.e[e].re, m.e[e].im);
  } else {
    stL(U, g, i, mu, m);
  }
}

// Gram-Schmidt on rows 0 and 1 of a link about to be stored as two rows.
// The third row is rebuilt from them on every load, so a link updated and
// stored again and again (the heatbath) feeds any non-unitarity of rows 0-1
// back through the rebuilt row: at strong coupling the rounding grew about
// 1.4-fold per sweep and reached O(1) after ~100 sweeps.  Re-orthonormalising
// on every store keeps it at the rounding of one update.
__device__ __forceinline__ void orthonormalize_rows01(M& m) {
  double n0 = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) n0 += m.e[c].re * m.e[c].re + m.e[c].im * m.e[c].im;
  const double i0 = 1.0 / sqrt(n0);
#pragma unroll
  for (int c = 0; c < 3; ++c) m.e[c] = {i0 * m.e[c].re, i0 * m.e[c].im};
  C d = {0.0, 0.0};  // <row0, row1>
#pragma unroll
  for (int c = 0; c < 3; ++c) cmacc(d, m.e[c], m.e[3 + c]);
  double n1 = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    m.e[3 + c] = m.e[3 + c] - d * m.e[c];
    n1 += m.e[3 + c].re * m.e[3 + c].re + m.e[3 + c].im * m.e[3 + c].im;
  }
  const double i1 = 1.0 / sqrt(n1);
#pragma unroll
  for (int c = 0; c < 3; ++c) m.e[3 + c] = {i1 * m.e[3 + c].re, i1 * m.e[3 + c].im};
}

// t <- t m and t <- t m^dag, one row at a time: only t, m and three entries
// are live, so the staple kernels keep several waves per SIMD.
__device__ __forceinline__ void rmul_n(M& t, const M& m) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const C a0 = t.e[3 * i], a1 = t.e[3 * i + 1], a2 = t.e[3 * i + 2];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      C s = a0 * m.e[j];
      cmac(s, a1, m.e[3 + j]);
      cmac(s, a2, m.e[6 + j]);
      t.e[3 * i + j] = s;
    }
  }
}
__device__ __forceinline__ void rmul_d(M& t, const M& m) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const C a0 = t.e[3 * i], a1 = t.e[3 * i + 1], a2 = t.e[3 * i + 2];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      C s = a0 * conj(m.e[3 * j]);
      cmac(s, a1, conj(m.e[3 * j + 1]));
      cmac(s, a2, conj(m.e[3 * j + 2]));
      t.e[3 * i + j] = s;
    }
  }
}

// Closed paths through link U_mu(n), as moves from n + mu back to n: d = 0
// moves along mu, d = 1 along the plane's other direction nu; a + move from
// x takes U_d(x), a - move takes U_d(x - d)^dag.  The product of a path is
// the staple whose ReTr(U_mu(n) staple) is the loop.
struct Move {
  int d, s;
};
// the two plaquettes of the (mu, nu) plane (Lattice.cpp:249-313): below, above
static __constant__ Move kPlaqPath[2][3] = {{{1, -1}, {0, -1}, {1, +1}}, {{1, +1}, {0, -1}, {1, -1}}};
// the six rectangles (Lattice.cpp:315-499): 2 x 1 in nu below and above
// (short side on the link), then 1 x 2 with the link first or second along
// the long side, below and above
static __constant__ Move kRectPath[6][5] = {
    {{1, -1}, {1, -1}, {0, -1}, {1, +1}, {1, +1}},
    {{1, +1}, {1, +1}, {0, -1}, {1, -1}, {1, -1}},
    {{0, +1}, {1, -1}, {0, -1}, {0, -1}, {1, +1}},
    {{1, +1}, {0, -1}, {0, -1}, {1, -1}, {0, +1}},
    {{1, -1}, {0, -1}, {0, -1}, {1, +1}, {0, +1}},
    {{0, +1}, {1, +1}, {0, -1}, {0, -1}, {1, -1}},
};

template <class IX, int N>
__device__ __forceinline__ M walk(const double2* __restrict__ U, const Geom& g, Site4 x, int mu, int nu,
                                  const Move (&path)[N]) {
  M t;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int d = path[k].d ? nu : mu;
    const int s = path[k].s;
    if (s < 0) x.move(g, d, -1);
    const M m = ldLx<IX>(U, g, IX::of(g, x), d);
    if (s > 0) x.move(g, d, +1);
    if (k == 0) t = s > 0 ? m : mdag(m);
    else if (s > 0) rmul_n(t, m);
    else rmul_d(t, m);
  }
  return t;
}

// c11 A1 + c12 A2 for link (n, mu): A1 the six plaquette staples, A2 the
// eighteen rectangle staples (Lattice.cpp:249-507); IX places sites in U.
template <class IX = EoIndex>
static __device__ M staples(const double2* __restrict__ U, const Geom& g, const Site4& n, int mu, double c11,
                            double c12) {
  M A = mzero<double>();
  Site4 x = n;
  x.move(g, mu, +1);
#pragma unroll 1
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
    if (c11 != 0.0) {
#pragma unroll 1
      for (int k = 0; k < 2; ++k) A = madd(A, mscale(c11, walk<IX>(U, g, x, mu, nu, kPlaqPath[k])));
    }
    if (c12 != 0.0) {
#pragma unroll 1
      for (int k = 0; k < 6; ++k) A = madd(A, mscale(c12, walk<IX>(U, g, x, mu, nu, kRectPath[k])));
    }
  }
  return A;
}

}  // namespace sfh
