// Closed-path staples of the Symanzik gauge action, shared by the force,
// action, heatbath and HEX kernels: sites as coordinates in registers, link
// loads and stores in the device layouts, SU(3) products, and the plaquette
// and rectangle paths walked from n + mu back to n with their shared factors.
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
  // the same move with every choice written as a flat select on the
  // direction (uniform in the staple kernels): nested choices became trees of
  // uniform branches, and the register allocator then kept the union of the
  // values of both sides of every branch alive
  __device__ __forceinline__ void move_f(const Geom& g, int d, int s) {
    const bool f0 = d == 0, f1 = d == 1, f2 = d == 2, f3 = d == 3;
    const int L = (f0 ? g.L[0] : 0) + (f1 ? g.L[1] : 0) + (f2 ? g.L[2] : 0) + (f3 ? g.L[3] : 0);
    int v = (f0 ? x0 : 0) + (f1 ? x1 : 0) + (f2 ? x2 : 0) + (f3 ? x3 : 0) + s;
    v += v < 0 ? L : 0;
    v -= v >= L ? L : 0;
    x0 = f0 ? v : x0;
    x1 = f1 ? v : x1;
    x2 = f2 ? v : x2;
    x3 = f3 ? v : x3;
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

// Link (i, d) with a wave-uniform base per element (d uniform, V a kernel
// argument) and a 32-bit lane offset: the nine loads take a scalar base and
// one vector offset, with no per-element 64-bit address arithmetic in the
// vector ALU.  i < 2^24, so i * 16 fits the 32-bit offset.
__device__ __forceinline__ M ldLs(const double2* __restrict__ U, const Geom& g, int i, int d) {
  const size_t es = (size_t)g.V * sizeof(double2);
  const char* base = reinterpret_cast<const char*>(U) + (size_t)(d * 9) * es;
  const unsigned off = (unsigned)i * (unsigned)sizeof(double2);
  M m;
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const double2 v = *reinterpret_cast<const double2*>(base + e * es + off);
    m.e[e] = {v.x, v.y};
  }
  return m;
}

// t <- t m and t <- t m^dag with every complex multiply-add written as two
// real fused multiply-adds per component (the plain expressions compile to a
// multiply, a fused multiply-add and an add): 12 FP64 instructions per
// element instead of 16.
__device__ __forceinline__ void rmul_nf(M& t, const M& m) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const C a0 = t.e[3 * i], a1 = t.e[3 * i + 1], a2 = t.e[3 * i + 2];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const C b0 = m.e[j], b1 = m.e[3 + j], b2 = m.e[6 + j];
      double re = a0.re * b0.re, im = a0.re * b0.im;
      re = fma(-a0.im, b0.im, re);
      im = fma(a0.im, b0.re, im);
      re = fma(a1.re, b1.re, re);
      re = fma(-a1.im, b1.im, re);
      im = fma(a1.re, b1.im, im);
      im = fma(a1.im, b1.re, im);
      re = fma(a2.re, b2.re, re);
      re = fma(-a2.im, b2.im, re);
      im = fma(a2.re, b2.im, im);
      im = fma(a2.im, b2.re, im);
      t.e[3 * i + j] = {re, im};
    }
  }
}
__device__ __forceinline__ void rmul_df(M& t, const M& m) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const C a0 = t.e[3 * i], a1 = t.e[3 * i + 1], a2 = t.e[3 * i + 2];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      // row j of m, conjugated
      const C b0 = m.e[3 * j], b1 = m.e[3 * j + 1], b2 = m.e[3 * j + 2];
      double re = a0.re * b0.re, im = -a0.re * b0.im;
      re = fma(a0.im, b0.im, re);
      im = fma(a0.im, b0.re, im);
      re = fma(a1.re, b1.re, re);
      re = fma(a1.im, b1.im, re);
      im = fma(-a1.re, b1.im, im);
      im = fma(a1.im, b1.re, im);
      re = fma(a2.re, b2.re, re);
      re = fma(a2.im, b2.im, re);
      im = fma(-a2.re, b2.im, im);
      im = fma(a2.im, b2.re, im);
      t.e[3 * i + j] = {re, im};
    }
  }
}

// ldLs whose lane offset passes through an empty asm statement that also
// takes `dep`: the load cannot issue before `dep` is computed.  Chaining each
// link load of the staple paths to the product before it keeps the compiler
// from hoisting the loads of later paths (independent of earlier ones) to the
// top of the loop body, which held dozens of links in registers at once and
// spilled.
// R12: SU(3) links, elements 0-5 read and row 2 rebuilt as conj(row0 x row1).
// STRIDE: stored elements per link and direction (9, or 6 for a layout of
// two-row links).
template <bool R12 = false, int STRIDE = 9>
__device__ __forceinline__ M ldLs_after(const double2* __restrict__ U, const Geom& g, int i, int d, double dep) {
  const size_t es = (size_t)g.V * sizeof(double2);
  const char* base = reinterpret_cast<const char*>(U) + (size_t)(d * STRIDE) * es;
  unsigned off = (unsigned)i * (unsigned)sizeof(double2);
  asm volatile("" : "+v"(off) : "v"(dep));
  M m;
#pragma unroll
  for (int e = 0; e < (R12 ? 6 : 9); ++e) {
    const double2 v = *reinterpret_cast<const double2*>(base + e * es + off);
    m.e[e] = {v.x, v.y};
  }
  if constexpr (R12) {
    m.e[6] = conj(m.e[1] * m.e[5] - m.e[2] * m.e[4]);
    m.e[7] = conj(m.e[2] * m.e[3] - m.e[0] * m.e[5]);
    m.e[8] = conj(m.e[0] * m.e[4] - m.e[1] * m.e[3]);
  }
  return m;
}

// a += s b (fused multiply-adds)
__device__ __forceinline__ void maxpy(M& a, double s, const M& b) {
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    a.e[k].re = fma(s, b.e[k].re, a.e[k].re);
    a.e[k].im = fma(s, b.e[k].im, a.e[k].im);
  }
}

__device__ __forceinline__ Site4 step(const Geom& g, Site4 x, int d, int s) {
  x.move(g, d, s);
  return x;
}

// The four coordinates of a site packed 8 bits each (x_d in bits 8d .. 8d + 7;
// every extent is below 256), moved by S along a wave-uniform direction d:
// the extent is a scalar select.  w = 1 when the move wrapped.  The wrap is
// arithmetic on a 0/1 flag: selects on it turned into uniform-looking
// branch trees, and the force kernel spilled.
template <int S> __device__ __forceinline__ unsigned packed_move(const Geom& g, unsigned c, int d, int& w) {
  const bool f0 = d == 0, f1 = d == 1, f2 = d == 2, f3 = d == 3;
  const int L = (f0 ? g.L[0] : 0) + (f1 ? g.L[1] : 0) + (f2 ? g.L[2] : 0) + (f3 ? g.L[3] : 0);
  const unsigned sh = 8u * (unsigned)d;
  int v = (int)((c >> sh) & 255u) + S;
  w = S > 0 ? (v >= L ? 1 : 0) : (v < 0 ? 1 : 0);
  v -= w * (S * L);
  return (c & ~(255u << sh)) | ((unsigned)v << sh);
}
__device__ __forceinline__ unsigned pack_site(const Site4& n) {
  return (unsigned)n.x0 | ((unsigned)n.x1 << 8) | ((unsigned)n.x2 << 16) | ((unsigned)n.x3 << 24);
}

// A site of the force's staple walks in the solver's even-odd order: packed
// coordinates, the lexicographic index and the parity's offset p V/2, all
// carried along by every move (strides are scalar selects), so a link index
// costs two vector instructions, where recomputing it from four coordinates
// after each move took about twenty-five.
struct SiteW {
  unsigned c;
  int lex, pv;
  __device__ __forceinline__ static SiteW of(const Geom& g, const Site4& n) {
    SiteW w;
    w.c = pack_site(n);
    w.lex = ((n.x0 * g.L[1] + n.x1) * g.L[2] + n.x2) * g.L[3] + n.x3;
    w.pv = ((n.x0 + n.x1 + n.x2 + n.x3) & 1) ? g.Vh : 0;
    return w;
  }
  template <int S> __device__ __forceinline__ void move(const Geom& g, int d) {
    const bool f0 = d == 0, f1 = d == 1, f2 = d == 2, f3 = d == 3;
    const int L = (f0 ? g.L[0] : 0) + (f1 ? g.L[1] : 0) + (f2 ? g.L[2] : 0) + (f3 ? g.L[3] : 0);
    const int st = (f0 ? 2 * g.S0 : 0) + (f1 ? 2 * g.S1 : 0) + (f2 ? 2 * g.S2 : 0) + (f3 ? 1 : 0);
    int w;
    c = packed_move<S>(g, c, d, w);
    lex += S * st - w * (S * st * L);
    pv = g.Vh - pv;  // every move changes the parity (all extents are even)
  }
  __device__ __forceinline__ int index(const Geom&) const { return pv + (lex >> 1); }
};

// A site of the heatbath's staple walks: packed coordinates, the index in
// the heatbath layout (x3 = 4 q + r; site (((r L0 + x0) L1 + x1) L2 + x2) Q + q,
// Q = L3 / 4) recomputed from them at each load.
struct SiteHb {
  unsigned c;
  __device__ __forceinline__ static SiteHb of(const Geom&, const Site4& n) { return SiteHb{pack_site(n)}; }
  template <int S> __device__ __forceinline__ void move(const Geom& g, int d) {
    int w;
    c = packed_move<S>(g, c, d, w);
  }
  __device__ __forceinline__ int index(const Geom& g) const {
    const int x0 = (int)(c & 255u), x1 = (int)((c >> 8) & 255u), x2 = (int)((c >> 16) & 255u), x3 = (int)(c >> 24);
    return ((((x3 & 3) * g.L[0] + x0) * g.L[1] + x1) * g.L[2] + x2) * (g.L[3] >> 2) + (x3 >> 2);
  }
};

// Link layouts of the fused staple walks: the site type and the number of
// stored elements per link and direction.  EoLinks: the solver's even-odd
// order, nine elements; HbLinks: the heatbath layout, six elements when its
// links are kept as two rows.
struct EoLinks {
  using Site = SiteW;
  static constexpr int kStride = 9;
};
template <bool TWO> struct HbLinks {
  using Site = SiteHb;
  static constexpr int kStride = TWO ? 6 : 9;
};

// conj(a b - c d) with fused multiply-adds: one entry of the cross product
// that rebuilds the third row (or column) of an SU(3) matrix.
__device__ __forceinline__ C cross_conj(C a, C b, C c, C d) {
  double re = a.re * b.re, im = a.re * b.im;
  re = fma(-a.im, b.im, re);
  im = fma(a.im, b.re, im);
  re = fma(-c.re, d.re, re);
  im = fma(-c.re, d.im, im);
  re = fma(c.im, d.im, re);
  im = fma(-c.im, d.re, im);
  return {re, -im};
}

// Column 2 of an SU(3) matrix from columns 0 and 1: conj(col0 x col1).
__device__ __forceinline__ void rebuild_col2(M& t) {
  t.e[2] = cross_conj(t.e[3], t.e[7], t.e[6], t.e[4]);
  t.e[5] = cross_conj(t.e[6], t.e[1], t.e[0], t.e[7]);
  t.e[8] = cross_conj(t.e[0], t.e[4], t.e[3], t.e[1]);
}

// t <- t m and t <- t m^dag for SU(3) t (a staple path's partial product)
// and an SU(3) link m of which rows 0 and 1 are loaded: columns 0 and 1 of
// the product, then column 2 rebuilt from them.  t m^dag needs only rows 0
// and 1 of m; t m needs m's (2,0) and (2,1) entries, rebuilt from its rows
// 0 and 1.  96 and 112 FP64 instructions, against 108 for a full product
// plus 24 to rebuild the link's third row.
__device__ __forceinline__ void rmul_n_su3(M& t, const M& m) {
  const C m6 = cross_conj(m.e[1], m.e[5], m.e[2], m.e[4]);
  const C m7 = cross_conj(m.e[2], m.e[3], m.e[0], m.e[5]);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const C a0 = t.e[3 * i], a1 = t.e[3 * i + 1], a2 = t.e[3 * i + 2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const C b0 = m.e[j], b1 = m.e[3 + j], b2 = j == 0 ? m6 : m7;
      double re = a0.re * b0.re, im = a0.re * b0.im;
      re = fma(-a0.im, b0.im, re);
      im = fma(a0.im, b0.re, im);
      re = fma(a1.re, b1.re, re);
      re = fma(-a1.im, b1.im, re);
      im = fma(a1.re, b1.im, im);
      im = fma(a1.im, b1.re, im);
      re = fma(a2.re, b2.re, re);
      re = fma(-a2.im, b2.im, re);
      im = fma(a2.re, b2.im, im);
      im = fma(a2.im, b2.re, im);
      t.e[3 * i + j] = {re, im};
    }
  }
  rebuild_col2(t);
}
__device__ __forceinline__ void rmul_d_su3(M& t, const M& m) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const C a0 = t.e[3 * i], a1 = t.e[3 * i + 1], a2 = t.e[3 * i + 2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      // row j of m, conjugated
      const C b0 = m.e[3 * j], b1 = m.e[3 * j + 1], b2 = m.e[3 * j + 2];
      double re = a0.re * b0.re, im = -a0.re * b0.im;
      re = fma(a0.im, b0.im, re);
      im = fma(a0.im, b0.re, im);
      re = fma(a1.re, b1.re, re);
      re = fma(a1.im, b1.im, re);
      im = fma(-a1.re, b1.im, im);
      im = fma(a1.im, b1.re, im);
      re = fma(a2.re, b2.re, re);
      re = fma(a2.im, b2.im, re);
      im = fma(-a2.re, b2.im, im);
      im = fma(a2.im, b2.re, im);
      t.e[3 * i + j] = {re, im};
    }
  }
  rebuild_col2(t);
}

// Rows 0 and 1 of SU(3) link (i, d) (elements 0-5 of STRIDE stored per
// direction), loaded after `dep` as ldLs_after does.
template <int STRIDE = 9>
__device__ __forceinline__ M ldLs2_after(const double2* __restrict__ U, const Geom& g, int i, int d, double dep) {
  const size_t es = (size_t)g.V * sizeof(double2);
  const char* base = reinterpret_cast<const char*>(U) + (size_t)(d * STRIDE) * es;
  unsigned off = (unsigned)i * (unsigned)sizeof(double2);
  asm volatile("" : "+v"(off) : "v"(dep));
  M m;
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const double2 v = *reinterpret_cast<const double2*>(base + e * es + off);
    m.e[e] = {v.x, v.y};
  }
  return m;
}

// The link of the step from x along direction d with sign S: U_d(x) for
// S > 0, U_d(x - d)^dag for S < 0; x moves to x + S d.  t <- t L, or t = L
// for the first step of a path.  The sign is a compile-time constant: a
// runtime sign would make the compiler evaluate both branches.
// Software pipelining of a path with prefetch distance two: the load of
// step k waits for `dep.y`, an element of the partial product as it stood
// before step k - 2 multiplied (dep.x holds the one before step k - 1), so
// the links of steps k and k + 1 can be in flight while step k - 1
// multiplies, but no later link is loaded early.
// PURE: t is a product of links (SU(3) when R12), so with SU(3) links the
// product takes the column-rebuild forms above.
template <int S, bool R12, bool PURE = true, class LP = EoLinks>
__device__ __forceinline__ void mul_step(M& t, const double2* __restrict__ U, const Geom& g, typename LP::Site& x,
                                         int d, double2& dep) {
  constexpr bool kSu3 = R12 && PURE;
  constexpr int ST = LP::kStride;
  if constexpr (S > 0) {
    const M m = kSu3 ? ldLs2_after<ST>(U, g, x.index(g), d, dep.y) : ldLs_after<R12, ST>(U, g, x.index(g), d, dep.y);
    x.template move<+1>(g, d);
    dep.y = dep.x;
    dep.x = t.e[8].im;
    if constexpr (kSu3) rmul_n_su3(t, m);
    else rmul_nf(t, m);
  } else {
    x.template move<-1>(g, d);
    const M m = kSu3 ? ldLs2_after<ST>(U, g, x.index(g), d, dep.y) : ldLs_after<R12, ST>(U, g, x.index(g), d, dep.y);
    dep.y = dep.x;
    dep.x = t.e[8].im;
    if constexpr (kSu3) rmul_d_su3(t, m);
    else rmul_df(t, m);
  }
}
// the first link of a path
template <int S, bool R12, class LP = EoLinks>
__device__ __forceinline__ M first_step(const double2* __restrict__ U, const Geom& g, typename LP::Site& x, int d,
                                        double2& dep) {
  constexpr int ST = LP::kStride;
  if constexpr (S > 0) {
    const M m = ldLs_after<R12, ST>(U, g, x.index(g), d, dep.y);
    x.template move<+1>(g, d);
    return m;
  } else {
    x.template move<-1>(g, d);
    return mdag(ldLs_after<R12, ST>(U, g, x.index(g), d, dep.y));
  }
}

// a[k * stride] += s b (a matrix kept lane-interleaved in LDS, 18 reals)
__device__ __forceinline__ void lds_axpy(double* a, int stride, double s, const M& b) {
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    a[(2 * k) * stride] = fma(s, b.e[k].re, a[(2 * k) * stride]);
    a[(2 * k + 1) * stride] = fma(s, b.e[k].im, a[(2 * k + 1) * stride]);
  }
}

// One half of the staple sum of plane (mu, nu) (see staples_fused), added to
// A; WA1: also the half's plaquette staple a (nu, -S), added to the matrix
// at sA1 (LDS, stride sst).
template <int S, bool R12, class LP, bool WA1>
__device__ __forceinline__ void staple_half(M& A, double* sA1, int sst, const double2* __restrict__ U, const Geom& g,
                                            const typename LP::Site& nm, int mu, int nu, double c11, double c12) {
  using Site = typename LP::Site;
  double2 dep = make_double2(A.e[8].im, A.e[8].im);
  Site x = nm;
  M t = first_step<S, R12, LP>(U, g, x, nu, dep);
  mul_step<-1, R12, true, LP>(t, U, g, x, mu, dep);  // a, x = n + S nu
  const Site xa = x;
  if constexpr (WA1) {
    Site xp = xa;
    M P = t;
    mul_step<-S, R12, true, LP>(P, U, g, xp, nu, dep);
    lds_axpy(sA1, sst, 1.0, P);
  }
  M Z = mscale(c11, t);
  if (c12 != 0.0) {
    mul_step<-1, R12, true, LP>(t, U, g, x, mu, dep);
    mul_step<-S, R12, true, LP>(t, U, g, x, nu, dep);
    mul_step<+1, R12, true, LP>(t, U, g, x, mu, dep);
    maxpy(A, c12, t);  // link second
    x = nm;
    t = first_step<+1, R12, LP>(U, g, x, mu, dep);
    mul_step<S, R12, true, LP>(t, U, g, x, nu, dep);
    mul_step<-1, R12, true, LP>(t, U, g, x, mu, dep);
    mul_step<-1, R12, true, LP>(t, U, g, x, mu, dep);
    maxpy(Z, c12, t);  // link first, up to its last step
    x = nm;
    t = first_step<S, R12, LP>(U, g, x, nu, dep);
    mul_step<S, R12, true, LP>(t, U, g, x, nu, dep);
    mul_step<-1, R12, true, LP>(t, U, g, x, mu, dep);
    mul_step<-S, R12, true, LP>(t, U, g, x, nu, dep);
    maxpy(Z, c12, t);  // 2 x 1 in nu, up to its last step
  }
  x = xa;
  mul_step<-S, R12, false, LP>(Z, U, g, x, nu, dep);  // Z is a sum of paths, not SU(3)
#pragma unroll
  for (int k = 0; k < 9; ++k) A.e[k] = A.e[k] + Z.e[k];
}

// c11 A1 + c12 A2 for link (n, mu), A1 the six plaquette staples and A2 the
// eighteen rectangle staples (Lattice.cpp:249-507), with the paths of each
// plane grouped by shared factors.  The
// eight paths of plane (mu, nu) split into the four that leave n + mu
// towards +nu or end on U_nu(n)^dag (the plaquette above and rectangle types
// 2, 4 and 6) and the four mirrored ones towards -nu (below, types 1, 5 and
// 3); with S = +1 or -1 for the two halves and steps written (direction,
// sign) from n + mu:
//   plaquette   (nu, S) (mu, -) (nu, -S)
//   2 x 1 in nu (nu, S) (nu, S) (mu, -) (nu, -S) (nu, -S)
//   link first  (mu, +) (nu, S) (mu, -) (mu, -) (nu, -S)
//   link second (nu, S) (mu, -) (mu, -) (nu, -S) (mu, +)
// so with a = (nu, S) (mu, -), the half sums to
//   [c11 a + c12 (mu, +) (nu, S) (mu, -) (mu, -) + c12 (nu, S) (nu, S) (mu, -) (nu, -S)] (nu, -S)
//   + c12 a (mu, -) (nu, -S) (mu, +):
// 11 SU(3) products and 14 link loads per half instead of 14 and 18.  Every
// product keeps its path's left-to-right order; only the sums are regrouped,
// so the result equals the path-by-path sum to rounding.  mu is uniform over the
// workgroup and nu over the loop, so every link base stays scalar.  Every
// extent must be below 256 (packed coordinates).  LP: the link layout; WA1:
// the plaquette staple sum A1 (six paths, one more product per half) is
// also accumulated into the LDS matrix at sA1 (stride sst), which the
// caller zeroes.
template <bool R12 = false, class LP = EoLinks, bool WA1 = false>
__device__ __forceinline__ M staples_fused(const double2* __restrict__ U, const Geom& g, const Site4& n, int mu,
                                          double c11, double c12, double* sA1 = nullptr, int sst = 0) {
  M A = mzero<double>();
  typename LP::Site nm = LP::Site::of(g, n);
  nm.template move<+1>(g, mu);  // n + mu
  // one half per loop iteration, the two signs in loops of their own: a
  // branch between the halves inside one loop kept both halves' values live
#pragma unroll 1
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
    staple_half<+1, R12, LP, WA1>(A, sA1, sst, U, g, nm, mu, nu, c11, c12);
  }
#pragma unroll 1
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
    staple_half<-1, R12, LP, WA1>(A, sA1, sst, U, g, nm, mu, nu, c11, c12);
  }
  return A;
}

}  // namespace sfh
