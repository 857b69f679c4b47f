// Shared device/host definitions for the gfx950 staggered-fermion library.
//
// Device layout (structure of arrays, site-parity ordered):
//   device site index  i = p * Vh + cb,  p = (n0+n1+n2+n3) & 1,
//   cb = ((n0*L1 + n1)*L2 + n2)*(L3/2) + (n3 >> 1)
// For the reference's site order s = ((n0*L1+n1)*L2+n2)*L3+n3 this is
// cb = s >> 1 (L3 even), so converting between layouts is a gather.
//   links   : complex element e = 3a+b of U_mu at site i is at [(mu*9+e)*V + i]
//   spinors : colour c at site i is at [c*V + i]
// All lattice extents are even; time (direction 3) is antiperiodic for the
// fermions.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#define SFH_HD __host__ __device__ __forceinline__

namespace sfh {

struct Geom {
  int L[4];
  int V, Vh, L3h;
  int S2, S1, S0;  // cb strides of directions 2, 1, 0 (direction 3 moves cb by 0 or 1)
};

inline Geom make_geom(const int* d) {
  Geom g;
  for (int k = 0; k < 4; ++k) g.L[k] = d[k];
  g.V = d[0] * d[1] * d[2] * d[3];
  g.Vh = g.V / 2;
  g.L3h = d[3] / 2;
  g.S2 = g.L3h;
  g.S1 = g.L3h * d[2];
  g.S0 = g.L3h * d[2] * d[1];
  return g;
}

// First checkerboard index of the range of `half` sites that logical
// workgroup lb covers (on each parity).  When a column of fixed (n0, n1)
// (S1 checkerboard sites) splits into whole ranges, workgroups walk the
// columns fastest, in tiles of kTileX x kTileY columns (y fastest inside a
// tile, then y-tiles, then x-tiles), one z-slab of `half` sites per column
// at a time.  Inside a tile, y-neighbour ranges are adjacent in the walk and
// x-neighbour ranges kTileY workgroups apart, so the backward links and
// spinors they share are still in the XCD's L2 (4 MB) even in FP64, where a
// workgroup touches about 240 KB; in plain checkerboard order x-neighbours
// are L1 L2 L3 / (2 half) workgroups apart and their links come from HBM a
// second time.  Extents that do not tile fall back to the plain column walk
// or to checkerboard order.  Always a bijection onto [0, Vh).
constexpr int kTileX = 8, kTileY = 4;
__device__ __forceinline__ int range_start(const Geom& g, int lb, int half) {
  if (g.S1 % half == 0) {
    const int ncol = g.L[0] * g.L[1];
    const int chunk = lb / ncol;
    int col = lb - chunk * ncol;
    if (g.L[0] % kTileX == 0 && g.L[1] % kTileY == 0) {
      const int tile = col / (kTileX * kTileY), w = col - tile * (kTileX * kTileY);
      const int ytiles = g.L[1] / kTileY;
      const int tx = tile / ytiles, ty = tile - tx * ytiles;
      col = (tx * kTileX + w / kTileY) * g.L[1] + ty * kTileY + w % kTileY;
    }
    return col * g.S1 + chunk * half;
  }
  return lb * half;
}

// Coordinates of checkerboard index cb at parity p.
SFH_HD void cb_coords(const Geom& g, int cb, int p, int n[4]) {
  int t = cb;
  const int n3h = t % g.L3h;
  t /= g.L3h;
  n[2] = t % g.L[2];
  t /= g.L[2];
  n[1] = t % g.L[1];
  n[0] = t / g.L[1];
  n[3] = 2 * n3h + ((p + n[0] + n[1] + n[2]) & 1);
}

SFH_HD int cb_of(const Geom& g, const int n[4]) {
  return ((n[0] * g.L[1] + n[1]) * g.L[2] + n[2]) * g.L3h + (n[3] >> 1);
}

// Reference site index of device site (p, cb).
SFH_HD int ref_site(const Geom& g, int cb, int p) {
  int n[4];
  cb_coords(g, cb, p, n);
  return ((n[0] * g.L[1] + n[1]) * g.L[2] + n[2]) * g.L[3] + n[3];
}

// eta_mu(n) = (-1)^(n0 + ... + n_{mu-1})  (Fermions.cpp:47-52), as 0/1 sign bit.
SFH_HD int eta_bit(const int n[4], int mu) {
  int s = 0;
  for (int k = 0; k < mu; ++k) s += n[k];
  return s & 1;
}

// Workgroup id remapped so that each of the 8 XCDs (which receive workgroups
// round-robin) works on one contiguous range of sites: neighbours then share
// that XCD's L2.
__device__ __forceinline__ int xcd_block(int nb) {
  const int bid = blockIdx.x;
  const int x = bid & 7, k = bid >> 3;
  const int q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + k;
}

// ------------------------------------------------------------------ complex
template <typename R> struct cpx { R re, im; };

template <typename R> SFH_HD cpx<R> cmake(R a, R b) { return cpx<R>{a, b}; }
template <typename R> SFH_HD cpx<R> operator+(cpx<R> a, cpx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> SFH_HD cpx<R> operator-(cpx<R> a, cpx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> SFH_HD cpx<R> operator*(cpx<R> a, cpx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> SFH_HD cpx<R> operator*(R s, cpx<R> a) { return {s * a.re, s * a.im}; }
template <typename R> SFH_HD cpx<R> conj(cpx<R> a) { return {a.re, -a.im}; }
// a += b * c
template <typename R> SFH_HD void cmac(cpx<R>& a, cpx<R> b, cpx<R> c) {
  a.re += b.re * c.re - b.im * c.im;
  a.im += b.re * c.im + b.im * c.re;
}
// a += conj(b) * c
template <typename R> SFH_HD void cmacc(cpx<R>& a, cpx<R> b, cpx<R> c) {
  a.re += b.re * c.re + b.im * c.im;
  a.im += b.re * c.im - b.im * c.re;
}

template <typename R> struct mat3 { cpx<R> e[9]; };
template <typename R> struct vec3 { cpx<R> c[3]; };

template <typename R> SFH_HD mat3<R> mzero() {
  mat3<R> m;
  for (int i = 0; i < 9; ++i) m.e[i] = {R(0), R(0)};
  return m;
}
template <typename R> SFH_HD mat3<R> meye() {
  mat3<R> m = mzero<R>();
  m.e[0].re = m.e[4].re = m.e[8].re = R(1);
  return m;
}
template <typename R> SFH_HD mat3<R> mmul(const mat3<R>& a, const mat3<R>& b) {
  mat3<R> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<R> s = a.e[3 * i] * b.e[j];
      cmac(s, a.e[3 * i + 1], b.e[3 + j]);
      cmac(s, a.e[3 * i + 2], b.e[6 + j]);
      r.e[3 * i + j] = s;
    }
  return r;
}
// a * b^dagger
template <typename R> SFH_HD mat3<R> mmul_nd(const mat3<R>& a, const mat3<R>& b) {
  mat3<R> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<R> s = a.e[3 * i] * conj(b.e[3 * j]);
      s = s + a.e[3 * i + 1] * conj(b.e[3 * j + 1]);
      s = s + a.e[3 * i + 2] * conj(b.e[3 * j + 2]);
      r.e[3 * i + j] = s;
    }
  return r;
}
// a^dagger * b
template <typename R> SFH_HD mat3<R> mmul_dn(const mat3<R>& a, const mat3<R>& b) {
  mat3<R> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<R> s = conj(a.e[i]) * b.e[j];
      s = s + conj(a.e[3 + i]) * b.e[3 + j];
      s = s + conj(a.e[6 + i]) * b.e[6 + j];
      r.e[3 * i + j] = s;
    }
  return r;
}
template <typename R> SFH_HD mat3<R> mdag(const mat3<R>& a) {
  mat3<R> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.e[3 * i + j] = conj(a.e[3 * j + i]);
  return r;
}
template <typename R> SFH_HD mat3<R> madd(const mat3<R>& a, const mat3<R>& b) {
  mat3<R> r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.e[i] = a.e[i] + b.e[i];
  return r;
}
template <typename R> SFH_HD mat3<R> msub(const mat3<R>& a, const mat3<R>& b) {
  mat3<R> r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.e[i] = a.e[i] - b.e[i];
  return r;
}
template <typename R> SFH_HD mat3<R> mscale(R s, const mat3<R>& a) {
  mat3<R> r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.e[i] = s * a.e[i];
  return r;
}
template <typename R> SFH_HD mat3<R> mscalec(cpx<R> s, const mat3<R>& a) {
  mat3<R> r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.e[i] = s * a.e[i];
  return r;
}
template <typename R> SFH_HD cpx<R> mtrace(const mat3<R>& a) { return a.e[0] + a.e[4] + a.e[8]; }
// Re Tr(a b)
template <typename R> SFH_HD R retr_mul(const mat3<R>& a, const mat3<R>& b) {
  R s = R(0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) s += a.e[3 * i + k].re * b.e[3 * k + i].re - a.e[3 * i + k].im * b.e[3 * k + i].im;
  return s;
}

// ------------------------------------------------------------------ reductions
// Sum of a double over the 64 lanes of a wavefront (butterfly: every lane gets
// the same, order-fixed result).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum in a fixed order; result valid in every thread.  blockDim.x must
// be a multiple of 64 and at most 1024.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double wsum[16];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wsum[w] = v;
  __syncthreads();
  const int nw = blockDim.x >> 6;
  double t = 0.0;
  for (int k = 0; k < nw; ++k) t += wsum[k];
  __syncthreads();
  return t;
}

}  // namespace sfh
