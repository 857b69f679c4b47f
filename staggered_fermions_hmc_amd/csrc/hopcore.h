// Per-site building blocks of the staggered hop shared by the single-system
// and batched kernels: storage precisions, link / spinor loads and stores in
// the SoA-parity layout, the two 3x3 complex mat-vec accumulations, and the
// neighbour index and sign of one direction.
#pragma once

#include <type_traits>

#include "common.h"

namespace sfh {

struct Half {};

// FP16 spinor of one site: three complex halves scaled by the site's
// max-norm, and that norm, in one 16-byte record (one load, one store).
struct alignas(16) H16 {
  __half2 c[3];
  float n;
};

template <class T> struct Prec;
template <> struct Prec<double> { using R = double; using L = double2; using S = double2; };
template <> struct Prec<float> { using R = float; using L = float2; using S = float2; };
// FP16 links: per direction mu a block of 36 V bytes holding, for each site,
// elements 0-3 (16-byte record), 4-5 (8 bytes), 6-7 (8 bytes) and 8 (4 bytes)
// in four arrays, so a two-row read is two loads and a full read four.
template <> struct Prec<Half> { using R = float; using L = __half2; using S = H16; };

// FP16 records with two-row FP16 links: the hop terms of hop_dir_h2.
template <class T, bool R12> constexpr bool kHalfR12 = std::is_same<T, Half>::value && R12;

// Spinor storage elements per site: three complex, or one FP16 record.
template <class T> constexpr int kSpinorElems = std::is_same<T, Half>::value ? 1 : 3;

__device__ __forceinline__ cpx<double> ldc(const double2* p) {
  const double2 v = *p;
  return {v.x, v.y};
}
__device__ __forceinline__ cpx<float> ldc(const float2* p) {
  const float2 v = *p;
  return {v.x, v.y};
}
__device__ __forceinline__ cpx<float> ldc(const __half2* p) {
  const float2 v = __half22float2(*p);
  return {v.x, v.y};
}

__device__ __forceinline__ cpx<float> h2c(unsigned w) {
  const float2 v = __half22float2(__builtin_bit_cast(__half2, w));
  return {v.x, v.y};
}

// FP32 complex arithmetic as chains of fused multiply-adds: the plain
// expressions compile to separate multiplies, adds and subtracts (252, 102
// and 54 of the FP32 hop's 944 VALU instructions per wave), and at small
// lattices, one wave per SIMD, a wave's own instruction stream is on the
// critical path.
// conj(a b - c d)
__device__ __forceinline__ cpx<float> cross_conj_f(cpx<float> a, cpx<float> b, cpx<float> c, cpx<float> d) {
  float re = __builtin_fmaf(a.re, b.re, 0.0f);
  re = __builtin_fmaf(-a.im, b.im, re);
  re = __builtin_fmaf(-c.re, d.re, re);
  re = __builtin_fmaf(c.im, d.im, re);
  float im = __builtin_fmaf(a.re, b.im, 0.0f);
  im = __builtin_fmaf(a.im, b.re, im);
  im = __builtin_fmaf(-c.re, d.im, im);
  im = __builtin_fmaf(-c.im, d.re, im);
  return {re, -im};
}
// t += a b, t += conj(a) b
__device__ __forceinline__ void mac_f(cpx<float>& t, cpx<float> a, cpx<float> b) {
  t.re = __builtin_fmaf(a.re, b.re, t.re);
  t.re = __builtin_fmaf(-a.im, b.im, t.re);
  t.im = __builtin_fmaf(a.re, b.im, t.im);
  t.im = __builtin_fmaf(a.im, b.re, t.im);
}
__device__ __forceinline__ void macc_f(cpx<float>& t, cpx<float> a, cpx<float> b) {
  t.re = __builtin_fmaf(a.re, b.re, t.re);
  t.re = __builtin_fmaf(a.im, b.im, t.re);
  t.im = __builtin_fmaf(a.re, b.im, t.im);
  t.im = __builtin_fmaf(-a.im, b.re, t.im);
}

// R12: read the first two rows only and rebuild the third as
// conj(row0 x row1), exact for SU(3) links (two thirds of the link bytes).
template <class T, bool R12 = false>
__device__ __forceinline__ void load_link(const typename Prec<T>::L* __restrict__ U, int V, int mu, int i,
                                          mat3<typename Prec<T>::R>& m) {
  if constexpr (std::is_same<T, Half>::value) {
    const char* base = reinterpret_cast<const char*>(U) + (size_t)mu * 36 * V;
    const uint4 a = reinterpret_cast<const uint4*>(base)[i];
    const uint2 b = reinterpret_cast<const uint2*>(base + (size_t)16 * V)[i];
    m.e[0] = h2c(a.x);
    m.e[1] = h2c(a.y);
    m.e[2] = h2c(a.z);
    m.e[3] = h2c(a.w);
    m.e[4] = h2c(b.x);
    m.e[5] = h2c(b.y);
    if constexpr (!R12) {
      const uint2 c = reinterpret_cast<const uint2*>(base + (size_t)24 * V)[i];
      const unsigned d = reinterpret_cast<const unsigned*>(base + (size_t)32 * V)[i];
      m.e[6] = h2c(c.x);
      m.e[7] = h2c(c.y);
      m.e[8] = h2c(d);
    }
  } else {
#pragma unroll
    for (int e = 0; e < (R12 ? 6 : 9); ++e) m.e[e] = ldc(U + (size_t)(mu * 9 + e) * V + i);
  }
  if constexpr (R12) {
    if constexpr (std::is_same<typename Prec<T>::R, float>::value) {
      m.e[6] = cross_conj_f(m.e[1], m.e[5], m.e[2], m.e[4]);
      m.e[7] = cross_conj_f(m.e[2], m.e[3], m.e[0], m.e[5]);
      m.e[8] = cross_conj_f(m.e[0], m.e[4], m.e[1], m.e[3]);
    } else {
      m.e[6] = conj(m.e[1] * m.e[5] - m.e[2] * m.e[4]);
      m.e[7] = conj(m.e[2] * m.e[3] - m.e[0] * m.e[5]);
      m.e[8] = conj(m.e[0] * m.e[4] - m.e[1] * m.e[3]);
    }
  }
}

// ---- FP16 hop terms without conversion instructions.
// In FP16 storage the hop is bound by vector-ALU issue, not bytes: with each
// half converted to FP32 before it is multiplied, and every neighbour spinor
// scaled by its norm component by component, one site cost 1105 VALU
// instructions per wave.  Here the stored halves go straight into
// mixed-precision fused multiply-adds (v_fma_mix_f32: FP16 operands, FP32
// product and accumulation; a product of two halves is exact in FP32), and
// a neighbour's norm multiplies the 3-vector U h once per link.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ f16x2 as_h2(unsigned w) { return __builtin_bit_cast(f16x2, w); }

// A two-row FP16 link as stored (rows 0 and 1, six complex halves) and its
// third row rebuilt in FP32, conj(row0 x row1).
struct LinkH2 {
  f16x2 u[6];
  cpx<float> r2[3];
};

// conj(a b - c d) for complex halves, in FP32 (the zero addend keeps the
// first product a fused multiply-add of halves)
__device__ __forceinline__ cpx<float> cross_conj_h(f16x2 a, f16x2 b, f16x2 c, f16x2 d) {
  float re = __builtin_fmaf((float)a.x, (float)b.x, 0.0f);
  re = __builtin_fmaf(-(float)a.y, (float)b.y, re);
  re = __builtin_fmaf(-(float)c.x, (float)d.x, re);
  re = __builtin_fmaf((float)c.y, (float)d.y, re);
  float im = __builtin_fmaf((float)a.x, (float)b.y, 0.0f);
  im = __builtin_fmaf((float)a.y, (float)b.x, im);
  im = __builtin_fmaf(-(float)c.x, (float)d.y, im);
  im = __builtin_fmaf(-(float)c.y, (float)d.x, im);
  return {re, -im};
}

__device__ __forceinline__ void load_link_h2(const __half2* __restrict__ U, int V, int mu, int i, LinkH2& m) {
  const char* base = reinterpret_cast<const char*>(U) + (size_t)mu * 36 * V;
  const uint4 a = reinterpret_cast<const uint4*>(base)[i];
  const uint2 b = reinterpret_cast<const uint2*>(base + (size_t)16 * V)[i];
  m.u[0] = as_h2(a.x);
  m.u[1] = as_h2(a.y);
  m.u[2] = as_h2(a.z);
  m.u[3] = as_h2(a.w);
  m.u[4] = as_h2(b.x);
  m.u[5] = as_h2(b.y);
  m.r2[0] = cross_conj_h(m.u[1], m.u[5], m.u[2], m.u[4]);
  m.r2[1] = cross_conj_h(m.u[2], m.u[3], m.u[0], m.u[5]);
  m.r2[2] = cross_conj_h(m.u[0], m.u[4], m.u[1], m.u[3]);
}

// t += a b and t += conj(a) b, a an FP16 or FP32 complex, b a complex half
__device__ __forceinline__ void mac_h(cpx<float>& t, f16x2 a, f16x2 b) {
  t.re = __builtin_fmaf((float)a.x, (float)b.x, t.re);
  t.re = __builtin_fmaf(-(float)a.y, (float)b.y, t.re);
  t.im = __builtin_fmaf((float)a.x, (float)b.y, t.im);
  t.im = __builtin_fmaf((float)a.y, (float)b.x, t.im);
}
__device__ __forceinline__ void mac_h(cpx<float>& t, cpx<float> a, f16x2 b) {
  t.re = __builtin_fmaf(a.re, (float)b.x, t.re);
  t.re = __builtin_fmaf(-a.im, (float)b.y, t.re);
  t.im = __builtin_fmaf(a.re, (float)b.y, t.im);
  t.im = __builtin_fmaf(a.im, (float)b.x, t.im);
}
__device__ __forceinline__ void macc_h(cpx<float>& t, f16x2 a, f16x2 b) {
  t.re = __builtin_fmaf((float)a.x, (float)b.x, t.re);
  t.re = __builtin_fmaf((float)a.y, (float)b.y, t.re);
  t.im = __builtin_fmaf((float)a.x, (float)b.y, t.im);
  t.im = __builtin_fmaf(-(float)a.y, (float)b.x, t.im);
}
__device__ __forceinline__ void macc_h(cpx<float>& t, cpx<float> a, f16x2 b) {
  t.re = __builtin_fmaf(a.re, (float)b.x, t.re);
  t.re = __builtin_fmaf(a.im, (float)b.y, t.re);
  t.im = __builtin_fmaf(a.re, (float)b.y, t.im);
  t.im = __builtin_fmaf(-a.im, (float)b.x, t.im);
}

__device__ __forceinline__ void acc_add(vec3<float>& acc, int a, float sn, cpx<float> t) {
  acc.c[a].re = __builtin_fmaf(sn, t.re, acc.c[a].re);
  acc.c[a].im = __builtin_fmaf(sn, t.im, acc.c[a].im);
}

// acc += s U v for the spinor v = n h of FP16 record r (h its halves, n its
// norm): U h is formed from the halves, then scaled by s n.
__device__ __forceinline__ void acc_mv_h(vec3<float>& acc, const LinkH2& U, const H16& r, float s) {
  const f16x2 h0 = __builtin_bit_cast(f16x2, r.c[0]), h1 = __builtin_bit_cast(f16x2, r.c[1]),
              h2 = __builtin_bit_cast(f16x2, r.c[2]);
  const float sn = s * r.n;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    cpx<float> t = {0.0f, 0.0f};
    mac_h(t, U.u[3 * a], h0);
    mac_h(t, U.u[3 * a + 1], h1);
    mac_h(t, U.u[3 * a + 2], h2);
    acc_add(acc, a, sn, t);
  }
  cpx<float> t = {0.0f, 0.0f};
  mac_h(t, U.r2[0], h0);
  mac_h(t, U.r2[1], h1);
  mac_h(t, U.r2[2], h2);
  acc_add(acc, 2, sn, t);
}
// acc += s U^dag v, v = n h as above: (U^dag h)_a = sum_b conj(U_ba) h_b.
__device__ __forceinline__ void acc_mdv_h(vec3<float>& acc, const LinkH2& U, const H16& r, float s) {
  const f16x2 h0 = __builtin_bit_cast(f16x2, r.c[0]), h1 = __builtin_bit_cast(f16x2, r.c[1]),
              h2 = __builtin_bit_cast(f16x2, r.c[2]);
  const float sn = s * r.n;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    cpx<float> t = {0.0f, 0.0f};
    macc_h(t, U.u[a], h0);
    macc_h(t, U.u[3 + a], h1);
    macc_h(t, U.r2[a], h2);
    acc_add(acc, a, sn, t);
  }
}

// One direction of the hop for FP16 records and two-row FP16 links:
// acc += sf U_mu(i) in(f) - sb U_mu(b)^dag in(b), the loads issued first.
__device__ __forceinline__ void hop_dir_h2(vec3<float>& acc, const __half2* __restrict__ U,
                                           const H16* __restrict__ in, int V, int mu, int i, int f, int b, float sf,
                                           float sb) {
  LinkH2 Uf, Ub;
  load_link_h2(U, V, mu, i, Uf);
  const H16 vf = in[f];
  load_link_h2(U, V, mu, b, Ub);
  const H16 vb = in[b];
  acc_mv_h(acc, Uf, vf, sf);
  acc_mdv_h(acc, Ub, vb, -sb);
}

// nrm: unused (the FP16 record carries its norm); kept so one call serves
// every precision.
template <class T>
__device__ __forceinline__ void load_spinor(const typename Prec<T>::S* __restrict__ f, const float* __restrict__ nrm,
                                            int V, int i, vec3<typename Prec<T>::R>& v) {
  (void)nrm;
  if constexpr (std::is_same<T, Half>::value) {
    const H16 s = f[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float2 t = __half22float2(s.c[c]);
      v.c[c] = {s.n * t.x, s.n * t.y};
    }
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) v.c[c] = ldc(f + (size_t)c * V + i);
  }
}

template <class T>
__device__ __forceinline__ void store_spinor(typename Prec<T>::S* __restrict__ f, float* __restrict__ nrm, int V,
                                             int i, const vec3<typename Prec<T>::R>& v) {
  (void)nrm;
  if constexpr (std::is_same<T, Half>::value) {
    float m = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) m = fmaxf(m, fmaxf(fabsf(v.c[c].re), fabsf(v.c[c].im)));
    const float inv = m > 0.f ? 1.0f / m : 0.f;
    H16 s;
#pragma unroll
    for (int c = 0; c < 3; ++c) s.c[c] = __floats2half2_rn(v.c[c].re * inv, v.c[c].im * inv);
    s.n = m;
    f[i] = s;
  } else {
    using S = typename Prec<T>::S;
#pragma unroll
    for (int c = 0; c < 3; ++c) f[(size_t)c * V + i] = S{v.c[c].re, v.c[c].im};
  }
}

// Storage words of one site's spinor, loaded now and decoded later: a kernel
// can issue the loads before a wait on something else (a barrier, the solve
// state) and convert after it, without the conversion pulling the wait for
// the data forward.
template <class T> struct SpinorWords {
  cpx<typename Prec<T>::R> c[3];
  __device__ __forceinline__ void load(const typename Prec<T>::S* __restrict__ f, int V, int i) {
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = ldc(f + (size_t)k * V + i);
  }
  __device__ __forceinline__ void decode(vec3<typename Prec<T>::R>& v) const {
#pragma unroll
    for (int k = 0; k < 3; ++k) v.c[k] = c[k];
  }
};
template <> struct SpinorWords<Half> {
  H16 s;
  __device__ __forceinline__ void load(const H16* __restrict__ f, int V, int i) {
    (void)V;
    s = f[i];
  }
  __device__ __forceinline__ void decode(vec3<float>& v) const {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float2 t = __half22float2(s.c[k]);
      v.c[k] = {s.n * t.x, s.n * t.y};
    }
  }
};

// acc += s * U v
template <class R>
__device__ __forceinline__ void acc_mv(vec3<R>& acc, const mat3<R>& U, const vec3<R>& v, R s) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if constexpr (std::is_same<R, float>::value) {
      cpx<float> t = {0.0f, 0.0f};
      mac_f(t, U.e[3 * a], v.c[0]);
      mac_f(t, U.e[3 * a + 1], v.c[1]);
      mac_f(t, U.e[3 * a + 2], v.c[2]);
      acc.c[a].re = __builtin_fmaf(s, t.re, acc.c[a].re);
      acc.c[a].im = __builtin_fmaf(s, t.im, acc.c[a].im);
    } else {
      cpx<R> t = U.e[3 * a] * v.c[0];
      cmac(t, U.e[3 * a + 1], v.c[1]);
      cmac(t, U.e[3 * a + 2], v.c[2]);
      acc.c[a].re += s * t.re;
      acc.c[a].im += s * t.im;
    }
  }
}
// acc += s * U^dag v
template <class R>
__device__ __forceinline__ void acc_mdv(vec3<R>& acc, const mat3<R>& U, const vec3<R>& v, R s) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if constexpr (std::is_same<R, float>::value) {
      cpx<float> t = {0.0f, 0.0f};
      macc_f(t, U.e[a], v.c[0]);
      macc_f(t, U.e[3 + a], v.c[1]);
      macc_f(t, U.e[6 + a], v.c[2]);
      acc.c[a].re = __builtin_fmaf(s, t.re, acc.c[a].re);
      acc.c[a].im = __builtin_fmaf(s, t.im, acc.c[a].im);
    } else {
      cpx<R> t = conj(U.e[a]) * v.c[0];
      cmacc(t, U.e[3 + a], v.c[1]);
      cmacc(t, U.e[6 + a], v.c[2]);
      acc.c[a].re += s * t.re;
      acc.c[a].im += s * t.im;
    }
  }
}

// Checkerboard indices (in the opposite parity) of the forward and backward
// neighbours of site (cb, coordinates n) in direction mu, and the signs
// eta_mu(n) bup(n, mu) and eta_mu(n) bdn(n, mu) (Fermions.cpp:47-87).
template <class R>
__device__ __forceinline__ void neighbours(const Geom& g, int cb, const int n[4], int mu, int& fw, int& bw, R& sf,
                                           R& sb) {
  const int es = (mu > 0 ? n[0] : 0) + (mu > 1 ? n[1] : 0) + (mu > 2 ? n[2] : 0);
  sf = (es & 1) ? R(-1) : R(1);
  sb = sf;
  if (mu == 3) {
    const int h = n[3] >> 1;
    const int up = n[3] + 1 == g.L[3] ? 0 : n[3] + 1;
    const int dn = n[3] == 0 ? g.L[3] - 1 : n[3] - 1;
    fw = cb - h + (up >> 1);
    bw = cb - h + (dn >> 1);
    if (n[3] == g.L[3] - 1) sf = -sf;
    if (n[3] == 0) sb = -sb;
  } else {
    const int S = mu == 0 ? g.S0 : (mu == 1 ? g.S1 : g.S2);
    const int L = mu == 0 ? g.L[0] : (mu == 1 ? g.L[1] : g.L[2]);
    const int nm = mu == 0 ? n[0] : (mu == 1 ? n[1] : n[2]);
    fw = cb + (nm + 1 == L ? -(L - 1) * S : S);
    bw = cb + (nm == 0 ? (L - 1) * S : -S);
  }
}

}  // namespace sfh
