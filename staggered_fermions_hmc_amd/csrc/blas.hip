// Fused CG vector kernels and FP64 reductions (Fermions.cpp:521-528,
// :588-606).  Every kernel works one lattice site (3 complex colours) per
// thread, so FP16 fields with one norm per site are updated in place.
// Reductions always accumulate in FP64 in a fixed order.
#include "internal.h"
#include "reduce.h"

namespace sfh {

constexpr int kB = 256;
static inline unsigned nblk(size_t n) { return (unsigned)((n + kB - 1) / kB); }

// ------------------------------------------------------------------ plain FP64 dot
__global__ __launch_bounds__(kB) void k_dot(const double* __restrict__ a, const double* __restrict__ b, size_t n,
                                            size_t per, double* partials, unsigned* cnt, double* out, int nb) {
  const size_t base = ((size_t)blockIdx.x * kB + threadIdx.x) * per;
  double s = 0.0;
  for (size_t k = 0; k < per; ++k) {
    const size_t j = base + k;
    if (j < n) s += a[j] * b[j];
  }
  grid_reduce(s, partials, cnt, nb, blockIdx.x, FinStore{out});
}

void dot_f64(hipStream_t s, const double* a, const double* b, size_t n, double* partials, unsigned* cnt,
             double* result) {
  size_t per = 4;
  while (nblk((n + per - 1) / per) > (unsigned)(kPartialSlot - kRedGroups)) per *= 2;
  const unsigned nb = nblk((n + per - 1) / per);
  hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kB), 0, s, a, b, n, per, partials, cnt, result, (int)nb);
}

// ------------------------------------------------------------------ FP64 / FP32 CG updates
// x += alpha p ; r -= alpha v ; reduce |r|^2 -> beta, convergence.
template <class S>
__global__ __launch_bounds__(kB) void k_cg_xr(int V, S* __restrict__ x, S* __restrict__ r, const S* __restrict__ p,
                                              const S* __restrict__ v, CGState* st, double* partials, int nb) {
  if (st->done) return;
  const int i = blockIdx.x * kB + threadIdx.x;
  const double alpha = st->alpha;
  double rr = 0.0;
  if (i < V) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t j = (size_t)c * V + i;
      S xx = x[j], rv = r[j];
      const S pp = p[j], vv = v[j];
      using R = decltype(xx.x);
      const R al = (R)alpha;
      xx.x += al * pp.x;
      xx.y += al * pp.y;
      rv.x -= al * vv.x;
      rv.y -= al * vv.y;
      x[j] = xx;
      r[j] = rv;
      rr += (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
    }
  }
  grid_reduce(rr, partials, ctr_of(st->ctr, CNT_RR), nb, blockIdx.x, FinRR{st});
}

// p = r + beta p
template <class S>
__global__ __launch_bounds__(kB) void k_cg_p(int V, S* __restrict__ p, const S* __restrict__ r, CGState* st) {
  if (st->done) return;
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= V) return;
  const double beta = st->beta;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t j = (size_t)c * V + i;
    S pp = p[j];
    const S rv = r[j];
    using R = decltype(pp.x);
    const R be = (R)beta;
    pp.x = rv.x + be * pp.x;
    pp.y = rv.y + be * pp.y;
    p[j] = pp;
  }
}

void cg_xr_f64(hipStream_t s, const Geom& g, double2* x, double2* r, const double2* p, const double2* v,
               CGState* st, double* partials) {
  const unsigned nb = nblk(g.V);
  hipLaunchKernelGGL(k_cg_xr<double2>, dim3(nb), dim3(kB), 0, s, g.V, x, r, p, v, st, partials, (int)nb);
}
void cg_p_f64(hipStream_t s, const Geom& g, double2* p, const double2* r, CGState* st) {
  hipLaunchKernelGGL(k_cg_p<double2>, dim3(nblk(g.V)), dim3(kB), 0, s, g.V, p, r, st);
}
void cg_p_f32(hipStream_t s, const Geom& g, float2* p, const float2* r, CGState* st) {
  hipLaunchKernelGGL(k_cg_p<float2>, dim3(nblk(g.V)), dim3(kB), 0, s, g.V, p, r, st);
}

// ------------------------------------------------------------------ FP16 (per-site norm) CG updates
__device__ __forceinline__ void st16(H16* f, int i, const float2 v[3]) {
  float m = 0.f;
#pragma unroll
  for (int c = 0; c < 3; ++c) m = fmaxf(m, fmaxf(fabsf(v[c].x), fabsf(v[c].y)));
  const float inv = m > 0.f ? 1.0f / m : 0.f;
  H16 s;
#pragma unroll
  for (int c = 0; c < 3; ++c) s.c[c] = __floats2half2_rn(v[c].x * inv, v[c].y * inv);
  s.n = m;
  f[i] = s;
}

// ------------------------------------------------------------------ precision conversion / FP64 BLAS
__global__ void k_f64_to_f32(const double2* __restrict__ in, float2* __restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) out[i] = make_float2((float)in[i].x, (float)in[i].y);
}
void f64_to_f32(hipStream_t s, const double2* in, float2* out, size_t n) {
  hipLaunchKernelGGL(k_f64_to_f32, dim3(nblk(n)), dim3(kB), 0, s, in, out, n);
}

__global__ void k_f64_to_f16(int V, const double2* __restrict__ in, H16* __restrict__ out) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= V) return;
  float2 v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double2 t = in[(size_t)c * V + i];
    v[c] = make_float2((float)t.x, (float)t.y);
  }
  st16(out, i, v);
}
void f64_to_f16(hipStream_t s, const Geom& g, const double2* in, H16* out) {
  hipLaunchKernelGGL(k_f64_to_f16, dim3(nblk(g.V)), dim3(kB), 0, s, g.V, in, out);
}

// Back to FP64: FP32 spinors and FP16 records (halves times the site's norm).
__global__ void k_f32_to_f64(const float2* __restrict__ in, double2* __restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) out[i] = make_double2((double)in[i].x, (double)in[i].y);
}
void f32_to_f64(hipStream_t s, const float2* in, double2* out, size_t n) {
  hipLaunchKernelGGL(k_f32_to_f64, dim3(nblk(n)), dim3(kB), 0, s, in, out, n);
}
__global__ void k_f16_to_f64(int V, const H16* __restrict__ in, double2* __restrict__ out) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= V) return;
  const H16 r = in[i];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float2 t = __half22float2(r.c[c]);
    out[(size_t)c * V + i] = make_double2((double)(r.n * t.x), (double)(r.n * t.y));
  }
}
void f16_to_f64(hipStream_t s, const Geom& g, const H16* in, double2* out) {
  hipLaunchKernelGGL(k_f16_to_f64, dim3(nblk(g.V)), dim3(kB), 0, s, g.V, in, out);
}

// FP16 records of (1 + i) in every colour (halves 1.0, norm 1.0).
__global__ void k_h16_ones(H16* __restrict__ f, int V) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= V) return;
  H16 s;
#pragma unroll
  for (int c = 0; c < 3; ++c) s.c[c] = __floats2half2_rn(1.0f, 1.0f);
  s.n = 1.0f;
  f[i] = s;
}
void h16_ones(hipStream_t s, H16* f, int V) { hipLaunchKernelGGL(k_h16_ones, dim3(nblk(V)), dim3(kB), 0, s, f, V); }

// The solve state copied to pinned host memory by one wave of dword stores:
// a kernel on the stream costs less than a device-to-host copy command.
__global__ __launch_bounds__(64) void k_state_out(const CGState* __restrict__ src, CGState* __restrict__ dst) {
  static_assert(sizeof(CGState) % 4 == 0 && sizeof(CGState) / 4 <= 64, "one dword per lane");
  const unsigned* a = reinterpret_cast<const unsigned*>(src);
  unsigned* b = reinterpret_cast<unsigned*>(dst);
  if (threadIdx.x < sizeof(CGState) / 4) b[threadIdx.x] = a[threadIdx.x];
}
void state_out(hipStream_t s, const CGState* d_state, CGState* h_state) {
  hipLaunchKernelGGL(k_state_out, dim3(1), dim3(64), 0, s, d_state, h_state);
}

__global__ void k_acc_f32(double2* __restrict__ xh, float2* __restrict__ xl, size_t n) {
  const size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const float2 l = xl[i];
  double2 h = xh[i];
  h.x += (double)l.x;
  h.y += (double)l.y;
  xh[i] = h;
  xl[i] = make_float2(0.f, 0.f);
}
void acc_f32_into_f64(hipStream_t s, double2* xh, float2* xl, size_t n) {
  hipLaunchKernelGGL(k_acc_f32, dim3(nblk(n)), dim3(kB), 0, s, xh, xl, n);
}

__global__ void k_xpay(double2* __restrict__ y, const double2* __restrict__ x, double a, size_t n) {
  const size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const double2 xv = x[i];
  double2 yv = y[i];
  yv.x = xv.x + a * yv.x;
  yv.y = xv.y + a * yv.y;
  y[i] = yv;
}
void xpay_f64(hipStream_t s, double2* y, const double2* x, double a, size_t n) {
  hipLaunchKernelGGL(k_xpay, dim3(nblk(n)), dim3(kB), 0, s, y, x, a, n);
}

__global__ void k_axpy(double2* __restrict__ y, const double2* __restrict__ x, double a, size_t n) {
  const size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const double2 xv = x[i];
  double2 yv = y[i];
  yv.x += a * xv.x;
  yv.y += a * xv.y;
  y[i] = yv;
}
void axpy_f64(hipStream_t s, double2* y, const double2* x, double a, size_t n) {
  hipLaunchKernelGGL(k_axpy, dim3(nblk(n)), dim3(kB), 0, s, y, x, a, n);
}

__global__ void k_sub(double2* __restrict__ out, const double2* __restrict__ a, const double2* __restrict__ b,
                      size_t n) {
  const size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const double2 av = a[i], bv = b[i];
  out[i] = make_double2(av.x - bv.x, av.y - bv.y);
}
void sub_f64(hipStream_t s, double2* out, const double2* a, const double2* b, size_t n) {
  hipLaunchKernelGGL(k_sub, dim3(nblk(n)), dim3(kB), 0, s, out, a, b, n);
}

}  // namespace sfh
