// Staggered hop kernels for gfx950 and the layout converters around them.
//
// One launch applies the operator to the whole lattice, both parities:
//   out(n) = a in(n) + (b/2) sum_mu eta_mu(n) [ bup(n,mu) U_mu(n) in(n+mu)
//                                             - bdn(n,mu) U_mu(n-mu)^dag in(n-mu) ]
// which is D (a=m, b=+1) or D^dagger (a=m, b=-1) of Fermions.cpp:218-288.
// Each workgroup takes the same checkerboard range on both parities (first
// half of its lanes even sites, second half odd), so a link loaded as the
// forward link of one site is reused from cache as the backward link of its
// neighbour.  Workgroups are remapped so each XCD sweeps a contiguous range,
// in an order that keeps x- and y-neighbour ranges close in time
// (range_start).
// The hop is HBM-bound (~0.9 flop/byte in FP64): no MFMA.
#include "hopcore.h"
#include "internal.h"
#include "reduce.h"

namespace sfh {

// UNR: directions processed per step of the mu loop (4: all links of the
// site in flight at once; 1 or 2: fewer registers, more waves per SIMD).
// R12: links read as two rows, the third rebuilt (SU(3) links only).
template <class T, bool DOT, int UNR, bool R12 = false>
__global__ __launch_bounds__(256) void k_hop(Geom g, const typename Prec<T>::L* __restrict__ U,
                                             const typename Prec<T>::S* __restrict__ in,
                                             const float* __restrict__ in_norm, typename Prec<T>::S* __restrict__ out,
                                             float* __restrict__ out_norm, typename Prec<T>::R a,
                                             typename Prec<T>::R hb, const typename Prec<T>::S* __restrict__ dw,
                                             const float* __restrict__ dw_norm, CGState* st, double* partials,
                                             int nb, int gate) {
  using R = typename Prec<T>::R;
  // gate 0: a CG step, skipped once the solve has stopped; gate 1: a
  // reliable-update step, run only while one is pending
  if (st != nullptr && (gate ? !st->ru_pending : st->done)) return;
  const int lb = xcd_block(nb);
  const int hw = blockDim.x >> 1;
  const int p = threadIdx.x >= hw;
  const int cb = range_start(g, lb, hw) + (threadIdx.x - p * hw);
  double dotv = 0.0;
  if (cb < g.Vh) {
    const int V = g.V;
    const int i = p * g.Vh + cb;
    const int q = (1 - p) * g.Vh;
    // FP64: the own site's input (and the dot's other operand) are loaded
    // first, in flight with the links and neighbours (16^4 FP64 CG 242 ->
    // 252 solves/s); the low-precision hops load them last, which keeps
    // their register count and measured the same or faster
    constexpr bool kEarly = std::is_same<T, double>::value;
    SpinorWords<T> sw, ww;
    if constexpr (kEarly) {
      sw.load(in, V, i);
      if constexpr (DOT) ww.load(dw, V, i);
    }
    int n[4];
    cb_coords(g, cb, p, n);
    vec3<R> acc;
#pragma unroll
    for (int c = 0; c < 3; ++c) acc.c[c] = {R(0), R(0)};
#pragma unroll UNR
    for (int mu = 0; mu < 4; ++mu) {
      const int esum = (mu > 0 ? n[0] : 0) + (mu > 1 ? n[1] : 0) + (mu > 2 ? n[2] : 0);
      const R eta = (esum & 1) ? R(-1) : R(1);
      int fw, bw;
      R sf = eta, sb = eta;
      if (mu == 3) {
        const int h = n[3] >> 1;
        const int up = n[3] + 1 == g.L[3] ? 0 : n[3] + 1;
        const int dn = n[3] == 0 ? g.L[3] - 1 : n[3] - 1;
        fw = cb - h + (up >> 1);
        bw = cb - h + (dn >> 1);
        if (n[3] == g.L[3] - 1) sf = -sf;
        if (n[3] == 0) sb = -sb;
      } else {
        // selects, not indexing: mu is not a constant when the loop is not unrolled
        const int S = mu == 0 ? g.S0 : (mu == 1 ? g.S1 : g.S2);
        const int L = mu == 0 ? g.L[0] : (mu == 1 ? g.L[1] : g.L[2]);
        const int nm = mu == 0 ? n[0] : (mu == 1 ? n[1] : n[2]);
        fw = cb + (nm + 1 == L ? -(L - 1) * S : S);
        bw = cb + (nm == 0 ? (L - 1) * S : -S);
      }
      if constexpr (kHalfR12<T, R12>) {
        hop_dir_h2(acc, U, in, V, mu, i, q + fw, q + bw, sf, sb);
      } else {
        mat3<R> Uf, Ub;
        vec3<R> vf, vb;
        load_link<T, R12>(U, V, mu, i, Uf);
        load_spinor<T>(in, in_norm, V, q + fw, vf);
        load_link<T, R12>(U, V, mu, q + bw, Ub);
        load_spinor<T>(in, in_norm, V, q + bw, vb);
        acc_mv(acc, Uf, vf, sf);
        acc_mdv(acc, Ub, vb, -sb);
      }
    }
    vec3<R> res;
    if constexpr (!kEarly) sw.load(in, V, i);
    if constexpr (kHalfR12<T, R12>) {
      // a n h + hb acc from the record's halves
      const H16& sr = sw.s;
      const R an = a * sr.n;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const f16x2 h = __builtin_bit_cast(f16x2, sr.c[c]);
        res.c[c].re = __builtin_fmaf(an, (float)h.x, hb * acc.c[c].re);
        res.c[c].im = __builtin_fmaf(an, (float)h.y, hb * acc.c[c].im);
      }
    } else {
      vec3<R> self;
      sw.decode(self);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        res.c[c].re = a * self.c[c].re + hb * acc.c[c].re;
        res.c[c].im = a * self.c[c].im + hb * acc.c[c].im;
      }
    }
    store_spinor<T>(out, out_norm, V, i, res);
    if constexpr (DOT) {
      if constexpr (!kEarly) ww.load(dw, V, i);
      vec3<R> w;
      ww.decode(w);
#pragma unroll
      for (int c = 0; c < 3; ++c) dotv += (double)w.c[c].re * (double)res.c[c].re + (double)w.c[c].im * (double)res.c[c].im;
    }
  }
  if constexpr (DOT) grid_reduce(dotv, partials, ctr_of(st->ctr, CNT_PV), nb, lb, FinPV{st});
}

// Four lanes per site, one per direction mu: each lane does the forward and
// backward hop of its direction, the four partial sums meet through two
// cross-lane exchanges, and lane mu < 3 writes colour mu.  Four times the
// waves of the one-lane form, for lattices too small to fill 256 CUs and
// for more loads in flight per CU.  Lane layout in a wavefront:
// lane = mu * 16 + site, so the 16 lanes of one direction load 16
// consecutive sites (coalesced).
template <class T, bool DOT>
__global__ __launch_bounds__(256) void k_hop4(Geom g, const typename Prec<T>::L* __restrict__ U,
                                              const typename Prec<T>::S* __restrict__ in,
                                              const float* __restrict__ in_norm, typename Prec<T>::S* __restrict__ out,
                                              float* __restrict__ out_norm, typename Prec<T>::R a,
                                              typename Prec<T>::R hb, const typename Prec<T>::S* __restrict__ dw,
                                              const float* __restrict__ dw_norm, CGState* st, double* partials,
                                              int nb, int gate) {
  using R = typename Prec<T>::R;
  // gate 0: a CG step, skipped once the solve has stopped; gate 1: a
  // reliable-update step, run only while one is pending
  if (st != nullptr && (gate ? !st->ru_pending : st->done)) return;
  const int lb = xcd_block(nb);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int mu = lane >> 4, sl = lane & 15;
  // 4 waves: waves 0,1 even sites, waves 2,3 odd sites, same cb range of 32
  const int p = wave >> 1;
  const int cb = lb * 32 + (wave & 1) * 16 + sl;
  const bool valid = cb < g.Vh;
  const int V = g.V;
  const int i = p * g.Vh + (valid ? cb : 0);
  const int q = (1 - p) * g.Vh;
  int n[4];
  cb_coords(g, valid ? cb : 0, p, n);
  vec3<R> acc;
#pragma unroll
  for (int c = 0; c < 3; ++c) acc.c[c] = {R(0), R(0)};
  if (valid) {
    const int es = (mu > 0 ? n[0] : 0) + (mu > 1 ? n[1] : 0) + (mu > 2 ? n[2] : 0);
    R sf = (es & 1) ? R(-1) : R(1), sb = sf;
    int fw, bw;
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
    mat3<R> Uf, Ub;
    vec3<R> vf, vb;
    load_link<T>(U, V, mu, i, Uf);
    load_spinor<T>(in, in_norm, V, q + fw, vf);
    load_link<T>(U, V, mu, q + bw, Ub);
    load_spinor<T>(in, in_norm, V, q + bw, vb);
    acc_mv(acc, Uf, vf, sf);
    acc_mdv(acc, Ub, vb, -sb);
  }
  // sum over the four direction lanes (xor 16, xor 32); fixed order
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    R re = acc.c[c].re, im = acc.c[c].im;
    re += __shfl_xor(re, 16, 64);
    im += __shfl_xor(im, 16, 64);
    re += __shfl_xor(re, 32, 64);
    im += __shfl_xor(im, 32, 64);
    acc.c[c] = {re, im};
  }
  double dotv = 0.0;
  if constexpr (std::is_same<T, Half>::value) {
    // FP16 output needs the whole site for its norm: lane mu == 0 stores it.
    if (valid && mu == 0) {
      vec3<R> self, res;
      load_spinor<T>(in, in_norm, V, i, self);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        res.c[c].re = a * self.c[c].re + hb * acc.c[c].re;
        res.c[c].im = a * self.c[c].im + hb * acc.c[c].im;
      }
      store_spinor<T>(out, out_norm, V, i, res);
      if constexpr (DOT) {
        vec3<R> w;
        load_spinor<T>(dw, dw_norm, V, i, w);
#pragma unroll
        for (int c = 0; c < 3; ++c)
          dotv += (double)w.c[c].re * (double)res.c[c].re + (double)w.c[c].im * (double)res.c[c].im;
      }
    }
  } else {
    if (valid && mu < 3) {
      using S = typename Prec<T>::S;
      const cpx<R> self = ldc(in + (size_t)mu * V + i);
      const cpx<R> ac = mu == 0 ? acc.c[0] : (mu == 1 ? acc.c[1] : acc.c[2]);
      const R re = a * self.re + hb * ac.re, im = a * self.im + hb * ac.im;
      out[(size_t)mu * V + i] = S{re, im};
      if constexpr (DOT) {
        const cpx<R> w = ldc(dw + (size_t)mu * V + i);
        dotv = (double)w.re * (double)re + (double)w.im * (double)im;
      }
    }
  }
  if constexpr (DOT) grid_reduce(dotv, partials, ctr_of(st->ctr, CNT_PV), nb, lb, FinPV{st});
}

constexpr int kHopBlock = 256;

static inline int hop_blocks(const Geom& g) { return (g.Vh + kHopBlock / 2 - 1) / (kHopBlock / 2); }
static inline int hop4_blocks(const Geom& g) { return (g.Vh + 31) / 32; }

// Variant 0: one lane per site, all four directions unrolled; 1: four lanes
// per site; 2, 3: one lane per site, the direction loop stepped one or two
// directions at a time.  r12 (low precision only, not the four-lane form):
// SU(3) links read as two rows.
template <class T>
static void hop_launch(hipStream_t s, int variant, bool r12, const Geom& g, const typename Prec<T>::L* U,
                       const typename Prec<T>::S* in, const float* in_norm, typename Prec<T>::S* out, float* out_norm,
                       typename Prec<T>::R a, typename Prec<T>::R hb, const typename Prec<T>::S* dw,
                       const float* dw_norm, CGState* st, double* partials, int gate) {
  if (variant == 1) {
    const int nb = hop4_blocks(g);
    if (dw)
      hipLaunchKernelGGL((k_hop4<T, true>), dim3(nb), dim3(256), 0, s, g, U, in, in_norm, out, out_norm, a, hb, dw,
                         dw_norm, st, partials, nb, gate);
    else
      hipLaunchKernelGGL((k_hop4<T, false>), dim3(nb), dim3(256), 0, s, g, U, in, in_norm, out, out_norm, a, hb,
                         dw, dw_norm, st, partials, nb, gate);
  } else {
    const int nb = hop_blocks(g);
    auto go = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(nb), dim3(kHopBlock), 0, s, g, U, in, in_norm, out, out_norm, a, hb, dw, dw_norm,
                         st, partials, nb, gate);
    };
    auto pick = [&](auto r) {
      constexpr bool R = decltype(r)::value;
      if (variant == 2) {
        if (dw) go(k_hop<T, true, 1, R>);
        else go(k_hop<T, false, 1, R>);
      } else if (variant == 3) {
        if (dw) go(k_hop<T, true, 2, R>);
        else go(k_hop<T, false, 2, R>);
      } else {
        if (dw) go(k_hop<T, true, 4, R>);
        else go(k_hop<T, false, 4, R>);
      }
    };
    if (r12) pick(std::true_type{});
    else pick(std::false_type{});
  }
}

static int g_hop_variant[3] = {-1, -1, -1};  // forced variant per precision, -1 = by size

static int pick_variant(int prec, const Geom& g) {
  if (g_hop_variant[prec] >= 0) return g_hop_variant[prec];
  // Measured on MI355X (16^4 and 32^4, all precisions): one lane per site is
  // faster; four lanes per site stays available through the tuner.
  (void)g;
  return 0;
}

void set_hop_variant(int prec, int variant) {
  if (prec >= 0 && prec < 3) g_hop_variant[prec] = variant;
}
int get_hop_variant(int prec, const Geom& g) { return pick_variant(prec, g); }

void hop_f64(hipStream_t s, const Geom& g, const double2* U, const double2* in, double2* out, double a, double b,
             const double2* dot_with, CGState* st, double* partials, int gate, bool r12) {
  hop_launch<double>(s, pick_variant(0, g), r12, g, U, in, nullptr, out, nullptr, a, 0.5 * b, dot_with, nullptr, st,
                     partials, gate);
}

void hop_f32(hipStream_t s, const Geom& g, const float2* U, const float2* in, float2* out, float a, float b,
             const float2* dot_with, CGState* st, double* partials, bool r12) {
  hop_launch<float>(s, pick_variant(1, g), r12, g, U, in, nullptr, out, nullptr, a, 0.5f * b, dot_with, nullptr, st,
                    partials, 0);
}

void hop_f16(hipStream_t s, const Geom& g, const __half2* U, const H16* in, H16* out, float a, float b,
             const H16* dot_with, CGState* st, double* partials, bool r12) {
  hop_launch<Half>(s, pick_variant(2, g), r12, g, U, in, nullptr, out, nullptr, a, 0.5f * b, dot_with, nullptr, st,
                   partials, 0);
}

// ------------------------------------------------------------------ layout converters
// One thread per device site; the reference site is computed from (p, cb).
__global__ void k_links_aos_to_soa(Geom g, const double* __restrict__ aos, double2* __restrict__ soa) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.V) return;
  const int p = i >= g.Vh, cb = i - p * g.Vh;
  const size_t s = (size_t)ref_site(g, cb, p);
  for (int mu = 0; mu < 4; ++mu)
    for (int e = 0; e < 9; ++e) {
      const double* src = aos + (s * 4 + mu) * 18 + 2 * e;
      soa[(size_t)(mu * 9 + e) * g.V + i] = make_double2(src[0], src[1]);
    }
}

__global__ void k_links_soa_to_aos(Geom g, const double2* __restrict__ soa, double* __restrict__ aos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.V) return;
  const int p = i >= g.Vh, cb = i - p * g.Vh;
  const size_t s = (size_t)ref_site(g, cb, p);
  for (int mu = 0; mu < 4; ++mu)
    for (int e = 0; e < 9; ++e) {
      const double2 v = soa[(size_t)(mu * 9 + e) * g.V + i];
      double* dst = aos + (s * 4 + mu) * 18 + 2 * e;
      dst[0] = v.x;
      dst[1] = v.y;
    }
}

__global__ void k_spinor_aos_to_soa(Geom g, const double* __restrict__ aos, double2* __restrict__ soa) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.V) return;
  const int p = i >= g.Vh, cb = i - p * g.Vh;
  const size_t s = (size_t)ref_site(g, cb, p);
  for (int c = 0; c < 3; ++c) soa[(size_t)c * g.V + i] = make_double2(aos[s * 6 + 2 * c], aos[s * 6 + 2 * c + 1]);
}

__global__ void k_spinor_soa_to_aos(Geom g, const double2* __restrict__ soa, double* __restrict__ aos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.V) return;
  const int p = i >= g.Vh, cb = i - p * g.Vh;
  const size_t s = (size_t)ref_site(g, cb, p);
  for (int c = 0; c < 3; ++c) {
    const double2 v = soa[(size_t)c * g.V + i];
    aos[s * 6 + 2 * c] = v.x;
    aos[s * 6 + 2 * c + 1] = v.y;
  }
}

__global__ void k_links_to_f32(size_t n, const double2* __restrict__ U, float2* __restrict__ u) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 v = U[i];
  u[i] = make_float2((float)v.x, (float)v.y);
}

// One lane per (site, direction): the nine elements in the packed FP16
// layout of Prec<Half> (records of elements 0-3, 4-5, 6-7 and 8).
__global__ void k_links_to_f16(Geom g, const double2* __restrict__ U, __half2* __restrict__ u) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (i >= g.V) return;
  unsigned w[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const double2 v = U[(size_t)(mu * 9 + e) * g.V + i];
    w[e] = __builtin_bit_cast(unsigned, __floats2half2_rn((float)v.x, (float)v.y));
  }
  char* base = reinterpret_cast<char*>(u) + (size_t)mu * 36 * g.V;
  reinterpret_cast<uint4*>(base)[i] = make_uint4(w[0], w[1], w[2], w[3]);
  reinterpret_cast<uint2*>(base + (size_t)16 * g.V)[i] = make_uint2(w[4], w[5]);
  reinterpret_cast<uint2*>(base + (size_t)24 * g.V)[i] = make_uint2(w[6], w[7]);
  reinterpret_cast<unsigned*>(base + (size_t)32 * g.V)[i] = w[8];
}

// One lane per link: |row2 - conj(row0 x row1)| in FP64, the block maximum
// folded into *dev by an integer max on the float's bits (order-preserving
// for non-negative floats; NaN compares above every finite value).
__global__ __launch_bounds__(256) void k_su3_dev(Geom g, const double2* __restrict__ U, unsigned* dev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  float d = 0.f;
  if (i < g.V) {
    cpx<double> m[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) m[e] = ldc(U + (size_t)(mu * 9 + e) * g.V + i);
    const cpx<double> r[3] = {conj(m[1] * m[5] - m[2] * m[4]), conj(m[2] * m[3] - m[0] * m[5]),
                              conj(m[0] * m[4] - m[1] * m[3])};
    // fmax drops NaN operands, so non-finite entries are flagged apart
    double x = 0.0;
    bool finite = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const cpx<double> e = m[6 + k] - r[k];
      finite = finite && isfinite(e.re) && isfinite(e.im);
      x = fmax(x, fmax(fabs(e.re), fabs(e.im)));
    }
    d = finite ? (float)x : __int_as_float(0x7f800000);
  }
  __shared__ float w[4];
  for (int o = 32; o > 0; o >>= 1) d = fmaxf(d, __shfl_xor(d, o));
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = 0.f;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) b = fmaxf(b, w[k]);
    atomicMax(dev, __float_as_uint(b));
  }
}

static inline unsigned grid1(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

void links_aos_to_soa(hipStream_t s, const Geom& g, const double* aos, double2* soa) {
  hipLaunchKernelGGL(k_links_aos_to_soa, dim3(grid1(g.V, 256)), dim3(256), 0, s, g, aos, soa);
}
void links_soa_to_aos(hipStream_t s, const Geom& g, const double2* soa, double* aos) {
  hipLaunchKernelGGL(k_links_soa_to_aos, dim3(grid1(g.V, 256)), dim3(256), 0, s, g, soa, aos);
}
void spinor_aos_to_soa(hipStream_t s, const Geom& g, const double* aos, double2* soa) {
  hipLaunchKernelGGL(k_spinor_aos_to_soa, dim3(grid1(g.V, 256)), dim3(256), 0, s, g, aos, soa);
}
void spinor_soa_to_aos(hipStream_t s, const Geom& g, const double2* soa, double* aos) {
  hipLaunchKernelGGL(k_spinor_soa_to_aos, dim3(grid1(g.V, 256)), dim3(256), 0, s, g, soa, aos);
}
void links_to_f32(hipStream_t s, const Geom& g, const double2* U, float2* u32) {
  const size_t n = (size_t)36 * g.V;
  hipLaunchKernelGGL(k_links_to_f32, dim3(grid1(n, 256)), dim3(256), 0, s, n, U, u32);
}
void su3_deviation(hipStream_t s, const Geom& g, const double2* U, unsigned* dev) {
  hipLaunchKernelGGL(k_su3_dev, dim3(grid1(g.V, 256), 4), dim3(256), 0, s, g, U, dev);
}

void links_to_f16(hipStream_t s, const Geom& g, const double2* U, __half2* u16) {
  hipLaunchKernelGGL(k_links_to_f16, dim3(grid1(g.V, 256), 4), dim3(256), 0, s, g, U, u16);
}

}  // namespace sfh
