// Two-kernel iteration of the mixed-precision CG on (D D^dag) x = phi
// (Fermions.cpp:553-617 in FP32 or FP16 storage, with FP64 reliable updates).
//
// The textbook iteration needs four passes (two hops, x/r update, p update)
// because alpha and beta are global reductions.  Here each iteration is two
// kernels:
//   A_k: r_k = r_{k-1} - alpha_{k-1} v_{k-1},  p_k = r_k + beta_{k-1} p_{k-1},
//        x += alpha_{k-1} p_{k-1} at the own site; r_k and p_k recomputed at
//        the 8 neighbours from their stored r, v, p; then t = D^dag p_k.
//   B_k: v_k = D t and, in one grid reduction, <p_k, v_k>, <r_k, v_k>,
//        <v_k, v_k>, <r_k, r_k>; the finisher forms alpha_k = |r_k|^2 /
//        <p,v> and |r_{k+1}|^2 = |r_k|^2 - 2 alpha <r,v> + alpha^2 <v,v>.
// A neighbour recomputes exactly what the owning site stores (same inputs,
// same fused multiply-adds, same FP16 rounding), so every site sees one
// consistent p_k.  Kernel launches per iteration drop from four to two and
// the r, p and v vectors are read once per iteration instead of twice.
#include "mixed.h"
#include "reduce.h"

namespace sfh {

namespace {

constexpr int kBlk = 256;

// FP16 storage of one site: halves scaled by the site's max-norm.  v is
// replaced by the value that a later load of the stored site returns.
struct Q16 {
  __half2 h[3];
  float m;
};
__device__ __forceinline__ void quant16(vec3<float>& v, Q16& q) {
  float m = 0.f;
#pragma unroll
  for (int c = 0; c < 3; ++c) m = fmaxf(m, fmaxf(fabsf(v.c[c].re), fabsf(v.c[c].im)));
  const float inv = m > 0.f ? 1.0f / m : 0.f;
  q.m = m;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    q.h[c] = __floats2half2_rn(v.c[c].re * inv, v.c[c].im * inv);
    const float2 t = __half22float2(q.h[c]);
    v.c[c] = {m * t.x, m * t.y};
  }
}

template <class T> constexpr bool is_half() { return std::is_same<T, Half>::value; }

// r <- r - al v ; p <- r + be p, each rounded to storage as the owner stores it.
template <class T>
__device__ __forceinline__ void advance(vec3<float>& r, const vec3<float>& v, vec3<float>& p, float al, float be,
                                        Q16& qr, Q16& qp) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    r.c[c].re = __fmaf_rn(-al, v.c[c].re, r.c[c].re);
    r.c[c].im = __fmaf_rn(-al, v.c[c].im, r.c[c].im);
  }
  if constexpr (is_half<T>()) quant16(r, qr);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    p.c[c].re = __fmaf_rn(be, p.c[c].re, r.c[c].re);
    p.c[c].im = __fmaf_rn(be, p.c[c].im, r.c[c].im);
  }
  if constexpr (is_half<T>()) quant16(p, qp);
}

template <class T>
__device__ __forceinline__ void store_q(typename Prec<T>::S* f, float* nrm, int V, int i, const vec3<float>& v,
                                        const Q16& q) {
  (void)nrm;
  if constexpr (is_half<T>()) {
    H16 s;
#pragma unroll
    for (int c = 0; c < 3; ++c) s.c[c] = q.h[c];
    s.n = q.m;
    f[i] = s;
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) f[(size_t)c * V + i] = make_float2(v.c[c].re, v.c[c].im);
  }
}

// Site of this lane: each workgroup takes one checkerboard range on both
// parities (as the single-system hop kernel), workgroups remapped per XCD.
struct Site {
  int lb, cb, p, i, q;
};
__device__ __forceinline__ Site site_of(const Geom& g, int nb) {
  Site s;
  s.lb = xcd_block(nb);
  const int hw = blockDim.x >> 1;
  s.p = threadIdx.x >= hw;
  s.cb = range_start(g, s.lb, hw) + (threadIdx.x - s.p * hw);
  s.i = s.p * g.Vh + s.cb;
  s.q = (1 - s.p) * g.Vh;
  return s;
}

// Sum of the four partials per workgroup of the previous kernel B, in one
// fixed order for a 256-lane block: every workgroup that calls it gets the
// same bits.  Split in two so the loads (a fixed count per lane, no loop, so
// no wait between them) are issued before anything else the kernel loads,
// and waited for only when summed.
// PIT: partials per lane and value, ceil(nb / 256) rounded up to 1 or 4 (a
// few loads per lane: the ones past nb are wasted and count against the 63
// loads a wave can have in flight).
constexpr int kPartIters = kDirectSum / kBlk;
template <int PIT> struct Partials4 {
  double v[PIT][4];
};
template <int PIT>
__device__ __forceinline__ void load_partials4(const double* __restrict__ partials, int nb, Partials4<PIT>& pl) {
#pragma unroll
  for (int it = 0; it < PIT; ++it) {
    const int k = threadIdx.x + it * kBlk;
    const int kc = k < nb ? k : nb - 1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double x = partials[(size_t)j * nb + kc];
      pl.v[it][j] = k < nb ? x : 0.0;
    }
  }
}
template <int PIT> __device__ __forceinline__ void add_partials4(const Partials4<PIT>& pl, double (&d)[4]) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    d[j] = 0.0;
#pragma unroll
    for (int it = 0; it < PIT; ++it) d[j] += pl.v[it][j];
  }
}

// The solve state read by one load per lane (a dword each) and written to
// LDS: one memory round trip, where a struct copy by one lane issues its
// loads one after another.
constexpr int kStateWords = (int)(sizeof(CGState) / 4);
static_assert(sizeof(CGState) % 4 == 0 && kStateWords <= kBlk, "one dword per lane");
__device__ __forceinline__ unsigned state_word(const CGState* st) {
  const int t = (int)threadIdx.x < kStateWords ? (int)threadIdx.x : kStateWords - 1;
  return reinterpret_cast<const unsigned*>(st)[t];
}
__device__ __forceinline__ void state_to_lds(CGState& sst, unsigned w) {
  if ((int)threadIdx.x < kStateWords) reinterpret_cast<unsigned*>(&sst)[threadIdx.x] = w;
}

// PIT > 0 (small grids, DEFER): the previous B left its dots as plain per-workgroup
// partials; every workgroup of A sums them and runs the finisher itself (no
// grid-wide atomics on the iteration's critical path).  A reads the state
// from st_in and writes the advanced state to st_out, a different copy, so
// no workgroup reads what another writes; successive iterations alternate
// the two copies.  Otherwise st_in == st_out and B has run the finisher.
// cur: the buffer this launch writes r_k and p_k into (the launch's position
// in the chunk, so known before the state is read); workgroup 0 records it
// in the state for the reliable update.
// Loads that do not depend on the state (the own site's r, v, p and x) are
// issued first and decoded after the state is known, so they are in flight
// while the state and the previous B's dots are read and finished.  (Issuing
// the eight links there as well was measured slower in the solve at 16^4:
// 2.91 against 2.89 ms.)
template <class T, int PIT, bool R12>
__global__ __launch_bounds__(kBlk) void k_mix_a(Geom g, const typename Prec<T>::L* __restrict__ U, MixBufs<T> b,
                                                float mass, const CGState* st_in, CGState* st_out,
                                                const double* __restrict__ partials, int nb, int cur) {
  constexpr bool DEFER = PIT > 0;
  // small grids: the previous B's partials and the state, loaded first
  Partials4<DEFER ? PIT : 1> pl;
  unsigned sw = 0;
  if constexpr (DEFER) {
    load_partials4(partials, nb, pl);
    sw = state_word(st_in);
  }
  const Site s = site_of(g, nb);
  const bool valid = s.cb < g.Vh;
  const int prv = cur ^ 1;
  const typename Prec<T>::S* __restrict__ rp = b.r[prv];
  const typename Prec<T>::S* __restrict__ pp = b.p[prv];
  const float* __restrict__ rpn = b.rn[prv];
  const float* __restrict__ ppn = b.pn[prv];
  const int V = g.V;
  // lanes past the lattice load site (p, 0) and store nothing
  const int cb = valid ? s.cb : 0;
  const int i = valid ? s.i : s.p * g.Vh;
  int n[4];
  cb_coords(g, cb, s.p, n);
  SpinorWords<T> rw, vw, pw;
  rw.load(rp, V, i);
  vw.load(b.v, V, i);
  pw.load(pp, V, i);
  float2 xx[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) xx[c] = b.x[(size_t)c * V + i];
  float al, be;
  if constexpr (DEFER) {
    __shared__ CGState sst;
    // the partials were read together with the state (one round trip); they
    // are used only if the state says a finisher is pending
    state_to_lds(sst, sw);
    __syncthreads();
    if (sst.fin_pending) {
      double d[4];
      add_partials4(pl, d);
      block_sum_n<4>(d);
      if (threadIdx.x == 0) {
        finish_fused(&sst, d);
        sst.fin_pending = 0;
      }
      __syncthreads();
    }
    const bool stop = sst.done;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      CGState o = sst;
      o.fin_pending = stop ? 0 : 1;  // the B that follows leaves partials
      if (!stop) o.pbuf = cur;
      *st_out = o;
    }
    if (stop) return;  // uniform over the block
    al = (float)sst.alpha;
    be = (float)sst.beta;
  } else {
    if (st_in->done) return;
    al = (float)st_in->alpha;
    be = (float)st_in->beta;
    if (blockIdx.x == 0 && threadIdx.x == 0) st_out->pbuf = cur;
  }
  if (!valid) return;
  Q16 qr, qp;
  // own site: advance, store r_k and p_k, accumulate x
  vec3<float> ro, vo, po;
  rw.decode(ro);
  vw.decode(vo);
  pw.decode(po);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    xx[c].x = __fmaf_rn(al, po.c[c].re, xx[c].x);
    xx[c].y = __fmaf_rn(al, po.c[c].im, xx[c].y);
    b.x[(size_t)c * V + i] = xx[c];
  }
  advance<T>(ro, vo, po, al, be, qr, qp);
  store_q<T>(b.r[cur], b.rn[cur], V, i, ro, qr);
  store_q<T>(b.p[cur], b.pn[cur], V, i, po, qp);
  // t = m p_k - (1/2) sum_mu eta [bup U p_k(n+mu) - bdn U^dag p_k(n-mu)]
  vec3<float> acc;
#pragma unroll
  for (int c = 0; c < 3; ++c) acc.c[c] = {0.f, 0.f};
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    int fw, bw;
    float sf, sb;
    neighbours<float>(g, s.cb, n, mu, fw, bw, sf, sb);
    mat3<float> Uf, Ub;
    vec3<float> rf, vf, pf, rb, vb, pb;
    load_link<T, R12>(U, V, mu, s.i, Uf);
    load_spinor<T>(rp, rpn, V, s.q + fw, rf);
    load_spinor<T>(b.v, b.vn, V, s.q + fw, vf);
    load_spinor<T>(pp, ppn, V, s.q + fw, pf);
    load_link<T, R12>(U, V, mu, s.q + bw, Ub);
    load_spinor<T>(rp, rpn, V, s.q + bw, rb);
    load_spinor<T>(b.v, b.vn, V, s.q + bw, vb);
    load_spinor<T>(pp, ppn, V, s.q + bw, pb);
    Q16 t0, t1;
    advance<T>(rf, vf, pf, al, be, t0, t1);
    advance<T>(rb, vb, pb, al, be, t0, t1);
    acc_mv(acc, Uf, pf, sf);
    acc_mdv(acc, Ub, pb, -sb);
  }
  vec3<float> res;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    res.c[c].re = mass * po.c[c].re - 0.5f * acc.c[c].re;
    res.c[c].im = mass * po.c[c].im - 0.5f * acc.c[c].im;
  }
  store_spinor<T>(b.t, b.tn, V, i, res);
}

// DM (how the four dots leave the kernel): kDotsFinish, a grid reduction
// whose finisher advances the state; kDotsPartials (small grids), plain
// per-workgroup partials that every workgroup of the next kernel A sums;
// kDotsGroups (split form), the group level of the grid reduction only, whose
// group sums every workgroup of the next update kernel adds up.
template <class T, int DM, bool R12>
__global__ __launch_bounds__(kBlk) void k_mix_b(Geom g, const typename Prec<T>::L* __restrict__ U, MixBufs<T> b,
                                                float mass, CGState* st, double* partials, int nb, int cur) {
  if (st->done) return;
  const Site s = site_of(g, nb);
  double d[4] = {0.0, 0.0, 0.0, 0.0};
  if (s.cb < g.Vh) {
    const int V = g.V;
    int n[4];
    cb_coords(g, s.cb, s.p, n);
    vec3<float> acc;
#pragma unroll
    for (int c = 0; c < 3; ++c) acc.c[c] = {0.f, 0.f};
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      int fw, bw;
      float sf, sb;
      neighbours<float>(g, s.cb, n, mu, fw, bw, sf, sb);
      if constexpr (kHalfR12<T, R12>) {
        hop_dir_h2(acc, U, b.t, V, mu, s.i, s.q + fw, s.q + bw, sf, sb);
      } else {
        mat3<float> Uf, Ub;
        vec3<float> tf, tb;
        load_link<T, R12>(U, V, mu, s.i, Uf);
        load_spinor<T>(b.t, b.tn, V, s.q + fw, tf);
        load_link<T, R12>(U, V, mu, s.q + bw, Ub);
        load_spinor<T>(b.t, b.tn, V, s.q + bw, tb);
        acc_mv(acc, Uf, tf, sf);
        acc_mdv(acc, Ub, tb, -sb);
      }
    }
    vec3<float> v, pk, rk;
    if constexpr (kHalfR12<T, R12>) {
      // m n h + acc / 2 from the record's halves
      const H16 sr = b.t[s.i];
      const float mn = mass * sr.n;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const f16x2 h = __builtin_bit_cast(f16x2, sr.c[c]);
        v.c[c].re = __builtin_fmaf(mn, (float)h.x, 0.5f * acc.c[c].re);
        v.c[c].im = __builtin_fmaf(mn, (float)h.y, 0.5f * acc.c[c].im);
      }
    } else {
      vec3<float> self;
      load_spinor<T>(b.t, b.tn, V, s.i, self);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        v.c[c].re = mass * self.c[c].re + 0.5f * acc.c[c].re;
        v.c[c].im = mass * self.c[c].im + 0.5f * acc.c[c].im;
      }
    }
    Q16 qv;
    if constexpr (is_half<T>()) quant16(v, qv);
    store_q<T>(b.v, b.vn, V, s.i, v, qv);
    load_spinor<T>(b.p[cur], b.pn[cur], V, s.i, pk);
    load_spinor<T>(b.r[cur], b.rn[cur], V, s.i, rk);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double vr = v.c[c].re, vi = v.c[c].im, pr = pk.c[c].re, pi = pk.c[c].im, rr = rk.c[c].re,
                   ri = rk.c[c].im;
      d[0] += pr * vr + pi * vi;
      d[1] += rr * vr + ri * vi;
      d[2] += vr * vr + vi * vi;
      d[3] += rr * rr + ri * ri;
    }
  }
  if constexpr (DM == kDotsPartials) {
    // per-workgroup partials only: the next kernel A sums them
    block_sum_n<4>(d);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int j = 0; j < 4; ++j) partials[(size_t)j * nb + s.lb] = d[j];
    }
  } else if constexpr (DM == kDotsGroups) {
    group_reduce_n<4>(d, partials, ctr_of(st->ctr, CNT_PV), nb, s.lb);
  } else {
    grid_reduce_n<4>(d, partials, ctr_of(st->ctr, CNT_PV), nb, s.lb, FinFused{st});
  }
}

// Split form, own-site part of kernel A: r_k, p_k and x updated in place
// (r and p single-buffered: both slots point to one array).  The hop
// t = D^dag p_k then runs as a separate launch.  The previous B left the
// group sums of its dots (kDotsGroups): every workgroup adds them up in one
// fixed order and runs the finisher itself, reading the state from st_in and
// (workgroup 0) writing the advanced state to st_out, a different copy; the
// own site's r, v, p are loaded before the state is waited for.  nbb: the
// workgroups of kernel B.
template <class T>
__global__ __launch_bounds__(kBlk) void k_mix_upd(int V, MixBufs<T> b, const CGState* st_in, CGState* st_out,
                                                  const double* __restrict__ partials, int nbb) {
  static_assert(kRedGroups <= 64, "the group sums fit one wave");
  double d[4];
  load_groups_n<4>(partials, nbb, d);
  const unsigned sw = state_word(st_in);
  const int i0 = blockIdx.x * kBlk + threadIdx.x;
  const bool valid = i0 < V;
  const int i = valid ? i0 : 0;
  SpinorWords<T> rw, vw, pw;
  rw.load(b.r[0], V, i);
  vw.load(b.v, V, i);
  pw.load(b.p[0], V, i);
  float2 xx[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) xx[c] = b.x[(size_t)c * V + i];
  __shared__ CGState sst;
  state_to_lds(sst, sw);
  __syncthreads();
  if (sst.fin_pending) {  // uniform over the block
    // the group sums sit in wave 0; its sum is block_sum_n's (the other
    // waves would add zeros)
    if (threadIdx.x < 64) {
      wave_sum4_all(d);
      if (threadIdx.x == 0) {
        finish_fused(&sst, d);
        sst.fin_pending = 0;
      }
    }
    __syncthreads();
  }
  const bool stop = sst.done;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CGState o = sst;
    o.fin_pending = stop ? 0 : 1;  // the B that follows leaves group sums
    *st_out = o;
  }
  if (stop || !valid) return;
  const float al = (float)sst.alpha, be = (float)sst.beta;
  vec3<float> ro, vo, po;
  rw.decode(ro);
  vw.decode(vo);
  pw.decode(po);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    xx[c].x = __fmaf_rn(al, po.c[c].re, xx[c].x);
    xx[c].y = __fmaf_rn(al, po.c[c].im, xx[c].y);
    b.x[(size_t)c * V + i] = xx[c];
  }
  Q16 qr, qp;
  advance<T>(ro, vo, po, al, be, qr, qp);
  store_q<T>(b.r[0], b.rn[0], V, i, ro, qr);
  store_q<T>(b.p[0], b.pn[0], V, i, po, qp);
}

// First step of the reliable update at a chunk's end: x64 += x + alpha p_k
// and x = 0 when an update is pending.  With deferred dots (DM not
// kDotsFinish) the chunk's last B left its dots for a finisher: every
// workgroup sums them in one fixed order (as kernel A or the update kernel
// would) and runs it, reading the state from st_in and (workgroup 0)
// writing the advanced state to st_out, the copy the rest of the update
// works on; otherwise st_in == st_out.  nbb: the workgroups of kernel B.
template <class T, int DM>
__global__ __launch_bounds__(kBlk) void k_mix_acc(int V, double2* __restrict__ xh, MixBufs<T> b, const CGState* st_in,
                                                  CGState* st_out, const double* __restrict__ partials, int nbb) {
  __shared__ CGState sst;
  if constexpr (DM == kDotsFinish) {
    if (threadIdx.x == 0) sst = *st_in;
    __syncthreads();
  } else {
    static_assert(kRedGroups <= 64, "the group sums fit one wave");
    Partials4<kPartIters> pl;
    double d[4];
    if constexpr (DM == kDotsGroups) load_groups_n<4>(partials, nbb, d);
    else load_partials4(partials, nbb, pl);
    state_to_lds(sst, state_word(st_in));
    __syncthreads();
    if (sst.fin_pending) {  // uniform over the block
      if constexpr (DM == kDotsGroups) {
        if (threadIdx.x < 64) {  // the group sums sit in wave 0
          wave_sum4_all(d);
          if (threadIdx.x == 0) {
            finish_fused(&sst, d);
            sst.fin_pending = 0;
          }
        }
      } else {
        add_partials4(pl, d);
        block_sum_n<4>(d);
        if (threadIdx.x == 0) {
          finish_fused(&sst, d);
          sst.fin_pending = 0;
        }
      }
      __syncthreads();
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // ru_apply is read only by the apply kernel of this chunk
    if constexpr (DM == kDotsFinish) {
      st_out->ru_apply = 0;  // in place: the only field that changes
    } else {
      CGState o = sst;
      o.ru_apply = 0;
      *st_out = o;
    }
  }
  if (!sst.ru_pending) return;  // uniform
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= V) return;
  const int k = sst.pbuf;  // buffers of p_k, the direction of the last iteration
  const double al = sst.alpha;
  vec3<float> pk;
  load_spinor<T>(b.p[k], b.pn[k], V, i, pk);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t j = (size_t)c * V + i;
    const float2 xl = b.x[j];
    double2 h = xh[j];
    h.x += (double)xl.x + al * (double)pk.c[c].re;
    h.y += (double)xl.y + al * (double)pk.c[c].im;
    xh[j] = h;
    b.x[j] = make_float2(0.f, 0.f);
  }
}

// Last step of the reliable update: r_k = r64 in sloppy storage when the
// solve restarts; workgroup 0 first copies the update's state copy st to
// st_out (the copy the next chunk starts from) when they differ.
template <class T>
__global__ __launch_bounds__(kBlk) void k_mix_apply(int V, const double2* __restrict__ r64, MixBufs<T> b,
                                                    const CGState* st, CGState* st_out) {
  if (st_out != st && blockIdx.x == 0 && threadIdx.x == 0) *st_out = *st;
  if (!st->ru_apply) return;
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= V) return;
  // the next chunk's first kernel A reads r_{k-1} and p_{k-1} from buffer 1;
  // p_k is in buffer pbuf (0 only when the solve paused at an odd position)
  vec3<float> r;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double2 t = r64[(size_t)c * V + i];
    r.c[c] = {(float)t.x, (float)t.y};
  }
  store_spinor<T>(b.r[1], b.rn[1], V, i, r);
  if (st->pbuf == 0 && b.p[0] != b.p[1]) {
#pragma unroll
    for (int c = 0; c < kSpinorElems<T>; ++c) b.p[1][(size_t)c * V + i] = b.p[0][(size_t)c * V + i];
  }
}

inline int mix_blocks(const Geom& g) { return (g.Vh + kBlk / 2 - 1) / (kBlk / 2); }

// Second FP64 hop of the reliable update with its residual: r = phi - D t
// (D = m + the hop, t = D^dag x) at every site, and |r|^2 reduced with
// FinRU (converged, out of budget, or restart), so D t never goes to
// memory.  Same site map and per-site arithmetic as the FP64 hop; gated on
// a pending update like the other steps of the chunk's tail.
template <bool R12>
__global__ __launch_bounds__(kBlk) void k_ru_resid(Geom g, const double2* __restrict__ U,
                                                   const double2* __restrict__ t, const double2* __restrict__ phi,
                                                   double2* __restrict__ r, double mass, CGState* st,
                                                   double* partials, int nb) {
  if (!st->ru_pending) return;
  const Site s = site_of(g, nb);
  double rr = 0.0;
  if (s.cb < g.Vh) {
    const int V = g.V;
    int n[4];
    cb_coords(g, s.cb, s.p, n);
    vec3<double> acc;
#pragma unroll
    for (int c = 0; c < 3; ++c) acc.c[c] = {0.0, 0.0};
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      int fw, bw;
      double sf, sb;
      neighbours<double>(g, s.cb, n, mu, fw, bw, sf, sb);
      mat3<double> Uf, Ub;
      vec3<double> vf, vb;
      load_link<double, R12>(U, V, mu, s.i, Uf);
      load_spinor<double>(t, nullptr, V, s.q + fw, vf);
      load_link<double, R12>(U, V, mu, s.q + bw, Ub);
      load_spinor<double>(t, nullptr, V, s.q + bw, vb);
      acc_mv(acc, Uf, vf, sf);
      acc_mdv(acc, Ub, vb, -sb);
    }
    vec3<double> self;
    load_spinor<double>(t, nullptr, V, s.i, self);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t j = (size_t)c * V + s.i;
      const double2 f = phi[j];
      const double2 d = make_double2(f.x - (mass * self.c[c].re + 0.5 * acc.c[c].re),
                                     f.y - (mass * self.c[c].im + 0.5 * acc.c[c].im));
      r[j] = d;
      rr += d.x * d.x + d.y * d.y;
    }
  }
  grid_reduce(rr, partials, ctr_of(st->ctr, CNT_RU), nb, s.lb, FinRU{st});
}

}  // namespace

bool mix_defers(const Geom& g) { return mix_blocks(g) <= kDirectSum; }

template <class T>
void mix_a(hipStream_t s, const Geom& g, const typename Prec<T>::L* U, const MixBufs<T>& b, float mass,
           const CGState* st_in, CGState* st_out, const double* partials, bool r12, int cur) {
  const int nb = mix_blocks(g);
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(nb), dim3(kBlk), 0, s, g, U, b, mass, st_in, st_out, partials, nb, cur);
  };
  if (st_in != st_out && nb <= kBlk) {
    if (r12) go(k_mix_a<T, 1, true>);
    else go(k_mix_a<T, 1, false>);
  } else if (st_in != st_out) {
    if (r12) go(k_mix_a<T, kPartIters, true>);
    else go(k_mix_a<T, kPartIters, false>);
  } else {
    if (r12) go(k_mix_a<T, 0, true>);
    else go(k_mix_a<T, 0, false>);
  }
}
template <class T>
void mix_b(hipStream_t s, const Geom& g, const typename Prec<T>::L* U, const MixBufs<T>& b, float mass, CGState* st,
           double* partials, int dm, bool r12, int cur) {
  const int nb = mix_blocks(g);
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(nb), dim3(kBlk), 0, s, g, U, b, mass, st, partials, nb, cur);
  };
  if (dm == kDotsPartials) {
    if (r12) go(k_mix_b<T, kDotsPartials, true>);
    else go(k_mix_b<T, kDotsPartials, false>);
  } else if (dm == kDotsGroups) {
    if (r12) go(k_mix_b<T, kDotsGroups, true>);
    else go(k_mix_b<T, kDotsGroups, false>);
  } else {
    if (r12) go(k_mix_b<T, kDotsFinish, true>);
    else go(k_mix_b<T, kDotsFinish, false>);
  }
}
void ru_resid(hipStream_t s, const Geom& g, const double2* U, const double2* t, const double2* phi, double2* r,
              double mass, CGState* st, double* partials, bool r12) {
  const int nb = mix_blocks(g);
  if (r12)
    hipLaunchKernelGGL(k_ru_resid<true>, dim3(nb), dim3(kBlk), 0, s, g, U, t, phi, r, mass, st, partials, nb);
  else
    hipLaunchKernelGGL(k_ru_resid<false>, dim3(nb), dim3(kBlk), 0, s, g, U, t, phi, r, mass, st, partials, nb);
}
template <class T>
void mix_upd(hipStream_t s, const Geom& g, const MixBufs<T>& b, const CGState* st_in, CGState* st_out,
             const double* partials) {
  hipLaunchKernelGGL(k_mix_upd<T>, dim3((g.V + kBlk - 1) / kBlk), dim3(kBlk), 0, s, g.V, b, st_in, st_out, partials,
                     mix_blocks(g));
}
template <class T>
void mix_acc(hipStream_t s, const Geom& g, double2* x64, const MixBufs<T>& b, const CGState* st_in, CGState* st_out,
             const double* partials, int dm) {
  const dim3 grid((g.V + kBlk - 1) / kBlk);
  const int nbb = mix_blocks(g);
  if (dm == kDotsGroups)
    hipLaunchKernelGGL((k_mix_acc<T, kDotsGroups>), grid, dim3(kBlk), 0, s, g.V, x64, b, st_in, st_out, partials, nbb);
  else if (dm == kDotsPartials)
    hipLaunchKernelGGL((k_mix_acc<T, kDotsPartials>), grid, dim3(kBlk), 0, s, g.V, x64, b, st_in, st_out, partials, nbb);
  else
    hipLaunchKernelGGL((k_mix_acc<T, kDotsFinish>), grid, dim3(kBlk), 0, s, g.V, x64, b, st_in, st_out, partials, nbb);
}
template <class T>
void mix_apply(hipStream_t s, const Geom& g, const double2* r64, const MixBufs<T>& b, const CGState* st,
               CGState* st_out) {
  hipLaunchKernelGGL(k_mix_apply<T>, dim3((g.V + kBlk - 1) / kBlk), dim3(kBlk), 0, s, g.V, r64, b, st, st_out);
}

template void mix_a<float>(hipStream_t, const Geom&, const float2*, const MixBufs<float>&, float, const CGState*,
                           CGState*, const double*, bool, int);
template void mix_a<Half>(hipStream_t, const Geom&, const __half2*, const MixBufs<Half>&, float, const CGState*,
                          CGState*, const double*, bool, int);
template void mix_b<float>(hipStream_t, const Geom&, const float2*, const MixBufs<float>&, float, CGState*, double*,
                           int, bool, int);
template void mix_b<Half>(hipStream_t, const Geom&, const __half2*, const MixBufs<Half>&, float, CGState*, double*,
                          int, bool, int);
template void mix_upd<float>(hipStream_t, const Geom&, const MixBufs<float>&, const CGState*, CGState*,
                             const double*);
template void mix_upd<Half>(hipStream_t, const Geom&, const MixBufs<Half>&, const CGState*, CGState*, const double*);
template void mix_acc<float>(hipStream_t, const Geom&, double2*, const MixBufs<float>&, const CGState*, CGState*,
                             const double*, int);
template void mix_acc<Half>(hipStream_t, const Geom&, double2*, const MixBufs<Half>&, const CGState*, CGState*,
                            const double*, int);
template void mix_apply<float>(hipStream_t, const Geom&, const double2*, const MixBufs<float>&, const CGState*,
                               CGState*);
template void mix_apply<Half>(hipStream_t, const Geom&, const double2*, const MixBufs<Half>&, const CGState*,
                              CGState*);

}  // namespace sfh
