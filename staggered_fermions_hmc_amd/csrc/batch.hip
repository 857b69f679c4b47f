// Batched CG engine: many independent solves (different lattices, masses,
// and up to 8 right-hand sides per lattice sharing its links) advanced
// together, one launch per CG phase for every active solve.
//
// * The hop kernel walks a block map of (system, block, right-hand-side
//   tile) entries, one lane per site; a lane loads each of its links once
//   and applies it to every active right-hand side of its tile, whose
//   accumulators it keeps in registers (multi-RHS: the links cross HBM once
//   per tile).
// * The vector kernels walk a map of (slot, block) pairs, a slot being one
//   right-hand side in flight.
// * Every slot has its own device CGState and reduction partials, and a
//   system's blocks and reduction order depend only on its own lattice, so a
//   solve gives the same result alone or in any mix.
// * A chunk of iterations is captured into a HIP graph and replayed until the
//   set of active slots changes.
// * Mixed precision (FP32 inner): a slot whose iterated residual has dropped
//   by the reliable-update factor has an update due and iterates on to the
//   chunk's end (one that meets the target or its budget pauses); the
//   chunk's tail runs the FP64 reliable update of every such slot on the
//   device (x accumulated in FP64, true residual with the FP64 links,
//   restart or stop) while the others keep their place.  All states then
//   reach the host in one copy.
// * Slots can be admitted between chunks (continuous batching): the solve
//   service feeds the engine as requests arrive; each system owns a fixed
//   pool of slot memory, so admission is by free memory.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "batch.h"
#include "hopcore.h"
#include "reduce.h"

namespace sfh {

namespace {

constexpr int kBlas = 256;
template <class T> constexpr bool kMixedT = std::is_same<T, float>::value;

// Block-map entry of this workgroup: workgroups reach the 8 XCDs round-robin,
// so runs of kRun consecutive entries go to one XCD (neighbouring ranges of a
// system share its L2) while the runs themselves alternate between XCDs, so
// entries of very different cost (1 to 8 right-hand sides) stay spread over
// all of them.  A bijection on [0, nb): the last nb mod (8 kRun) entries keep
// their own index.
constexpr int kRun = 16;
__device__ __forceinline__ int map_entry(int nb) {
  const int k = blockIdx.x;
  if (k >= nb - nb % (8 * kRun)) return k;
  const int x = k & 7, j = k >> 3;
  return ((j / kRun) * 8 + x) * kRun + j % kRun;
}

// One lane per site over (system, block, right-hand-side tile) entries; each
// block covers one checkerboard range on both parities (as the single-system
// hop).  A lane loads the two links of a direction once and applies them to
// every active right-hand side of its tile (KT at most), whose accumulators
// it holds in registers: the links cross HBM once per tile, not once per
// right-hand side.  Each right-hand side's sum runs over the directions in
// one fixed order (forward, then backward term), so the tile changes no
// result.
// phase DOT=false: t = D^dag p ; DOT=true: v = D t and <p, v>.
template <class T, bool DOT, bool R12, int KT>
__global__ __launch_bounds__(512) void kb_hop(const BSys<T>* __restrict__ sys, const BSlot<T>* __restrict__ slots,
                                              const int4* __restrict__ bmap, typename Prec<T>::R hb) {
  using R = typename Prec<T>::R;
  using SP = typename Prec<T>::S;
  // system, block, right-hand sides [z, w) of the system
  const int4 m = bmap[map_entry(gridDim.x)];
  const BSys<T>& S = sys[m.x];  // read in place: a private copy with a dynamically indexed slot[] spills
  const int lb = m.y;
  const Geom& g = S.g;
  const int hw = blockDim.x >> 1;
  const int p = threadIdx.x >= hw;
  const int cb = range_start(g, lb, hw) + (threadIdx.x - p * hw);
  const bool valid = cb < g.Vh;
  const int V = g.V;
  const int i = p * g.Vh + (valid ? cb : 0);
  const int q = (1 - p) * g.Vh;
  const int nj = m.w - m.z;
  __shared__ int sdone[KT];
  if (threadIdx.x < KT) sdone[threadIdx.x] = threadIdx.x < nj ? slots[S.slot[m.z + threadIdx.x]].st->done : 1;
  __syncthreads();
  bool act[KT];
  const SP* in[KT];
#pragma unroll
  for (int j = 0; j < KT; ++j) {
    act[j] = !sdone[j];
    const BSlot<T>& L = slots[S.slot[m.z + (act[j] ? j : 0)]];
    in[j] = DOT ? L.t : L.p;
  }
  vec3<R> acc[KT];
#pragma unroll
  for (int j = 0; j < KT; ++j)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[j].c[c] = {R(0), R(0)};
  if (valid) {
    int n[4];
    cb_coords(g, cb, p, n);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      int fw, bw;
      R sf, sb;
      neighbours(g, cb, n, mu, fw, bw, sf, sb);
      mat3<R> Uf, Ub;
      load_link<T, R12>(S.U, V, mu, i, Uf);
      load_link<T, R12>(S.U, V, mu, q + bw, Ub);
#pragma unroll
      for (int j = 0; j < KT; ++j) {
        if (!act[j]) continue;
        vec3<R> vf, vb;
        load_spinor<T>(in[j], nullptr, V, q + fw, vf);
        load_spinor<T>(in[j], nullptr, V, q + bw, vb);
        acc_mv(acc[j], Uf, vf, sf);
        acc_mdv(acc[j], Ub, vb, -sb);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < KT; ++j) {
    if (!act[j]) continue;
    const BSlot<T>& L = slots[S.slot[m.z + j]];
    SP* out = DOT ? L.v : L.t;
    // mixed precision: <p,v>, <r,v>, <v,v>, <r,r> for FinFused; FP64: <p,v>
    double d[4] = {0.0, 0.0, 0.0, 0.0};
    if (valid) {
      vec3<R> self, res;
      load_spinor<T>(in[j], nullptr, V, i, self);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        res.c[c].re = S.mass * self.c[c].re + hb * acc[j].c[c].re;
        res.c[c].im = S.mass * self.c[c].im + hb * acc[j].c[c].im;
      }
      store_spinor<T>(out, nullptr, V, i, res);
      if constexpr (DOT) {
        vec3<R> w, rk;
        load_spinor<T>(L.p, nullptr, V, i, w);
        if constexpr (kMixedT<T>) load_spinor<T>(L.r, nullptr, V, i, rk);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double vr = res.c[c].re, vi = res.c[c].im;
          d[0] += (double)w.c[c].re * vr + (double)w.c[c].im * vi;
          if constexpr (kMixedT<T>) {
            const double rr = rk.c[c].re, ri = rk.c[c].im;
            d[1] += rr * vr + ri * vi;
            d[2] += vr * vr + vi * vi;
            d[3] += rr * rr + ri * ri;
          }
        }
      }
    }
    if constexpr (DOT) {
      if constexpr (kMixedT<T>) grid_reduce_n<4>(d, L.part, ctr_of(L.st->ctr, CNT_PV), S.nb_hop, lb, FinFused{L.st});
      else grid_reduce(d[0], L.part, ctr_of(L.st->ctr, CNT_PV), S.nb_hop, lb, FinPV{L.st});
    }
  }
}

// Launch of the batched hop with the tile length as a compile-time bound.
template <class T, bool DOT, bool R12>
void launch_kb_hop(int tile, int nblocks, int threads, hipStream_t st, const BSys<T>* sys, const BSlot<T>* slots,
                   const int4* bmap, typename Prec<T>::R hb) {
  switch (tile) {
    case 1: hipLaunchKernelGGL((kb_hop<T, DOT, R12, 1>), dim3(nblocks), dim3(threads), 0, st, sys, slots, bmap, hb); break;
    case 2: hipLaunchKernelGGL((kb_hop<T, DOT, R12, 2>), dim3(nblocks), dim3(threads), 0, st, sys, slots, bmap, hb); break;
    case 4: hipLaunchKernelGGL((kb_hop<T, DOT, R12, 4>), dim3(nblocks), dim3(threads), 0, st, sys, slots, bmap, hb); break;
    default: hipLaunchKernelGGL((kb_hop<T, DOT, R12, 8>), dim3(nblocks), dim3(threads), 0, st, sys, slots, bmap, hb); break;
  }
}

// Mixed precision, first kernel of an iteration (all updates of the previous
// one, in place): x += alpha p, r -= alpha v, p = r + beta p.  With alpha =
// beta = 0 (a fresh solve, p = v = 0) it leaves p = r; after a reliable
// update (alpha = 0, r the true residual) it forms p = r + beta p.
template <class T>
__global__ __launch_bounds__(kBlas) void kb_upd(const BSlot<T>* __restrict__ slots, const int2* __restrict__ bmap) {
  const int2 m = bmap[blockIdx.x];
  const BSlot<T> L = slots[m.x];
  if (L.st->done) return;
  const int i = m.y * kBlas + threadIdx.x;
  if (i >= L.V) return;
  const float al = (float)L.st->alpha, be = (float)L.st->beta;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t k = (size_t)c * L.V + i;
    float2 xx = L.x[k], rv = L.r[k], pp = L.p[k];
    const float2 vv = L.v[k];
    xx.x = __fmaf_rn(al, pp.x, xx.x);
    xx.y = __fmaf_rn(al, pp.y, xx.y);
    rv.x = __fmaf_rn(-al, vv.x, rv.x);
    rv.y = __fmaf_rn(-al, vv.y, rv.y);
    pp.x = __fmaf_rn(be, pp.x, rv.x);
    pp.y = __fmaf_rn(be, pp.y, rv.y);
    L.x[k] = xx;
    L.r[k] = rv;
    L.p[k] = pp;
  }
}

template <class T>
__global__ __launch_bounds__(kBlas) void kb_xr(const BSlot<T>* __restrict__ slots, const int2* __restrict__ bmap) {
  const int2 m = bmap[blockIdx.x];
  const BSlot<T> L = slots[m.x];
  CGState* st = L.st;
  if (st->done) return;
  const int V = L.V;
  const int nb = (V + kBlas - 1) / kBlas;
  const int i = m.y * kBlas + threadIdx.x;
  const double alpha = st->alpha;
  double rr = 0.0;
  if (i < V) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t k = (size_t)c * V + i;
      auto xx = L.x[k], rv = L.r[k];
      const auto pp = L.p[k], vv = L.v[k];
      using R = decltype(xx.x);
      const R al = (R)alpha;
      xx.x += al * pp.x;
      xx.y += al * pp.y;
      rv.x -= al * vv.x;
      rv.y -= al * vv.y;
      L.x[k] = xx;
      L.r[k] = rv;
      rr += (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
    }
  }
  grid_reduce(rr, L.part, ctr_of(st->ctr, CNT_RR), nb, m.y, FinRR{st});
}

template <class T>
__global__ __launch_bounds__(kBlas) void kb_p(const BSlot<T>* __restrict__ slots, const int2* __restrict__ bmap) {
  const int2 m = bmap[blockIdx.x];
  const BSlot<T> L = slots[m.x];
  if (L.st->done) return;
  const int V = L.V;
  const int i = m.y * kBlas + threadIdx.x;
  if (i >= V) return;
  const double beta = L.st->beta;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t k = (size_t)c * V + i;
    auto pp = L.p[k];
    const auto rv = L.r[k];
    using R = decltype(pp.x);
    const R be = (R)beta;
    pp.x = rv.x + be * pp.x;
    pp.y = rv.y + be * pp.y;
    L.p[k] = pp;
  }
}

// Reliable update of every mixed solve that has one due, on the device (gated
// on each slot's state, so the other solves cost one state read):
//   kb_ru_acc   x64 += x + alpha p, x = 0
//   kb_ru_hop   t64 = D^dag x64, then ax = D t64 (FP64 links, loaded once per
//               block for all such right-hand sides of the system)
//   kb_ru_sub   r64 = b - ax, |r64|^2 -> FinRU (restart, converged or out of budget)
//   kb_ru_apply r = r64 in FP32 (the next kb_upd restarts p from it)
template <class T>
__global__ __launch_bounds__(kBlas) void kb_ru_acc(const BSlot<T>* __restrict__ slots, const int2* __restrict__ bmap) {
  const int2 m = bmap[blockIdx.x];
  const BSlot<T> L = slots[m.x];
  if (m.y == 0 && threadIdx.x == 0) L.st->ru_apply = 0;  // read only by kb_ru_apply of this chunk
  if (!L.st->ru_pending) return;
  const int i = m.y * kBlas + threadIdx.x;
  if (i >= L.V) return;
  const double al = L.st->alpha;  // the last iteration's step, not yet in x
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t k = (size_t)c * L.V + i;
    const auto xl = L.x[k];
    const auto pk = L.p[k];
    double2 h = L.x64[k];
    h.x += (double)xl.x + al * (double)pk.x;
    h.y += (double)xl.y + al * (double)pk.y;
    L.x64[k] = h;
    L.x[k] = {0, 0};
  }
}

// FP64 hops of the reliable update (t64 = D^dag x64, then ax = D t64) for
// every slot of the tile that has one pending, with the same register tile
// as kb_hop over the FP64 links.
template <class T, int PHASE, bool R12, int KT>
__global__ __launch_bounds__(256) void kb_ru_hop(const BSys<T>* __restrict__ sys, const BSlot<T>* __restrict__ slots,
                                                 const int4* __restrict__ bmap) {
  const int4 m = bmap[map_entry(gridDim.x)];
  const BSys<T>& S = sys[m.x];  // read in place: a private copy with a dynamically indexed slot[] spills
  const int nj = m.w - m.z;
  __shared__ int spend[KT];
  if (threadIdx.x < KT) spend[threadIdx.x] = threadIdx.x < nj ? slots[S.slot[m.z + threadIdx.x]].st->ru_pending : 0;
  __syncthreads();
  bool act[KT];
  int any = 0;
#pragma unroll
  for (int j = 0; j < KT; ++j) {
    act[j] = spend[j] != 0;
    any |= spend[j];
  }
  if (!any) return;  // uniform over the block
  const int lb = m.y;
  const Geom& g = S.g;
  const int hw = blockDim.x >> 1;
  const int p = threadIdx.x >= hw;
  const int cb = range_start(g, lb, hw) + (threadIdx.x - p * hw);
  if (cb >= g.Vh) return;  // no barrier below
  const int V = g.V;
  const int i = p * g.Vh + cb;
  const int q = (1 - p) * g.Vh;
  const double2* in[KT];
#pragma unroll
  for (int j = 0; j < KT; ++j) {
    const BSlot<T>& L = slots[S.slot[m.z + (act[j] ? j : 0)]];
    in[j] = PHASE == 0 ? L.x64 : L.t64;
  }
  vec3<double> acc[KT];
#pragma unroll
  for (int j = 0; j < KT; ++j)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[j].c[c] = {0.0, 0.0};
  int n[4];
  cb_coords(g, cb, p, n);
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    int fw, bw;
    double sf, sb;
    neighbours(g, cb, n, mu, fw, bw, sf, sb);
    mat3<double> Uf, Ub;
    load_link<double, R12>(S.U64, V, mu, i, Uf);
    load_link<double, R12>(S.U64, V, mu, q + bw, Ub);
#pragma unroll
    for (int j = 0; j < KT; ++j) {
      if (!act[j]) continue;
      vec3<double> vf, vb;
      load_spinor<double>(in[j], nullptr, V, q + fw, vf);
      load_spinor<double>(in[j], nullptr, V, q + bw, vb);
      acc_mv(acc[j], Uf, vf, sf);
      acc_mdv(acc[j], Ub, vb, -sb);
    }
  }
  const double hb = PHASE == 0 ? -0.5 : 0.5;
#pragma unroll
  for (int j = 0; j < KT; ++j) {
    if (!act[j]) continue;
    const BSlot<T>& L = slots[S.slot[m.z + j]];
    double2* out = PHASE == 0 ? L.t64 : L.ax;
    vec3<double> self;
    load_spinor<double>(in[j], nullptr, V, i, self);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      out[(size_t)c * V + i] = make_double2(S.mass64 * self.c[c].re + hb * acc[j].c[c].re,
                                            S.mass64 * self.c[c].im + hb * acc[j].c[c].im);
  }
}

template <class T, int PHASE, bool R12>
void launch_kb_ru_hop(int tile, int nblocks, hipStream_t st, const BSys<T>* sys, const BSlot<T>* slots,
                      const int4* bmap) {
  switch (tile) {
    case 1: hipLaunchKernelGGL((kb_ru_hop<T, PHASE, R12, 1>), dim3(nblocks), dim3(256), 0, st, sys, slots, bmap); break;
    case 2: hipLaunchKernelGGL((kb_ru_hop<T, PHASE, R12, 2>), dim3(nblocks), dim3(256), 0, st, sys, slots, bmap); break;
    case 4: hipLaunchKernelGGL((kb_ru_hop<T, PHASE, R12, 4>), dim3(nblocks), dim3(256), 0, st, sys, slots, bmap); break;
    default: hipLaunchKernelGGL((kb_ru_hop<T, PHASE, R12, 8>), dim3(nblocks), dim3(256), 0, st, sys, slots, bmap); break;
  }
}

template <class T>
__global__ __launch_bounds__(kBlas) void kb_ru_sub(const BSlot<T>* __restrict__ slots, const int2* __restrict__ bmap) {
  const int2 m = bmap[blockIdx.x];
  const BSlot<T> L = slots[m.x];
  CGState* st = L.st;
  if (!st->ru_pending) return;
  const int V = L.V;
  const int nb = (V + kBlas - 1) / kBlas;
  const int i = m.y * kBlas + threadIdx.x;
  double rr = 0.0;
  if (i < V) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t k = (size_t)c * V + i;
      const double2 a = L.b[k], b = L.ax[k];
      const double2 d = make_double2(a.x - b.x, a.y - b.y);
      L.r64[k] = d;
      rr += d.x * d.x + d.y * d.y;
    }
  }
  grid_reduce(rr, L.part, ctr_of(st->ctr, CNT_RU), nb, m.y, FinRU{st});
}

template <class T>
__global__ __launch_bounds__(kBlas) void kb_ru_apply(const BSlot<T>* __restrict__ slots, const int2* __restrict__ bmap) {
  const int2 m = bmap[blockIdx.x];
  const BSlot<T> L = slots[m.x];
  if (!L.st->ru_apply) return;
  const int i = m.y * kBlas + threadIdx.x;
  if (i >= L.V) return;
  // alpha = 0 and beta = |r_true|^2 / |r_k|^2 in the state: the next kb_upd
  // forms p = r + beta p from this r
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t k = (size_t)c * L.V + i;
    const double2 t = L.r64[k];
    L.r[k] = make_float2((float)t.x, (float)t.y);
  }
}

// All solve states of the launch tables into one array, for one copy to the host.
template <class T>
__global__ void kb_gather(const BSlot<T>* __restrict__ slots, const int* __restrict__ ids, int n,
                          CGState* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = *slots[ids[k]].st;
}

inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

template <class T>
size_t Engine<T>::slot_bytes(int V) {
  const size_t n = 3 * (size_t)V;
  const int nbh = (V / 2 + 127) / 128, nbb = (V + kBlas - 1) / kBlas;
  return align256(n * sizeof(S)) * (kMixed ? 5 : 4) + (kMixed ? 3 * align256(n * sizeof(double2)) : 0) +
         align256(4 * (std::max(nbh, nbb) + kRedGroups) * sizeof(double)) + align256(sizeof(CGState)) +
         align256(kCtrBytes);
}

template <class T>
Engine<T>::Engine(sfh_ctx* c, double tol, int max_iter, double delta, int chunk, bool use_graph)
    : c_(c), tol_(tol), max_iter_(max_iter), delta_(delta > 0 ? delta : kDeltaF32), chunk_(chunk), use_graph_(use_graph) {}

template <class T>
void Engine<T>::launch_chunk() {
  // mixed: update, hop, hop + four dots (three kernels per iteration);
  // FP64: the reference recurrence, hop, hop + <p,v>, x/r update, p update
  for (int it = 0; it < chunk_; ++it) {
    if constexpr (kMixed) hipLaunchKernelGGL((kb_upd<T>), dim3(nbb_), dim3(kBlas), 0, c_->stream, dslot_, dmb_);
    const int tile = cfg_.rhs_tile;
    if (r12_) {
      launch_kb_hop<T, false, true>(tile, nbh0_, cfg_.hop_threads, c_->stream, dsys_, dslot_, dmh0_, (R)(-0.5));
      launch_kb_hop<T, true, true>(tile, nbh_, 256, c_->stream, dsys_, dslot_, dmh_, (R)(0.5));
    } else {
      launch_kb_hop<T, false, false>(tile, nbh0_, cfg_.hop_threads, c_->stream, dsys_, dslot_, dmh0_, (R)(-0.5));
      launch_kb_hop<T, true, false>(tile, nbh_, 256, c_->stream, dsys_, dslot_, dmh_, (R)(0.5));
    }
    if constexpr (!kMixed) {
      hipLaunchKernelGGL((kb_xr<T>), dim3(nbb_), dim3(kBlas), 0, c_->stream, dslot_, dmb_);
      hipLaunchKernelGGL((kb_p<T>), dim3(nbb_), dim3(kBlas), 0, c_->stream, dslot_, dmb_);
    }
  }
}

// Reliable updates of the mixed solves that have one due, then the state gather and its
// copy to pinned host memory.
template <class T>
void Engine<T>::launch_tail() {
  if constexpr (kMixed) {
    hipLaunchKernelGGL((kb_ru_acc<T>), dim3(nbb_), dim3(kBlas), 0, c_->stream, dslot_, dmb_);
    if (r12_64_) {
      launch_kb_ru_hop<T, 0, true>(cfg_.rhs_tile, nbh_, c_->stream, dsys_, dslot_, dmh_);
      launch_kb_ru_hop<T, 1, true>(cfg_.rhs_tile, nbh_, c_->stream, dsys_, dslot_, dmh_);
    } else {
      launch_kb_ru_hop<T, 0, false>(cfg_.rhs_tile, nbh_, c_->stream, dsys_, dslot_, dmh_);
      launch_kb_ru_hop<T, 1, false>(cfg_.rhs_tile, nbh_, c_->stream, dsys_, dslot_, dmh_);
    }
    hipLaunchKernelGGL((kb_ru_sub<T>), dim3(nbb_), dim3(kBlas), 0, c_->stream, dslot_, dmb_);
    hipLaunchKernelGGL((kb_ru_apply<T>), dim3(nbb_), dim3(kBlas), 0, c_->stream, dslot_, dmb_);
  }
  const int n = (int)mapped_.size();
  hipLaunchKernelGGL((kb_gather<T>), dim3((n + 63) / 64), dim3(64), 0, c_->stream, dslot_, dids_, n, dstates_);
  hipMemcpyAsync(hstates_, dstates_, n * sizeof(CGState), hipMemcpyDeviceToHost, c_->stream);
}

template <class T>
Engine<T>::~Engine() {
  if (gexec_) hipGraphExecDestroy(gexec_);
  if (graph_) hipGraphDestroy(graph_);
  for (void* p : owned_) hipFree(p);
  if (hstates_) hipHostFree(hstates_);
}

template <class T>
int Engine<T>::bind_system(SysH& s, const sfh_gauge* gauge, double mass) {
  sfh_gauge* gk = const_cast<sfh_gauge*>(gauge);
  if (int e = ensure_su3(c_, gk)) return e;
  s.gauge = gauge;
  s.g = gauge->g;
  s.U64 = gauge->U;
  s.mass = mass;
  s.su3_64 = gk->f64_su3;
  if (kMixed) {
    if (int e = ensure_lp(c_, gk, SFH_PREC_F32)) return e;
    s.U = reinterpret_cast<const typename Prec<T>::L*>(gk->lp.u32);
    s.su3_lp = gk->lp_su3;
  } else {
    s.U = reinterpret_cast<const typename Prec<T>::L*>(gk->U);
    s.su3_lp = gk->f64_su3;
  }
  return hip_check(c_, hipStreamSynchronize(c_->stream), "system links");
}

template <class T>
int Engine<T>::add_system(const sfh_gauge* gauge, double mass, int cap, int prealloc) {
  if (cap < 1 || cap > kMaxRhs || prealloc < 0 || prealloc > cap)
    return -set_err(c_, SFH_E_INVALID, "slot pool must be 1..8");
  SysH s;
  s.cap = cap;
  if (int e = bind_system(s, gauge, mass)) return -e;
  sys_.push_back(s);
  const int id = (int)sys_.size() - 1;
  for (int k = 0; k < prealloc; ++k) {
    const int slot = grow(id);
    if (slot < 0) return slot;
  }
  dirty_ = true;
  return id;
}

template <class T>
int Engine<T>::update_system(int sys, const sfh_gauge* gauge, double mass) {
  if (sys < 0 || sys >= (int)sys_.size() || sys_[sys].removed) return set_err(c_, SFH_E_INVALID, "no such system");
  for (int k : sys_[sys].pool)
    if (slots_[k].used) return set_err(c_, SFH_E_INVALID, "system has solves in flight");
  for (int k = 0; k < 4; ++k)
    if (gauge->g.L[k] != sys_[sys].g.L[k]) return set_err(c_, SFH_E_INVALID, "system lattice changed shape");
  // the links may have changed in place: their SU(3) check and low-precision
  // copy are measured again
  sfh_gauge* gk = const_cast<sfh_gauge*>(gauge);
  gk->su3_known = false;
  gk->lp32_valid = false;
  gk->lp16_valid = false;
  if (int e = bind_system(sys_[sys], gauge, mass)) return e;
  dirty_ = true;
  return SFH_OK;
}

template <class T>
int Engine<T>::remove_system(int sys) {
  if (sys < 0 || sys >= (int)sys_.size() || sys_[sys].removed) return set_err(c_, SFH_E_INVALID, "no such system");
  for (int k : sys_[sys].pool)
    if (slots_[k].used) return set_err(c_, SFH_E_INVALID, "system has solves in flight");
  for (int k : sys_[sys].pool) spare_.push_back(k);
  sys_[sys].pool.clear();
  sys_[sys].removed = true;
  dirty_ = true;
  return SFH_OK;
}

// Carves one slot's arrays out of its memory for lattice volume V; the
// arrays are exactly those counted by slot_bytes(V).  The reduction counters
// must start at zero.
template <class T>
int Engine<T>::layout_slot(SlotH& h, int V) {
  char* p = static_cast<char*>(h.mem);
  auto take = [&](size_t bytes) {
    char* q = p;
    p += align256(bytes);
    return q;
  };
  BSlot<T>& L = h.dev;
  L = BSlot<T>{};
  L.V = V;
  const size_t n = 3 * (size_t)V;
  // FP64 solves iterate directly in the caller's x (set at admission); the
  // mixed ones keep an FP32 x and an FP64 residual of their own
  if (kMixed) L.x = reinterpret_cast<S*>(take(n * sizeof(S)));
  L.r = reinterpret_cast<S*>(take(n * sizeof(S)));
  L.p = reinterpret_cast<S*>(take(n * sizeof(S)));
  L.t = reinterpret_cast<S*>(take(n * sizeof(S)));
  L.v = reinterpret_cast<S*>(take(n * sizeof(S)));
  h.r64 = nullptr;
  if (kMixed) {
    L.r64 = reinterpret_cast<double2*>(take(n * sizeof(double2)));
    L.t64 = reinterpret_cast<double2*>(take(n * sizeof(double2)));
    L.ax = reinterpret_cast<double2*>(take(n * sizeof(double2)));
    h.r64 = L.r64;
  }
  L.st = reinterpret_cast<CGState*>(take(sizeof(CGState)));
  h.ctr = reinterpret_cast<unsigned*>(take(kCtrBytes));
  L.part = reinterpret_cast<double*>(p);
  const int nbh = (V / 2 + 127) / 128, nbb = (V + kBlas - 1) / kBlas;
  const size_t used = (size_t)(p - static_cast<char*>(h.mem)) +
                      align256(4 * (std::max(nbh, nbb) + kRedGroups) * sizeof(double));
  if (used > h.capacity) return set_err(c_, SFH_E_INVALID, "slot layout exceeds its allocation");
  if (hipMemsetAsync(h.ctr, 0, kCtrBytes, c_->stream) != hipSuccess ||
      hipStreamSynchronize(c_->stream) != hipSuccess)
    return set_err(c_, SFH_E_HIP, "slot counters clear");
  return SFH_OK;
}

template <class T>
int Engine<T>::grow(int sys) {
  SysH& s = sys_[sys];
  if ((int)s.pool.size() >= s.cap) return -set_err(c_, SFH_E_INVALID, "slot pool is at its cap");
  const int V = s.g.V;
  const size_t sb = slot_bytes(V);
  // a spare slot (freed by a removed system) large enough for V, else new memory
  int id = -1;
  for (size_t j = 0; j < spare_.size(); ++j)
    if (slots_[spare_[j]].capacity >= sb) {
      id = spare_[j];
      spare_.erase(spare_.begin() + (long)j);
      break;
    }
  if (id < 0) {
    void* mem = nullptr;
    if (hipMalloc(&mem, sb) != hipSuccess) return -set_err(c_, SFH_E_NOMEM, "slot allocation");
    owned_.push_back(mem);
    bytes_ += sb;
    SlotH h;
    h.mem = mem;
    h.capacity = sb;
    id = (int)slots_.size();
    slots_.push_back(h);
  }
  SlotH& h = slots_[id];
  h.sys = sys;
  h.used = false;
  h.finished = false;
  if (int e = layout_slot(h, V)) return -e;
  s.pool.push_back(id);
  dirty_ = true;
  return id;
}

template <class T>
int Engine<T>::slots_of(int sys) const {
  return (int)sys_[sys].pool.size();
}

template <class T>
int Engine<T>::cap_of(int sys) const {
  return sys_[sys].cap;
}

template <class T>
bool Engine<T>::has_spare(int V) const {
  const size_t sb = slot_bytes(V);
  for (int k : spare_)
    if (slots_[k].capacity >= sb) return true;
  return false;
}

template <class T>
int Engine<T>::free_slots(int sys) const {
  int n = 0;
  for (int k : sys_[sys].pool) n += slots_[k].used ? 0 : 1;
  return n;
}

template <class T>
int Engine<T>::admit(int sys, const double2* b, double2* x, long long ticket) {
  int id = -1;
  for (int k : sys_[sys].pool)
    if (!slots_[k].used) {
      id = k;
      break;
    }
  if (id < 0) return -1;
  SlotH& h = slots_[id];
  const SysH& s = sys_[sys];
  const size_t n = 3 * (size_t)s.gauge->g.V;
  h.used = true;
  h.finished = false;
  h.status = SFH_OK;
  h.rupd = 0;
  h.b = b;
  h.xh = x;
  h.ticket = ticket;
  h.t_admit = std::chrono::steady_clock::now();
  BSlot<T>& L = h.dev;
  if (!kMixed) L.x = reinterpret_cast<S*>(x);
  L.b = b;
  L.x64 = x;
  hipMemsetAsync(x, 0, n * sizeof(double2), c_->stream);
  const double phi_sq = dev_dot(c_, b, b, n);
  h.phi_sq = phi_sq;
  h.rsq_true = phi_sq;
  CGState z{};
  z.rsq = phi_sq;
  z.target = tol_ * tol_ * phi_sq;
  z.max_iter = max_iter_;
  z.mixed = kMixed ? 1 : 0;
  z.delta = delta_;
  z.rmax = std::sqrt(phi_sq);
  z.ctr = h.ctr;
  if (phi_sq < 1e-28) {  // x = 0 (Fermions.cpp:562-564)
    z.done = 1;
    z.converged = 1;
    h.finished = true;
  }
  h.last = z;
  if (kMixed) {
    f64_to_f32(c_->stream, b, reinterpret_cast<float2*>(L.r), n);
    hipMemsetAsync(L.x, 0, n * sizeof(S), c_->stream);
    hipMemsetAsync(L.v, 0, n * sizeof(S), c_->stream);  // first kb_upd: r - 0 v must stay finite
  } else {
    hipMemcpyAsync(L.r, b, n * sizeof(S), hipMemcpyDeviceToDevice, c_->stream);
  }
  hipMemcpyAsync(L.p, L.r, n * sizeof(S), hipMemcpyDeviceToDevice, c_->stream);
  hipMemcpyAsync(L.st, &z, sizeof(CGState), hipMemcpyHostToDevice, c_->stream);
  if (hip_check(c_, hipStreamSynchronize(c_->stream), "admit")) return -2;
  if (h.finished) {
    done_.push_back(id);
  }
  dirty_ = true;
  return id;
}

// Block map of one hop: (system, block, right-hand-side range) entries, one
// lane per site at `threads` lanes per block, right-hand sides in tiles of
// `tile` (a lane holds a tile's accumulators in registers and loads its links
// once per tile); the tiles of one block index are adjacent, so a block's
// links are still in L2 for the next tile.
template <class T>
void Engine<T>::hop_map(int threads, int tile, std::vector<int4>& out) const {
  out.clear();
  const int hw = threads / 2;
  for (size_t s = 0; s < hsys_.size(); ++s) {
    const BSys<T>& d = hsys_[s];
    if (d.nslot == 0) continue;
    const int nb = (d.g.Vh + hw - 1) / hw;
    for (int b = 0; b < nb; ++b)
      for (int j0 = 0; j0 < d.nslot; j0 += tile) out.push_back(make_int4((int)s, b, j0, std::min(j0 + tile, d.nslot)));
  }
}

template <class T>
int Engine<T>::rebuild() {
  std::vector<int4> mh0, mh;
  std::vector<int2> mb;
  std::vector<BSlot<T>> hl(slots_.size());
  for (size_t k = 0; k < slots_.size(); ++k) hl[k] = slots_[k].dev;
  mapped_.clear();
  hsys_.clear();
  r12_ = r12_64_ = true;
  for (const SysH& sh : sys_) {
    if (sh.removed) continue;
    r12_ = r12_ && sh.su3_lp;
    r12_64_ = r12_64_ && sh.su3_64;
  }
  for (size_t s = 0; s < sys_.size(); ++s) {
    BSys<T> d{};
    d.g = sys_[s].g;
    d.U = sys_[s].U;
    d.U64 = sys_[s].U64;
    d.mass = (R)sys_[s].mass;
    d.mass64 = sys_[s].mass;
    d.nb_hop = (d.g.Vh + 127) / 128;
    d.nslot = 0;
    for (int k : sys_[s].pool)
      if (slots_[k].used && !slots_[k].finished) d.slot[d.nslot++] = k;
    hsys_.push_back(d);
    for (int j = 0; j < d.nslot; ++j) {
      mapped_.push_back(d.slot[j]);
      const int nbb = (d.g.V + kBlas - 1) / kBlas;
      for (int b = 0; b < nbb; ++b) mb.push_back(make_int2(d.slot[j], b));
    }
  }
  const std::vector<BSys<T>>& hs = hsys_;
  hop_map(cfg_.hop_threads, cfg_.rhs_tile, mh0);
  hop_map(256, cfg_.rhs_tile, mh);
  nbh0_ = (int)mh0.size();
  nbh_ = (int)mh.size();
  nbb_ = (int)mb.size();
  finished_mapped_ = 0;
  dirty_ = false;
  if (gexec_) {
    hipGraphExecDestroy(gexec_);
    hipGraphDestroy(graph_);
    gexec_ = nullptr;
    graph_ = nullptr;
  }
  if (!nbh_) return SFH_OK;
  const char* tag = kMixed ? "eng32_" : "eng64_";
  auto key = [&](const char* what) { return std::string(tag) + what; };
  dsys_ = static_cast<BSys<T>*>(scratch(c_, key("sys"), hs.size() * sizeof(BSys<T>)));
  dslot_ = static_cast<BSlot<T>*>(scratch(c_, key("slot"), hl.size() * sizeof(BSlot<T>)));
  dmh0_ = static_cast<int4*>(scratch(c_, key("mh0"), mh0.size() * sizeof(int4)));
  dmh_ = static_cast<int4*>(scratch(c_, key("mh"), mh.size() * sizeof(int4)));
  dmb_ = static_cast<int2*>(scratch(c_, key("mb"), mb.size() * sizeof(int2)));
  dids_ = static_cast<int*>(scratch(c_, key("ids"), mapped_.size() * sizeof(int)));
  dstates_ = static_cast<CGState*>(scratch(c_, key("states"), mapped_.size() * sizeof(CGState)));
  if (!dsys_ || !dslot_ || !dmh0_ || !dmh_ || !dmb_ || !dids_ || !dstates_) return SFH_E_NOMEM;
  if (mapped_.size() > hstates_cap_) {
    if (hstates_) hipHostFree(hstates_);
    hstates_ = nullptr;
    hstates_cap_ = 0;
    if (hipHostMalloc(reinterpret_cast<void**>(&hstates_), slots_.size() * sizeof(CGState)) != hipSuccess)
      return set_err(c_, SFH_E_NOMEM, "engine state staging");
    hstates_cap_ = slots_.size();
  }
  hipMemcpyAsync(dsys_, hs.data(), hs.size() * sizeof(BSys<T>), hipMemcpyHostToDevice, c_->stream);
  hipMemcpyAsync(dslot_, hl.data(), hl.size() * sizeof(BSlot<T>), hipMemcpyHostToDevice, c_->stream);
  hipMemcpyAsync(dmh0_, mh0.data(), mh0.size() * sizeof(int4), hipMemcpyHostToDevice, c_->stream);
  hipMemcpyAsync(dmh_, mh.data(), mh.size() * sizeof(int4), hipMemcpyHostToDevice, c_->stream);
  hipMemcpyAsync(dmb_, mb.data(), mb.size() * sizeof(int2), hipMemcpyHostToDevice, c_->stream);
  hipMemcpyAsync(dids_, mapped_.data(), mapped_.size() * sizeof(int), hipMemcpyHostToDevice, c_->stream);
  if (int e = hip_check(c_, hipStreamSynchronize(c_->stream), "engine tables")) return e;
  if (!use_graph_ || !c_->graph_replay) return SFH_OK;
  if (hipStreamBeginCapture(c_->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return set_err(c_, SFH_E_HIP, "engine graph capture");
  launch_chunk();
  launch_tail();
  if (hipStreamEndCapture(c_->stream, &graph_) != hipSuccess) return set_err(c_, SFH_E_HIP, "engine graph end");
  if (hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0) != hipSuccess)
    return set_err(c_, SFH_E_HIP, "engine graph instantiate");
  return SFH_OK;
}

template <class T>
int Engine<T>::time_hop(const BatchConfig& cfg, int n, double* us) {
  if (dirty_)
    if (int e = rebuild()) return e;
  if (!nbh_) {
    *us = 0.0;
    return SFH_OK;
  }
  std::vector<int4> m;
  hop_map(cfg.hop_threads, cfg.rhs_tile, m);
  int4* dm = static_cast<int4*>(scratch(c_, kMixed ? "eng32_tune" : "eng64_tune", m.size() * sizeof(int4)));
  if (!dm) return SFH_E_NOMEM;
  hipMemcpyAsync(dm, m.data(), m.size() * sizeof(int4), hipMemcpyHostToDevice, c_->stream);
  auto launch = [&]() {
    if (r12_)
      launch_kb_hop<T, false, true>(cfg.rhs_tile, (int)m.size(), cfg.hop_threads, c_->stream, dsys_, dslot_, dm,
                                    (R)(-0.5));
    else
      launch_kb_hop<T, false, false>(cfg.rhs_tile, (int)m.size(), cfg.hop_threads, c_->stream, dsys_, dslot_, dm,
                                     (R)(-0.5));
  };
  launch();
  hipEventRecord(c_->ev0, c_->stream);
  for (int k = 0; k < n; ++k) launch();
  hipEventRecord(c_->ev1, c_->stream);
  if (int e = hip_check(c_, hipEventSynchronize(c_->ev1), "batch hop timing")) return e;
  float ms = 0.f;
  hipEventElapsedTime(&ms, c_->ev0, c_->ev1);
  *us = 1e3 * ms / n;
  return SFH_OK;
}

template <class T>
int Engine<T>::active() const {
  int n = 0;
  for (const auto& h : slots_) n += (h.used && !h.finished) ? 1 : 0;
  return n;
}

// One chunk for every mapped solve, its reliable updates and the state copy;
// then completions on the host.  Finished solves stay in the launch tables
// (their kernels return at once) until a quarter of the tables has finished
// or a solve is admitted, so the captured chunk is replayed, not rebuilt,
// as solves complete.
template <class T>
int Engine<T>::step() {
  if (dirty_)
    if (int e = rebuild()) return e;
  if (!nbh_) return SFH_OK;
  {
    KTimer kt(c_, kMixed ? "batch_chunk_f32" : "batch_chunk_f64");
    if (gexec_) {
      if (hipGraphLaunch(gexec_, c_->stream) != hipSuccess) return set_err(c_, SFH_E_HIP, "engine graph launch");
    } else {
      launch_chunk();
      launch_tail();
    }
  }
  ++chunks_;
  if (int e = hip_check(c_, hipStreamSynchronize(c_->stream), "engine chunk")) return e;
  for (size_t j = 0; j < mapped_.size(); ++j) {
    const int k = mapped_[j];
    SlotH& h = slots_[k];
    const CGState& z = hstates_[j];
    if (!h.used || h.finished || !z.done) continue;
    // stopped: converged on the true residual (mixed), met the iterated
    // target (FP64), or an error status
    h.last = z;
    h.finished = true;
    h.rupd = z.rupd;
    h.status = z.status;
    h.rsq_true = kMixed ? z.rsq_true : z.rsq;
    if (kMixed && z.status == SFH_OK && !z.converged) h.status = h.last.status = SFH_E_NOCONV;
    done_.push_back(k);
    ++finished_mapped_;
  }
  if (4 * finished_mapped_ >= (int)mapped_.size()) dirty_ = true;
  return SFH_OK;
}

template <class T>
std::vector<typename Engine<T>::Done> Engine<T>::take_done() {
  std::vector<Done> out;
  for (int k : done_) {
    SlotH& h = slots_[k];
    Done d;
    d.ticket = h.ticket;
    d.slot = k;
    d.sys = h.sys;
    sfh_cg_stats& o = d.stats;
    std::memset(&o, 0, sizeof o);
    o.iterations = h.last.iter;
    o.reliable_updates = h.rupd;
    o.status = h.status;
    o.rsq = h.last.rsq;
    o.true_rel_resid = h.phi_sq > 0 ? std::sqrt(h.rsq_true / h.phi_sq) : 0.0;
    o.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - h.t_admit).count();
    o.hbm_bytes = (double)h.last.iter * (kMixed ? 888.0 : 1824.0) * h.dev.V;
    d.state = h.last;
    out.push_back(d);
    h.used = false;
  }
  done_.clear();
  return out;
}

template class Engine<float>;
template class Engine<double>;

namespace {

// Batch shape: precision, then lattice extents and right-hand-side count of
// every system, in order.
std::string batch_shape(int nsys, const BatchSys* sys, int precision) {
  std::string k = std::to_string(precision);
  for (int j = 0; j < nsys; ++j) {
    const Geom& g = sys[j].gauge->g;
    k += "|" + std::to_string(g.L[0]) + "," + std::to_string(g.L[1]) + "," + std::to_string(g.L[2]) + "," +
         std::to_string(g.L[3]) + "x" + std::to_string(sys[j].nrhs);
  }
  return k;
}

}  // namespace

int tune_batch(sfh_ctx* c, int nsys, const BatchSys* sys, const sfh_cg_params* p, int* hop_threads, int* rhs_tile,
               double* us) {
  auto tune = [&](auto& eng) -> int {
    long long t = 0;
    for (int k = 0; k < nsys; ++k) {
      const int id = eng.add_system(sys[k].gauge, sys[k].mass, sys[k].nrhs);
      if (id < 0) return -id;
      for (int j = 0; j < sys[k].nrhs; ++j, ++t)
        if (eng.admit(id, sys[k].rhs[j], sys[k].x[j], t) < 0) return SFH_E_HIP;
    }
    BatchConfig best;
    double best_us = 1e30;
    // a shape given on input is recorded as is (timed once for the caller)
    const bool fixed = hop_threads && rhs_tile && *hop_threads > 0 && *rhs_tile > 0;
    for (int threads : {128, 256, 512})
      for (int tile : {1, 2, 4, 8}) {
        if (fixed && (threads != *hop_threads || tile != *rhs_tile)) continue;
        BatchConfig cfg;
        cfg.hop_threads = threads;
        cfg.rhs_tile = tile;
        double u = 0.0;
        if (int e = eng.time_hop(cfg, 5, &u)) return e;
        if (u < best_us) {
          best_us = u;
          best = cfg;
        }
      }
    if (best_us >= 1e30) return set_err(c, SFH_E_INVALID, "batch shape: 128, 256 or 512 sites; 1, 2, 4 or 8 RHS");
    c->batch_cfg[batch_shape(nsys, sys, p->precision)] = {best.hop_threads, best.rhs_tile};
    if (hop_threads) *hop_threads = best.hop_threads;
    if (rhs_tile) *rhs_tile = best.rhs_tile;
    if (us) *us = best_us;
    return SFH_OK;
  };
  if (p->precision == SFH_PREC_F32) {
    Engine<float> eng(c, p->tol, p->max_iter, p->reliable_delta, kBatchChunk);
    return tune(eng);
  }
  if (p->precision == SFH_PREC_F64) {
    Engine<double> eng(c, p->tol, p->max_iter, p->reliable_delta, 8);
    return tune(eng);
  }
  return set_err(c, SFH_E_INVALID, "batched CG supports FP64 and FP32-inner precision");
}

int cg_batch(sfh_ctx* c, int nsys, const BatchSys* sys, const sfh_cg_params* p, sfh_cg_stats* stats) {
  const auto tuned = c->batch_cfg.find(batch_shape(nsys, sys, p->precision));
  auto run = [&](auto& eng) -> int {
    if (tuned != c->batch_cfg.end()) {
      BatchConfig cfg;
      cfg.hop_threads = tuned->second.first;
      cfg.rhs_tile = tuned->second.second;
      eng.set_config(cfg);
    }
    std::vector<long long> first(nsys);
    long long t = 0;
    for (int k = 0; k < nsys; ++k) {
      const int id = eng.add_system(sys[k].gauge, sys[k].mass, sys[k].nrhs);
      if (id < 0) return -id;
      first[k] = t;
      for (int j = 0; j < sys[k].nrhs; ++j, ++t)
        if (eng.admit(id, sys[k].rhs[j], sys[k].x[j], t) < 0) return SFH_E_HIP;
    }
    std::vector<sfh_cg_stats> all((size_t)t);
    std::vector<CGState> states((size_t)t);
    long guard = 0;
    int rc = SFH_OK;
    for (;;) {
      for (auto& d : eng.take_done()) {
        all[d.ticket] = d.stats;
        states[d.ticket] = d.state;
      }
      if (!eng.active()) break;
      if ((rc = eng.step())) return rc;
      if (++guard > (long)p->max_iter / 4 + 1000) return set_err(c, SFH_E_HIP, "batch did not advance");
    }
    int err = SFH_OK;
    for (long long k = 0; k < t; ++k) {
      if (stats) stats[k] = all[k];
      if (all[k].status != SFH_OK && err == SFH_OK) err = map_status(c, states[k], p->max_iter);
    }
    return err;
  };
  if (p->precision == SFH_PREC_F32) {
    Engine<float> eng(c, p->tol, p->max_iter, p->reliable_delta, kBatchChunk);
    return run(eng);
  }
  if (p->precision == SFH_PREC_F64) {
    Engine<double> eng(c, p->tol, p->max_iter, p->reliable_delta, 8);
    return run(eng);
  }
  return set_err(c, SFH_E_INVALID, "batched CG supports FP64 and FP32-inner precision");
}

}  // namespace sfh
