// Deterministic grid reductions: every workgroup writes its partial sum, the
// last workgroup to arrive (arrivals counted with device-scope atomics)
// sums all partials in index order and runs the finisher.  The result does
// not depend on arrival order, only on the grid shape, so it is reproducible
// run to run and independent of what else shares the GPU.
#pragma once

#include "common.h"
#include "internal.h"

namespace sfh {

//
// Coherence across the 8 XCDs (each with its own L2) without cache-wide
// fences: partial sums are written with agent-scope atomic stores (coherent
// at device scope), the wave waits for each store to complete, then counts
// its arrival with an agent-scope atomic add; partials are read back with
// agent-scope atomic loads.  No L2 write-back or invalidate is issued, so the
// kernel's other traffic keeps its cache reuse.  What the finisher writes is
// read by later kernels, after the kernel boundary.
//
// Arrivals are counted in two levels: workgroup `slot` belongs to group
// slot % ng (ng = min(nblocks, kRedGroups)), whose counter sits on a cache
// line of its own; the last arrival of each group counts on one top counter.
// Same-address atomics serialise in the memory system, so one counter for
// thousands of workgroups would add microseconds to every reduction.  Up to
// kDirectSum workgroups, the last one sums all partials itself (in slot
// order); above that, the last arrival of each group first sums its group's
// partials (in slot order) into a group partial, and the last group sums the
// ng group partials (in group order), so no workgroup reads more than a few
// partials per lane.
//
// N values are reduced together; value j of workgroup b is kept at
// partials[j * (nblocks + kRedGroups) + b], so partials must hold
// N * (nblocks + kRedGroups) doubles.  fin receives the N sums.
constexpr int kDirectSum = 1024;

// Four sums over the 64 lanes of a wave: the lanes first exchange half of
// their values with the partner across lane bit 0 and then across bit 1, so
// each lane carries one value into the four butterfly steps that remain:
// 7 exchanges of a double instead of 24 (each a cross-lane permute through
// the LDS crossbar).  Every lane gets the total of value
// slot = 2 (lane & 1) + ((lane >> 1) & 1), the same bits in every lane of
// that slot; the order of the additions is fixed.
__device__ __forceinline__ double wave_sum4(const double (&v)[4], int& slot) {
  const int l = threadIdx.x & 63;
  const bool b0 = l & 1, b1 = (l >> 1) & 1;
  const double w0 = (b0 ? v[2] : v[0]) + __shfl_xor(b0 ? v[0] : v[2], 1, 64);
  const double w1 = (b0 ? v[3] : v[1]) + __shfl_xor(b0 ? v[1] : v[3], 1, 64);
  double x = (b1 ? w1 : w0) + __shfl_xor(b1 ? w0 : w1, 2, 64);
#pragma unroll
  for (int off = 4; off < 64; off <<= 1) x += __shfl_xor(x, off, 64);
  slot = (b0 ? 2 : 0) + (b1 ? 1 : 0);
  return x;
}
// The four totals of wave_sum4 in every lane of the wave (lane 2 a + b
// holds slot 2 b + a).
__device__ __forceinline__ void wave_sum4_all(double (&v)[4]) {
  int slot;
  const double x = wave_sum4(v, slot);
  v[0] = __shfl(x, 0, 64);
  v[1] = __shfl(x, 2, 64);
  v[2] = __shfl(x, 1, 64);
  v[3] = __shfl(x, 3, 64);
}

template <int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N]) {
  __shared__ double ws[16][N];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if constexpr (N == 4) {
    int slot;
    const double x = wave_sum4(v, slot);
    if (lane < 4) ws[w][slot] = x;
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = wave_sum(v[j]);
    if (lane == 0) {
#pragma unroll
      for (int j = 0; j < N; ++j) ws[w][j] = v[j];
    }
  }
  __syncthreads();
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double t = 0.0;
    for (int k = 0; k < nw; ++k) t += ws[k][j];
    v[j] = t;
  }
  __syncthreads();
}

template <int N, class Fin>
__device__ __forceinline__ void grid_reduce_n(double (&v)[N], double* partials, unsigned* ctr, int nblocks, int slot,
                                              Fin fin) {
  block_sum_n<N>(v);
  const int ng = nblocks < kRedGroups ? nblocks : kRedGroups;
  const int grp = slot % ng;
  const size_t stride = (size_t)nblocks + kRedGroups;
  const bool two = nblocks > kDirectSum;
  __shared__ int group_last, grid_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < N; ++j)
      __hip_atomic_store(&partials[j * stride + slot], v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);  // the partials' stores have completed before the arrival is counted
    const unsigned gsize = unsigned(nblocks / ng + (grp < nblocks % ng ? 1 : 0));
    const unsigned t = __hip_atomic_fetch_add(&ctr[grp * kCtrStride], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    group_last = (t == gsize - 1);
  }
  __syncthreads();
  if (!group_last) return;
  if (two) {
    double s[N];
#pragma unroll
    for (int j = 0; j < N; ++j) s[j] = 0.0;
    for (int k = grp + ng * (int)threadIdx.x; k < nblocks; k += ng * (int)blockDim.x) {
#pragma unroll
      for (int j = 0; j < N; ++j)
        s[j] += __hip_atomic_load(&partials[j * stride + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum_n<N>(s);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int j = 0; j < N; ++j)
        __hip_atomic_store(&partials[j * stride + nblocks + grp], s[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (threadIdx.x == 0) {
    __hip_atomic_store(&ctr[grp * kCtrStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    const unsigned u =
        __hip_atomic_fetch_add(&ctr[kRedGroups * kCtrStride], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    grid_last = (u == unsigned(ng - 1));
  }
  __syncthreads();
  if (!grid_last) return;
  double s[N];
#pragma unroll
  for (int j = 0; j < N; ++j) s[j] = 0.0;
  const int base = two ? nblocks : 0, cnt = two ? ng : nblocks;
  for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
#pragma unroll
    for (int j = 0; j < N; ++j)
      s[j] += __hip_atomic_load(&partials[j * stride + base + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum_n<N>(s);
  if (threadIdx.x == 0) {
    fin(s);
    __hip_atomic_store(&ctr[kRedGroups * kCtrStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The group level of grid_reduce_n alone: the last arrival of each group sums
// its group's partials (in slot order, the same sums as grid_reduce_n's) into
// partials[j * (nblocks + kRedGroups) + nblocks + grp] and resets the group's
// counter; the min(nblocks, kRedGroups) group partials are summed by a later
// kernel (load_groups_n + block_sum_n gives grid_reduce_n's sums bit for bit
// when nblocks > kDirectSum).  No workgroup waits on a grid-wide counter.
template <int N>
__device__ __forceinline__ void group_reduce_n(double (&v)[N], double* partials, unsigned* ctr, int nblocks,
                                               int slot) {
  block_sum_n<N>(v);
  const int ng = nblocks < kRedGroups ? nblocks : kRedGroups;
  const int grp = slot % ng;
  const size_t stride = (size_t)nblocks + kRedGroups;
  __shared__ int group_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < N; ++j)
      __hip_atomic_store(&partials[j * stride + slot], v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);  // the partials' stores have completed before the arrival is counted
    const unsigned gsize = unsigned(nblocks / ng + (grp < nblocks % ng ? 1 : 0));
    const unsigned t = __hip_atomic_fetch_add(&ctr[grp * kCtrStride], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    group_last = (t == gsize - 1);
  }
  __syncthreads();
  if (!group_last) return;
  double s[N];
#pragma unroll
  for (int j = 0; j < N; ++j) s[j] = 0.0;
  for (int k = grp + ng * (int)threadIdx.x; k < nblocks; k += ng * (int)blockDim.x) {
#pragma unroll
    for (int j = 0; j < N; ++j)
      s[j] += __hip_atomic_load(&partials[j * stride + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum_n<N>(s);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < N; ++j)
      __hip_atomic_store(&partials[j * stride + nblocks + grp], s[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ctr[grp * kCtrStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The group partials group_reduce_n left, one group per lane (lanes past the
// group count hold zeros), for block_sum_n in a kernel after it.
template <int N>
__device__ __forceinline__ void load_groups_n(const double* __restrict__ partials, int nblocks, double (&v)[N]) {
  const int ng = nblocks < kRedGroups ? nblocks : kRedGroups;
  const size_t stride = (size_t)nblocks + kRedGroups;
  const int k = (int)threadIdx.x;
#pragma unroll
  for (int j = 0; j < N; ++j) v[j] = 0.0;
  if (k < ng) {  // whole waves past the group count load nothing
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = partials[j * stride + nblocks + k];
  }
}

template <class Fin> struct FinOne {
  Fin f;
  __device__ void operator()(const double* s) const { f(s[0]); }
};

// One value: partials must hold nblocks + kRedGroups doubles.
template <class Fin>
__device__ __forceinline__ void grid_reduce(double v, double* partials, unsigned* ctr, int nblocks, int slot,
                                            Fin fin) {
  double a[1] = {v};
  grid_reduce_n<1>(a, partials, ctr, nblocks, slot, FinOne<Fin>{fin});
}

// Finisher of <p, DDdag p>: alpha = rsq / pv with the reference's breakdown
// checks (Fermions.cpp:573-586).
struct FinPV {
  CGState* st;
  __device__ void operator()(double pv) const {
    st->pv = pv;
    if (!isfinite(pv) || fabs(pv) < 1e-30) {
      st->status = SFH_E_BREAKDOWN;
      st->done = 1;
      return;
    }
    const double a = st->rsq / pv;
    if (!isfinite(a)) {
      st->status = SFH_E_NONFINITE_ALPHA;
      st->done = 1;
      return;
    }
    st->alpha = a;
  }
};

// Finisher of |r|^2 after x += alpha p, r -= alpha v: beta, convergence and
// the iteration budget (Fermions.cpp:568, :594-600, :609-614).  In mixed
// precision it also marks a reliable update due, and pauses the solve when
// the iterated residual meets the target, for the FP64 update to decide.
// Iteration bookkeeping once |r_{k+1}|^2 = rr is known: beta, the
// convergence test, and in mixed precision the reliable-update bookkeeping.
__device__ __forceinline__ void finish_rr(CGState* st, double rr) {
  const double b = rr / st->rsq;
  st->iter += 1;
  st->rsq_prev = st->rsq;
  st->rsq = rr;
  if (!isfinite(b)) {
    st->status = SFH_E_NONFINITE_BETA;
    st->done = 1;
    return;
  }
  st->beta = b;
  if (st->mixed) {
    // FP64 reliable update, run at the end of the chunk: due when the
    // residual has fallen by delta since the last one (the iterations go on
    // to the chunk's end), and the solve pauses for it when the iterated
    // residual meets the target or the budget is spent (the update decides
    // convergence and the budget on the true residual)
    const double r = sqrt(rr);
    int reason = 0;
    if (rr < st->target) reason = 2;
    else if (st->iter >= st->max_iter) reason = 3;
    else if (r < st->delta * st->rmax) reason = 1;
    if (reason) {
      if (reason > st->pause_reason) st->pause_reason = reason;
      st->ru_pending = 1;
      if (reason != 1) st->done = 1;
    }
    if (r > st->rmax) st->rmax = r;
    return;
  }
  if (rr < st->target) {
    st->done = 1;
  } else if (st->iter >= st->max_iter) {
    st->status = SFH_E_NOCONV;
    st->done = 1;
  }
}

// Finisher of |r_{k+1}|^2 (Fermions.cpp:589-614).
struct FinRR {
  CGState* st;
  __device__ void operator()(double rr) const { finish_rr(st, rr); }
};

// Finisher of the two-kernel mixed-precision iteration: s = {<p,v>, <r,v>,
// <v,v>, <r,r>} of the stored p_k, r_k and v = DDdag p_k.  alpha = |r_k|^2 /
// <p,v> (breakdown checks of Fermions.cpp:573-586) and |r_{k+1}|^2 =
// |r_k - alpha v|^2 expanded, from this iteration's dots only.
__device__ __forceinline__ void finish_fused(CGState* st, const double* s) {
  const double pv = s[0], rv = s[1], vv = s[2], rr = s[3];
  st->pv = pv;
  st->rsq = rr;
  if (!isfinite(pv) || fabs(pv) < 1e-30) {
    st->status = SFH_E_BREAKDOWN;
    st->done = 1;
    return;
  }
  const double a = rr / pv;
  if (!isfinite(a)) {
    st->status = SFH_E_NONFINITE_ALPHA;
    st->done = 1;
    return;
  }
  st->alpha = a;
  double rn = rr - 2.0 * a * rv + a * a * vv;
  // cancellation can only take it below zero when r_{k+1} is tiny: the
  // reliable update that this triggers measures the true residual
  if (!(rn > 0.0)) rn = 0.0;
  finish_rr(st, rn);
}

struct FinFused {
  CGState* st;
  __device__ void operator()(const double* s) const { finish_fused(st, s); }
};

// Finisher of the FP64 true residual |phi - DDdag x|^2 of a reliable update
// (Clark et al., reliable updates): converged, out of budget, or restart the
// sloppy residual from the true one with beta = |r_true|^2 / |r_prev|^2.
struct FinRU {
  CGState* st;
  __device__ void operator()(double rr) const {
    st->rsq_true = rr;
    st->rupd += 1;
    st->ru_pending = 0;
    if (rr < st->target) {
      st->converged = 1;
      st->done = 1;
      return;
    }
    if (!isfinite(rr)) {
      st->status = SFH_E_NONFINITE_BETA;
      st->done = 1;
      return;
    }
    if (st->iter >= st->max_iter) {
      st->status = SFH_E_NOCONV;
      st->rsq = rr;
      st->done = 1;
      return;
    }
    // the true residual replaces the iterated one; the direction is kept
    // (beta = |r_true|^2 / |r_k|^2) unless the two residuals have drifted
    // apart (|r_true| > 10 |r_iterated|, e.g. when no update fell due before
    // the iterated residual met the target): the old direction then no
    // longer fits the residual and the iteration starts again from p = r_true
    st->beta = rr > 100.0 * st->rsq ? 0.0 : rr / st->rsq_prev;
    st->alpha = 0.0;  // the two-kernel iteration then restarts from r_true
    st->rsq = rr;
    st->rmax = sqrt(rr);
    st->ru_apply = 1;
    st->pause_reason = 0;
    st->done = 0;
  }
};

struct FinStore {
  double* out;
  __device__ void operator()(double s) const { *out = s; }
};

}  // namespace sfh
