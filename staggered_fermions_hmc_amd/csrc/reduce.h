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
// fences: the partial is written with an agent-scope atomic store (coherent
// at device scope), the wave waits for that store to complete, then counts
// its arrival with an agent-scope atomic add; the last workgroup reads the
// partials with agent-scope atomic loads.  No L2 write-back or invalidate is
// issued, so the kernel's other traffic keeps its cache reuse.  What the
// finisher writes is read by later kernels, after the kernel boundary.
//
// Arrivals are counted in two levels (kRedGroups group counters on separate
// cache lines, then one top counter): same-address atomics serialise in the
// memory system, so one counter for thousands of workgroups would add
// microseconds to every reduction.
template <class Fin>
__device__ __forceinline__ void grid_reduce(double v, double* partials, unsigned* ctr, int nblocks, int slot,
                                            Fin fin) {
  const double bs = block_sum(v);
  const int ng = nblocks < kRedGroups ? nblocks : kRedGroups;
  __shared__ int is_last;
  if (threadIdx.x == 0) {
    __hip_atomic_store(&partials[slot], bs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);  // the partial's store has completed before the arrival is counted
    const int grp = slot % ng;
    const unsigned gsize = unsigned(nblocks / ng + (grp < nblocks % ng ? 1 : 0));
    int last = 0;
    const unsigned t = __hip_atomic_fetch_add(&ctr[grp * kCtrStride], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t == gsize - 1) {
      const unsigned u =
          __hip_atomic_fetch_add(&ctr[kRedGroups * kCtrStride], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = (u == unsigned(ng - 1));
    }
    is_last = last;
  }
  __syncthreads();
  if (is_last) {
    double s = 0.0;
    for (int k = threadIdx.x; k < nblocks; k += blockDim.x)
      s += __hip_atomic_load(&partials[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s = block_sum(s);
    for (int k = threadIdx.x; k < ng; k += blockDim.x)
      __hip_atomic_store(&ctr[k * kCtrStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x == 0) {
      fin(s);
      __hip_atomic_store(&ctr[kRedGroups * kCtrStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
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
// precision it also pauses the solve when a reliable update is due or the
// iterated residual meets the target, for the host to run the FP64 update.
struct FinRR {
  CGState* st;
  __device__ void operator()(double rr) const {
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
      // pause for an FP64 reliable update: it decides convergence and the
      // budget on the true residual
      const double r = sqrt(rr);
      int reason = 0;
      if (rr < st->target) reason = 2;
      else if (r < st->delta * st->rmax) reason = 1;
      else if (st->iter >= st->max_iter) reason = 3;
      if (reason) {
        st->pause_reason = reason;
        st->ru_pending = 1;
        st->done = 1;
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
    st->beta = rr / st->rsq_prev;
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
