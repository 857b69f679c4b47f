// Batched CG engine: many independent solves (different lattices, masses,
// and up to 8 right-hand sides per lattice sharing its links) advanced
// together, one launch per CG phase for every active solve.
//
// * The hop kernel walks a block map of (system, block) pairs; each block
//   loads its links once and applies them to every active right-hand side
//   of its system (multi-RHS: the link bytes are shared by the RHS block).
// * The vector kernels walk a map of (slot, block) pairs, a slot being one
//   right-hand side in flight.
// * Every slot has its own device CGState and reduction partials, and a
//   system's blocks and reduction order depend only on its own lattice, so a
//   solve gives the same result alone or in any mix.
// * A chunk of iterations is captured into a HIP graph and replayed until the
//   set of active slots changes.
// * Mixed precision (FP32 inner): a slot whose iterated residual has dropped
//   by the reliable-update factor pauses; between chunks the host runs its
//   FP64 reliable update and resumes it while the other slots keep going.
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

// phase DOT=false: t = D^dag p ; DOT=true: v = D t and <p, v>.
template <class T, bool DOT>
__global__ __launch_bounds__(256) void kb_hop(const BSys<T>* __restrict__ sys, const BSlot<T>* __restrict__ slots,
                                              const int2* __restrict__ bmap, typename Prec<T>::R hb) {
  using R = typename Prec<T>::R;
  const int2 m = bmap[blockIdx.x];
  const BSys<T> S = sys[m.x];
  const int lb = m.y;
  const Geom& g = S.g;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int mu = lane >> 4, sl = lane & 15;
  const int p = wave >> 1;
  const int cb = lb * 32 + (wave & 1) * 16 + sl;
  const bool valid = cb < g.Vh;
  const int V = g.V;
  const int i = p * g.Vh + (valid ? cb : 0);
  const int q = (1 - p) * g.Vh;
  int n[4];
  cb_coords(g, valid ? cb : 0, p, n);
  int fw = 0, bw = 0;
  R sf = R(1), sb = R(1);
  mat3<R> Uf, Ub;
  if (valid) {
    neighbours(g, cb, n, mu, fw, bw, sf, sb);
    load_link<T>(S.U, V, mu, i, Uf);
    load_link<T>(S.U, V, mu, q + bw, Ub);
  }
  __shared__ int sdone;
  for (int j = 0; j < S.nslot; ++j) {
    const BSlot<T> L = slots[S.slot[j]];
    if (threadIdx.x == 0) sdone = L.st->done;
    __syncthreads();
    const int dn = sdone;
    __syncthreads();
    if (dn) continue;
    const typename Prec<T>::S* in = DOT ? L.t : L.p;
    typename Prec<T>::S* out = DOT ? L.v : L.t;
    vec3<R> acc;
#pragma unroll
    for (int c = 0; c < 3; ++c) acc.c[c] = {R(0), R(0)};
    if (valid) {
      vec3<R> vf, vb;
      load_spinor<T>(in, nullptr, V, q + fw, vf);
      load_spinor<T>(in, nullptr, V, q + bw, vb);
      acc_mv(acc, Uf, vf, sf);
      acc_mdv(acc, Ub, vb, -sb);
    }
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
    if (valid && mu < 3) {
      using SS = typename Prec<T>::S;
      const cpx<R> self = ldc(in + (size_t)mu * V + i);
      const cpx<R> ac = mu == 0 ? acc.c[0] : (mu == 1 ? acc.c[1] : acc.c[2]);
      const R re = S.mass * self.re + hb * ac.re, im = S.mass * self.im + hb * ac.im;
      out[(size_t)mu * V + i] = SS{re, im};
      if constexpr (DOT) {
        const cpx<R> w = ldc(L.p + (size_t)mu * V + i);
        dotv = (double)w.re * (double)re + (double)w.im * (double)im;
      }
    }
    if constexpr (DOT) grid_reduce(dotv, L.part, ctr_of(L.st->ctr, CNT_PV), S.nb_hop, lb, FinPV{L.st});
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

inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

template <class T>
size_t Engine<T>::slot_bytes(int V) {
  const size_t n = 3 * (size_t)V;
  const int nbh = (V / 2 + 31) / 32, nbb = (V + kBlas - 1) / kBlas;
  return align256(n * sizeof(S)) * (kMixed ? 5 : 4) + (kMixed ? align256(n * sizeof(double2)) : 0) +
         align256((std::max(nbh, nbb) + kRedGroups) * sizeof(double)) + align256(sizeof(CGState)) + align256(kCtrBytes);
}

template <class T>
Engine<T>::Engine(sfh_ctx* c, double tol, int max_iter, double delta, int chunk, bool use_graph)
    : c_(c), tol_(tol), max_iter_(max_iter), delta_(delta > 0 ? delta : 0.1), chunk_(chunk), use_graph_(use_graph) {}

template <class T>
void Engine<T>::launch_chunk() {
  for (int it = 0; it < chunk_; ++it) {
    hipLaunchKernelGGL((kb_hop<T, false>), dim3(nbh_), dim3(256), 0, c_->stream, dsys_, dslot_, dmh_, (R)(-0.5));
    hipLaunchKernelGGL((kb_hop<T, true>), dim3(nbh_), dim3(256), 0, c_->stream, dsys_, dslot_, dmh_, (R)(0.5));
    hipLaunchKernelGGL((kb_xr<T>), dim3(nbb_), dim3(kBlas), 0, c_->stream, dslot_, dmb_);
    hipLaunchKernelGGL((kb_p<T>), dim3(nbb_), dim3(kBlas), 0, c_->stream, dslot_, dmb_);
  }
}

template <class T>
Engine<T>::~Engine() {
  if (gexec_) hipGraphExecDestroy(gexec_);
  if (graph_) hipGraphDestroy(graph_);
  for (void* p : owned_) hipFree(p);
}

template <class T>
int Engine<T>::add_system(const sfh_gauge* gauge, double mass, int pool) {
  if (pool < 1 || pool > kMaxRhs) return set_err(c_, SFH_E_INVALID, "slot pool must be 1..8");
  sfh_gauge* gk = const_cast<sfh_gauge*>(gauge);
  SysH s;
  s.gauge = gauge;
  s.mass = mass;
  if (kMixed) {
    if (int e = ensure_lp(c_, gk, SFH_PREC_F32)) return -e;
    s.U = reinterpret_cast<const typename Prec<T>::L*>(gk->lp.u32);
  } else {
    s.U = reinterpret_cast<const typename Prec<T>::L*>(gk->U);
  }
  const int V = gauge->g.V;
  const size_t sb = slot_bytes(V);
  char* base = nullptr;
  if (hipMalloc(&base, sb * pool) != hipSuccess) return -set_err(c_, SFH_E_NOMEM, "slot pool allocation");
  owned_.push_back(base);
  // reduction counters must start at zero
  if (hipMemsetAsync(base, 0, sb * pool, c_->stream) != hipSuccess || hipStreamSynchronize(c_->stream) != hipSuccess)
    return -set_err(c_, SFH_E_HIP, "slot pool clear");
  bytes_ += sb * pool;
  const size_t n = 3 * (size_t)V;
  for (int k = 0; k < pool; ++k) {
    char* p = base + sb * k;
    auto take = [&](size_t bytes) {
      char* q = p;
      p += align256(bytes);
      return q;
    };
    SlotH h;
    h.sys = (int)sys_.size();
    BSlot<T>& L = h.dev;
    L.V = V;
    // FP64 solves iterate directly in the caller's x (set at admission); the
    // mixed ones keep an FP32 x and an FP64 residual of their own.  The
    // arrays taken here are exactly those counted by slot_bytes().
    if (kMixed) L.x = reinterpret_cast<S*>(take(n * sizeof(S)));
    L.r = reinterpret_cast<S*>(take(n * sizeof(S)));
    L.p = reinterpret_cast<S*>(take(n * sizeof(S)));
    L.t = reinterpret_cast<S*>(take(n * sizeof(S)));
    L.v = reinterpret_cast<S*>(take(n * sizeof(S)));
    if (kMixed) h.r64 = reinterpret_cast<double2*>(take(n * sizeof(double2)));
    L.st = reinterpret_cast<CGState*>(take(sizeof(CGState)));
    h.ctr = reinterpret_cast<unsigned*>(take(kCtrBytes));
    L.part = reinterpret_cast<double*>(p);
    const int nbh = (V / 2 + 31) / 32, nbb = (V + kBlas - 1) / kBlas;
    const size_t used = (size_t)(p - (base + sb * k)) + align256((std::max(nbh, nbb) + kRedGroups) * sizeof(double));
    if (used > sb) return -set_err(c_, SFH_E_INVALID, "slot layout exceeds its allocation");
    s.pool.push_back((int)slots_.size());
    slots_.push_back(h);
  }
  sys_.push_back(s);
  dirty_ = true;
  return (int)sys_.size() - 1;
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
    h.finished = true;
  }
  if (kMixed) {
    hipMemcpyAsync(h.r64, b, n * sizeof(double2), hipMemcpyDeviceToDevice, c_->stream);
    f64_to_f32(c_->stream, h.r64, reinterpret_cast<float2*>(L.r), n);
    hipMemsetAsync(L.x, 0, n * sizeof(S), c_->stream);
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

template <class T>
int Engine<T>::rebuild() {
  std::vector<BSys<T>> hs;
  std::vector<int2> mh, mb;
  std::vector<BSlot<T>> hl(slots_.size());
  for (size_t k = 0; k < slots_.size(); ++k) hl[k] = slots_[k].dev;
  for (size_t s = 0; s < sys_.size(); ++s) {
    BSys<T> d{};
    d.g = sys_[s].gauge->g;
    d.U = sys_[s].U;
    d.mass = (R)sys_[s].mass;
    d.nb_hop = (d.g.Vh + 31) / 32;
    d.nslot = 0;
    for (int k : sys_[s].pool)
      if (slots_[k].used && !slots_[k].finished) d.slot[d.nslot++] = k;
    const int sidx = (int)hs.size();
    hs.push_back(d);
    if (d.nslot == 0) continue;
    for (int b = 0; b < d.nb_hop; ++b) mh.push_back(make_int2(sidx, b));
    for (int j = 0; j < d.nslot; ++j) {
      const int nbb = (d.g.V + kBlas - 1) / kBlas;
      for (int b = 0; b < nbb; ++b) mb.push_back(make_int2(d.slot[j], b));
    }
  }
  nbh_ = (int)mh.size();
  nbb_ = (int)mb.size();
  dirty_ = false;
  if (gexec_) {
    hipGraphExecDestroy(gexec_);
    hipGraphDestroy(graph_);
    gexec_ = nullptr;
    graph_ = nullptr;
  }
  if (!nbh_) return SFH_OK;
  dsys_ = static_cast<BSys<T>*>(scratch(c_, kMixed ? "eng32_sys" : "eng64_sys", hs.size() * sizeof(BSys<T>)));
  dslot_ = static_cast<BSlot<T>*>(scratch(c_, kMixed ? "eng32_slot" : "eng64_slot", hl.size() * sizeof(BSlot<T>)));
  dmh_ = static_cast<int2*>(scratch(c_, kMixed ? "eng32_mh" : "eng64_mh", mh.size() * sizeof(int2)));
  dmb_ = static_cast<int2*>(scratch(c_, kMixed ? "eng32_mb" : "eng64_mb", mb.size() * sizeof(int2)));
  if (!dsys_ || !dslot_ || !dmh_ || !dmb_) return SFH_E_NOMEM;
  hipMemcpyAsync(dsys_, hs.data(), hs.size() * sizeof(BSys<T>), hipMemcpyHostToDevice, c_->stream);
  hipMemcpyAsync(dslot_, hl.data(), hl.size() * sizeof(BSlot<T>), hipMemcpyHostToDevice, c_->stream);
  hipMemcpyAsync(dmh_, mh.data(), mh.size() * sizeof(int2), hipMemcpyHostToDevice, c_->stream);
  hipMemcpyAsync(dmb_, mb.data(), mb.size() * sizeof(int2), hipMemcpyHostToDevice, c_->stream);
  if (int e = hip_check(c_, hipStreamSynchronize(c_->stream), "engine tables")) return e;
  if (!use_graph_) return SFH_OK;
  if (hipStreamBeginCapture(c_->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return set_err(c_, SFH_E_HIP, "engine graph capture");
  launch_chunk();
  if (hipStreamEndCapture(c_->stream, &graph_) != hipSuccess) return set_err(c_, SFH_E_HIP, "engine graph end");
  if (hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0) != hipSuccess)
    return set_err(c_, SFH_E_HIP, "engine graph instantiate");
  return SFH_OK;
}

template <class T>
int Engine<T>::active() const {
  int n = 0;
  for (const auto& h : slots_) n += (h.used && !h.finished) ? 1 : 0;
  return n;
}

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
    }
  }
  ++chunks_;
  std::vector<CGState> hst(slots_.size());
  for (size_t k = 0; k < slots_.size(); ++k)
    if (slots_[k].used && !slots_[k].finished)
      hipMemcpyAsync(&hst[k], slots_[k].dev.st, sizeof(CGState), hipMemcpyDeviceToHost, c_->stream);
  if (int e = hip_check(c_, hipStreamSynchronize(c_->stream), "engine chunk")) return e;
  for (size_t k = 0; k < slots_.size(); ++k) {
    SlotH& h = slots_[k];
    if (!h.used || h.finished || !hst[k].done) continue;
    CGState& z = hst[k];
    h.last = z;
    if (z.status != SFH_OK || !kMixed) {
      h.finished = true;
      h.status = z.status;
      h.rsq_true = z.rsq;
      done_.push_back((int)k);
      dirty_ = true;
      continue;
    }
    // FP64 reliable update
    const SysH& s = sys_[h.sys];
    const size_t n = 3 * (size_t)s.gauge->g.V;
    acc_f32_into_f64(c_->stream, h.xh, reinterpret_cast<float2*>(h.dev.x), n);
    double2* ax = static_cast<double2*>(scratch(c_, "engine_ax", n * sizeof(double2)));
    if (!ax) return SFH_E_NOMEM;
    if (int e = apply_ddag_dev(c_, s.gauge, s.mass, h.xh, ax)) return e;
    sub_f64(c_->stream, h.r64, h.b, ax, n);
    const double rt = dev_dot(c_, h.r64, h.r64, n);
    h.rsq_true = rt;
    h.rupd++;
    if (rt < z.target || z.iter >= max_iter_) {
      h.finished = true;
      h.status = rt < z.target ? SFH_OK : SFH_E_NOCONV;
      h.last.rsq = rt;
      if (h.status) h.last.status = h.status;
      done_.push_back((int)k);
      dirty_ = true;
      continue;
    }
    f64_to_f32(c_->stream, h.r64, reinterpret_cast<float2*>(h.dev.r), n);
    CGState ns = z;
    ns.beta = rt / z.rsq_prev;
    ns.rsq = rt;
    ns.rmax = std::sqrt(rt);
    ns.done = 0;
    ns.pause_reason = 0;
    ns.ru_pending = 0;
    hipMemcpyAsync(h.dev.st, &ns, sizeof(CGState), hipMemcpyHostToDevice, c_->stream);
    cg_p_f32(c_->stream, s.gauge->g, reinterpret_cast<float2*>(h.dev.p), reinterpret_cast<const float2*>(h.dev.r),
             h.dev.st);
    if (int e = hip_check(c_, hipStreamSynchronize(c_->stream), "engine reliable update")) return e;
  }
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
    sfh_cg_stats& o = d.stats;
    std::memset(&o, 0, sizeof o);
    o.iterations = h.last.iter;
    o.reliable_updates = h.rupd;
    o.status = h.status;
    o.rsq = h.last.rsq;
    o.true_rel_resid = h.phi_sq > 0 ? std::sqrt(h.rsq_true / h.phi_sq) : 0.0;
    o.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - h.t_admit).count();
    o.hbm_bytes = (double)h.last.iter * (kMixed ? 912.0 : 1824.0) * h.dev.V;
    d.state = h.last;
    out.push_back(d);
    h.used = false;
  }
  if (!done_.empty()) dirty_ = true;
  done_.clear();
  return out;
}

template class Engine<float>;
template class Engine<double>;

int cg_batch(sfh_ctx* c, int nsys, const BatchSys* sys, const sfh_cg_params* p, sfh_cg_stats* stats) {
  auto run = [&](auto& eng) -> int {
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
    Engine<float> eng(c, p->tol, p->max_iter, p->reliable_delta, 8);
    return run(eng);
  }
  if (p->precision == SFH_PREC_F64) {
    Engine<double> eng(c, p->tol, p->max_iter, p->reliable_delta, 8);
    return run(eng);
  }
  return set_err(c, SFH_E_INVALID, "batched CG supports FP64 and FP32-inner precision");
}

}  // namespace sfh
