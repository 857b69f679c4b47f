// CG drivers (host side).  The iteration runs entirely on the device: the
// step lengths alpha and beta, the stop test and the error checks are computed
// by the last workgroup of each reduction and kept in a device CGState, and
// every kernel returns at once when the state says the solve has stopped.
// The host issues iterations in chunks and reads the state back one chunk
// behind, so the GPU never waits on the host and the result is exactly that
// of a loop which tests convergence after every iteration
// (Fermions.cpp:553-617).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "internal.h"
#include "mixed.h"

namespace sfh {

int map_status(sfh_ctx* c, const CGState& s, int max_iter) {
  char msg[256];
  switch (s.status) {
    case SFH_OK:
      return SFH_OK;
    case SFH_E_BREAKDOWN:
      snprintf(msg, sizeof msg, "CG breakdown: p_dot_v = %f", s.pv);
      return set_err(c, s.status, msg);
    case SFH_E_NONFINITE_ALPHA:
      return set_err(c, s.status, "CG produced non-finite alpha.");
    case SFH_E_NONFINITE_BETA:
      return set_err(c, s.status, "CG produced non-finite beta.");
    case SFH_E_NOCONV:
      snprintf(msg, sizeof msg, "CG Solver failed to converge after %d iterations. Final r_sq = %f, Target = %f",
               max_iter, s.rsq, s.target);
      return set_err(c, s.status, msg);
    default:
      return set_err(c, s.status, "CG failed");
  }
}

double dev_dot(sfh_ctx* c, const double2* a, const double2* b, size_t ncomplex) {
  dot_f64(c->stream, reinterpret_cast<const double*>(a), reinterpret_cast<const double*>(b), 2 * ncomplex,
          c->d_partials + 2 * kPartialSlot, c->aux_ctr(), &c->d_state[1].alpha);
  hipMemcpyAsync(c->h_scalar, &c->d_state[1].alpha, sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (hip_check(c, hipStreamSynchronize(c->stream), "dot")) return NAN;
  return *c->h_scalar;
}

int ensure_su3(sfh_ctx* c, sfh_gauge* gauge) {
  if (gauge->su3_known) return SFH_OK;
  unsigned* dev = static_cast<unsigned*>(scratch(c, "su3_dev", sizeof(unsigned)));
  if (!dev) return SFH_E_NOMEM;
  hipMemsetAsync(dev, 0, sizeof(unsigned), c->stream);
  su3_deviation(c->stream, gauge->g, gauge->U, dev);
  unsigned* h = reinterpret_cast<unsigned*>(c->h_scalar);
  hipMemcpyAsync(h, dev, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream);
  if (int e = hip_check(c, hipStreamSynchronize(c->stream), "SU(3) check")) return e;
  float d;
  std::memcpy(&d, h, sizeof d);
  gauge->lp_su3 = d < kSu3Tol;
  gauge->f64_su3 = d < kSu3TolF64;
  gauge->su3_known = true;
  return SFH_OK;
}

// Low-precision links are derived lazily from the FP64 links and invalidated
// whenever those change; the SU(3) check (ensure_su3) decides whether the
// low-precision kernels read two rows and rebuild the third.
int ensure_lp(sfh_ctx* c, sfh_gauge* gauge, int precision) {
  const Geom& g = gauge->g;
  if (int e = ensure_su3(c, gauge)) return e;
  if (precision == SFH_PREC_F32 && !gauge->lp32_valid) {
    if (!gauge->lp.u32 && hipMalloc(&gauge->lp.u32, 36 * (size_t)g.V * sizeof(float2)) != hipSuccess)
      return set_err(c, SFH_E_NOMEM, "FP32 link copy");
    links_to_f32(c->stream, g, gauge->U, gauge->lp.u32);
    gauge->lp32_valid = true;
  }
  if (precision == SFH_PREC_F16 && !gauge->lp16_valid) {
    if (!gauge->lp.u16 && hipMalloc(&gauge->lp.u16, 36 * (size_t)g.V * sizeof(__half2)) != hipSuccess)
      return set_err(c, SFH_E_NOMEM, "FP16 link copy");
    links_to_f16(c->stream, g, gauge->U, gauge->lp.u16);
    gauge->lp16_valid = true;
  }
  return hip_check(c, hipGetLastError(), "link copy");
}

namespace {

// Iterations per chunk.  Mixed precision: a reliable update that falls due
// runs at the chunk's end, so the chunk length bounds how long the sloppy
// residual drifts past the update point: FP16 storage drifts fastest and
// takes short chunks (32^4, m = 0.01, three right-hand sides and 24^4:
// 10 iterations per chunk 1 % faster than 8 on average, 12 mixed, 14 and
// 16 up to 9 % slower from the drift); FP32 holds its recursion over the
// 300-fold drop of its update threshold and takes longer chunks, which pay
// the chunk's tail and launch less often (24 against 16: the same
// iteration counts from 16^4 to 32^4, 16^4 3 % faster; 48 let a 24^4
// solve drift by 7 iterations).
constexpr int kChunk = 16;
template <bool HALF> constexpr int kChunkMixed = HALF ? 10 : 24;

struct Bufs64 {
  double2 *r, *p, *t, *v;
};

template <class T> T* buf(sfh_ctx* c, const char* key, size_t count) {
  return static_cast<T*>(scratch(c, key, count * sizeof(T)));
}

int upload_state(sfh_ctx* c, const CGState& s) {
  *c->h_state = s;
  c->h_state->ctr = c->d_ctr;  // d_state[0]'s counters (zero between reductions)
  hipMemcpyAsync(c->d_state, c->h_state, sizeof(CGState), hipMemcpyHostToDevice, c->stream);
  return hip_check(c, hipStreamSynchronize(c->stream), "state upload");
}

// A chunk (a fixed number of iterations and, in mixed precision, the gated
// reliable update) captured once into a HIP graph and replayed: one launch
// per chunk instead of one per kernel.  Graphs are cached per context by the exact
// buffers, lattice and parameters they were captured with.  With kernel
// profiling on, chunks are launched directly so every kernel can be timed.
template <class F> int launch_chunk(sfh_ctx* c, const std::string& key, F direct) {
  if (c->profiling || !c->graph_replay) {
    direct();
    return SFH_OK;
  }
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    if (c->graphs.size() >= 32) {
      hipStreamSynchronize(c->stream);
      for (auto& kv : c->graphs) {
        hipGraphExecDestroy(kv.second.second);
        hipGraphDestroy(kv.second.first);
      }
      c->graphs.clear();
    }
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return set_err(c, SFH_E_HIP, "CG graph capture");
    direct();
    if (hipStreamEndCapture(c->stream, &graph) != hipSuccess) return set_err(c, SFH_E_HIP, "CG graph capture end");
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      hipGraphDestroy(graph);
      return set_err(c, SFH_E_HIP, "CG graph instantiate");
    }
    it = c->graphs.emplace(key, std::make_pair(graph, exec)).first;
  }
  return hip_check(c, hipGraphLaunch(it->second.second, c->stream), "CG graph launch");
}

std::string graph_key(const char* kind, const Geom& g, std::initializer_list<const void*> ptrs, double mass,
                      int variant) {
  std::string k = kind;
  char buf[64];
  snprintf(buf, sizeof buf, "|%d,%d,%d,%d|%.17g|%d", g.L[0], g.L[1], g.L[2], g.L[3], mass, variant);
  k += buf;
  for (const void* p : ptrs) {
    snprintf(buf, sizeof buf, "|%p", p);
    k += buf;
  }
  return k;
}

// Runs chunks until the device state says stop; issue_chunk launches one of
// `per_chunk` iterations.  Up to two chunks are in flight and the state is
// read back one chunk behind, so the GPU does not wait for the host.  When
// the residual history predicts that the chunk in flight ends the solve, the
// next chunk is held back until that chunk's state is read: a chunk issued
// after convergence would only run early-exit kernels.  Returns the final
// state through out.
template <class F> int run_chunks(sfh_ctx* c, F issue_chunk, int per_chunk, const CGState& init, CGState& out) {
  CGState* ring[2] = {c->h_state + 1, c->h_state + 2};
  hipEvent_t* ev = c->chunk_ev;  // created with the context, reused by every solve
  long next = 0, head = 0;  // next chunk to issue, oldest chunk in flight
  int rc = SFH_OK;
  auto issue = [&]() {
    const int slot = (int)(next & 1);
    if (int e = issue_chunk()) rc = e;
    state_out(c->stream, c->d_state, ring[slot]);
    if (!rc) rc = hip_check(c, hipEventRecord(ev[slot], c->stream), "CG chunk event");
    ++next;
  };
  const double lr0 = std::log(init.rsq > 0 ? init.rsq : 1.0);
  bool hold = false;
  // every chunk advances the iteration count, runs a reliable update or
  // stops the solve; a state that shows none of these for several polls
  // means the kernels are not running at all
  int last_iter = init.iter, last_rupd = init.rupd, stalled = 0;
  for (;;) {
    while (!rc && (next == head || (!hold && next - head < 2))) issue();
    if (rc) break;
    rc = hip_check(c, hipEventSynchronize(ev[head & 1]), "CG chunk");
    if (!rc) rc = hip_check(c, hipGetLastError(), "CG launch");
    if (rc) break;
    const CGState s = *ring[head & 1];
    ++head;
    if (s.done) break;
    if (s.iter == last_iter && s.rupd == last_rupd) {
      if (++stalled >= 4) {
        rc = set_err(c, SFH_E_HIP, "CG state did not advance");
        break;
      }
    } else {
      stalled = 0;
      last_iter = s.iter;
      last_rupd = s.rupd;
    }
    // iterations still needed at the mean convergence rate so far
    hold = false;
    if (s.iter > 0 && s.rsq > 0 && s.target > 0) {
      const double rate = (lr0 - std::log(s.rsq)) / s.iter;
      if (rate > 0) hold = std::log(s.rsq / s.target) / rate < per_chunk - 0.5;
    }
  }
  const int rs = hip_check(c, hipStreamSynchronize(c->stream), "CG chunks");
  // the final state through the pinned ring, on the context stream
  state_out(c->stream, c->d_state, ring[0]);
  const int rf = hip_check(c, hipStreamSynchronize(c->stream), "CG state");
  out = *ring[0];
  if (rc) return rc;
  if (rs) return rs;
  return rf ? rf : hip_check(c, hipGetLastError(), "CG launches");
}

// Form of the mixed iteration per precision (1 f32, 2 f16): 0 fused kernel
// A, 1 split (in-place update + plain hop), -1 by lattice size.
int g_mix_form[3] = {-1, -1, -1};

bool mix_fused(int prec, const Geom& g) {
  const int f = g_mix_form[prec];
  return f < 0 ? g.V < kMixFusedMaxV : f == 0;
}

// How kernel B hands its dots on: the fused form defers them to the next
// kernel A when every workgroup can read all partials, the split form always
// defers the last level of the reduction to the next update kernel.
int dots_mode(bool fused, const Geom& g) {
  if (!fused) return kDotsGroups;
  return mix_defers(g) ? kDotsPartials : kDotsFinish;
}

// Sloppy vectors of a mixed-precision solve on lattice g.  Fused form:
// kernel A recomputes r_k and p_k at the neighbours, so r and p are double
// buffers; split form: in-place update, then the plain hop.
// (FP16 vectors are one record per site that carries its norm; the norm
// pointers stay null.)
template <class T> int mix_bufs(sfh_ctx* c, const Geom& g, bool fused, MixBufs<T>& b) {
  using S = typename Prec<T>::S;
  const size_t n = 3 * (size_t)g.V;
  const char* names[6] = {"mx_r0", "mx_p0", "mx_v", "mx_t", "mx_r1", "mx_p1"};
  S* vecs[6] = {};
  for (int k = 0; k < (fused ? 6 : 4); ++k) {
    vecs[k] = buf<S>(c, names[k], (size_t)kSpinorElems<T> * g.V);
    if (!vecs[k]) return SFH_E_NOMEM;
  }
  b.r[0] = vecs[0]; b.p[0] = vecs[1]; b.v = vecs[2]; b.t = vecs[3];
  b.r[1] = fused ? vecs[4] : vecs[0];
  b.p[1] = fused ? vecs[5] : vecs[1];
  b.rn[0] = b.rn[1] = b.pn[0] = b.pn[1] = b.vn = b.tn = nullptr;
  b.x = buf<float2>(c, "cgl_x", n);
  return b.x ? SFH_OK : SFH_E_NOMEM;
}

// Step A of one mixed iteration: the fused kernel, or the in-place update
// and the plain hop t = D^dag p_k.
// st_in / st_out: the state copies the step reads and writes (different
// whenever the previous B deferred its finisher, see dots_mode).
template <class T>
void mix_step_a(sfh_ctx* c, const sfh_gauge* gauge, bool fused, const MixBufs<T>& b, float m, const CGState* st_in,
                CGState* st_out, const double* partials, int cur) {
  constexpr bool half = std::is_same<T, Half>::value;
  const Geom& g = gauge->g;
  if (fused) {
    KTimer kt(c, half ? "mix_a_f16" : "mix_a_f32");
    mix_a<T>(c->stream, g, reinterpret_cast<const typename Prec<T>::L*>(half ? (const void*)gauge->lp.u16
                                                                             : (const void*)gauge->lp.u32),
             b, m, st_in, st_out, partials, gauge->lp_su3, cur);
    return;
  }
  CGState* st = st_out;
  {
    KTimer kt(c, "mix_upd");
    mix_upd<T>(c->stream, g, b, st_in, st_out, partials);
  }
  KTimer kt(c, half ? "hop_f16" : "hop_f32");
  if (half)
    hop_f16(c->stream, g, gauge->lp.u16, reinterpret_cast<const H16*>(b.p[0]), reinterpret_cast<H16*>(b.t), m, -1.f,
            nullptr, st, nullptr, gauge->lp_su3);
  else
    hop_f32(c->stream, g, gauge->lp.u32, reinterpret_cast<const float2*>(b.p[0]), reinterpret_cast<float2*>(b.t), m,
            -1.f, nullptr, st, nullptr, gauge->lp_su3);
}

}  // namespace

int apply_ddag_dev(sfh_ctx* c, const sfh_gauge* gauge, double mass, const double2* in, double2* out) {
  const Geom& g = gauge->g;
  double2* t = buf<double2>(c, "ddag_tmp", 3 * (size_t)g.V);
  if (!t) return SFH_E_NOMEM;
  if (int e = ensure_su3(c, const_cast<sfh_gauge*>(gauge))) return e;
  KTimer kt(c, "hop_f64");
  hop_f64(c->stream, g, gauge->U, in, t, mass, -1.0, nullptr, nullptr, nullptr, 0, gauge->f64_su3);
  hop_f64(c->stream, g, gauge->U, t, out, mass, +1.0, nullptr, nullptr, nullptr, 0, gauge->f64_su3);
  return hip_check(c, hipGetLastError(), "DDdag");
}

static int cg_f64(sfh_ctx* c, const sfh_gauge* gauge, const sfh_cg_params* prm, const double2* phi, double2* x,
                  double phi_sq, sfh_cg_stats* stats) {
  const Geom& g = gauge->g;
  const size_t n = 3 * (size_t)g.V;
  if (int e = ensure_su3(c, const_cast<sfh_gauge*>(gauge))) return e;
  Bufs64 b{buf<double2>(c, "cg_r", n), buf<double2>(c, "cg_p", n), buf<double2>(c, "cg_t", n),
           buf<double2>(c, "cg_v", n)};
  if (!b.r || !b.p || !b.t || !b.v) return SFH_E_NOMEM;
  hipMemcpyAsync(b.r, phi, n * sizeof(double2), hipMemcpyDeviceToDevice, c->stream);
  hipMemcpyAsync(b.p, phi, n * sizeof(double2), hipMemcpyDeviceToDevice, c->stream);
  CGState s{};
  s.rsq = phi_sq;
  s.target = prm->tol * prm->tol * phi_sq;
  s.max_iter = prm->max_iter;
  s.done = (s.rsq < s.target || prm->max_iter <= 0) ? 1 : 0;
  if (prm->max_iter <= 0 && !(s.rsq < s.target)) s.status = SFH_E_NOCONV;
  if (int e = upload_state(c, s)) return e;
  CGState* st = c->d_state;
  double* part = c->d_partials;
  CGState fin = s;
  if (!s.done) {
    auto iter = [&]() {
      {
        KTimer kt(c, "hop_f64");
        hop_f64(c->stream, g, gauge->U, b.p, b.t, prm->mass, -1.0, nullptr, st, nullptr, 0, gauge->f64_su3);
      }
      {
        KTimer kt(c, "hop_dot_f64");
        hop_f64(c->stream, g, gauge->U, b.t, b.v, prm->mass, +1.0, b.p, st, part, 0, gauge->f64_su3);
      }
      KTimer kt(c, "cg_blas");
      cg_xr_f64(c->stream, g, x, b.r, b.p, b.v, st, part + kPartialSlot);
      cg_p_f64(c->stream, g, b.p, b.r, st);
    };
    auto direct = [&]() {
      for (int j = 0; j < kChunk; ++j) iter();
    };
    const std::string key = graph_key("f64", g, {gauge->U, x, b.r, b.p, b.t, b.v, st, part}, prm->mass,
                                      get_hop_variant(0, g) * 2 + (gauge->f64_su3 ? 1 : 0));
    if (int e = run_chunks(c, [&]() { return launch_chunk(c, key, direct); }, kChunk, s, fin)) return e;
    // Stop on the FP64 true residual (compute_true_residual >= 2): once the
    // iterated residual meets the target, r = phi - D D^dag x is formed and,
    // while it misses the target, the recurrence restarts from it (p = r)
    // with the iterations counted on.  The reference stops on the iterated
    // residual alone (Fermions.cpp:594-600), which is mode 0 / 1.
    for (int restarts = 0; prm->compute_true_residual >= 2 && fin.status == SFH_OK; ++restarts) {
      if (int e = apply_ddag_dev(c, gauge, prm->mass, x, b.v)) return e;
      sub_f64(c->stream, b.r, phi, b.v, n);
      const double rr = dev_dot(c, b.r, b.r, n);
      if (!std::isfinite(rr)) return set_err(c, SFH_E_NONFINITE_BETA, "CG produced non-finite beta.");
      fin.rsq_true = rr;
      fin.rupd += 1;
      if (rr < s.target) break;
      if (fin.iter >= prm->max_iter || restarts >= 64) {
        fin.status = SFH_E_NOCONV;
        fin.rsq = rr;
        break;
      }
      hipMemcpyAsync(b.p, b.r, n * sizeof(double2), hipMemcpyDeviceToDevice, c->stream);
      CGState r0 = fin;
      r0.rsq = rr;
      r0.alpha = r0.beta = 0.0;
      r0.done = 0;
      if (int e = upload_state(c, r0)) return e;
      if (int e = run_chunks(c, [&]() { return launch_chunk(c, key, direct); }, kChunk, r0, fin)) return e;
    }
  }
  if (stats) {
    stats->iterations = fin.iter;
    stats->reliable_updates = prm->compute_true_residual >= 2 ? fin.rupd : 0;
    stats->rsq = fin.rsq;
    if (prm->compute_true_residual >= 2 && fin.rupd > 0) stats->true_rel_resid = std::sqrt(fin.rsq_true / phi_sq);
    stats->hbm_bytes = (double)fin.iter * 1824.0 * g.V + fin.rupd * 1700.0 * g.V;
  }
  return map_status(c, fin, prm->max_iter);
}

// Mixed-precision CG: the two-kernel iteration of cgmixed.hip in FP32 or
// FP16 storage, with the FP64 reliable update (Clark et al.) run on the
// device at the end of every chunk, gated on the solve state:
//   mix_acc   : (deferred dots: finish the chunk's last iteration first)
//               x += x_lo + alpha p_k, x_lo = 0
//   hop_f64   : t64 = D^dag x                     (gate 1)
//   ru_resid  : r64 = phi - D t64, |r64|^2 -> FinRU (converged, out of
//               budget, or restart with beta = |r_true|^2 / |r_k|^2 and
//               alpha = 0)
//   mix_apply : r_k = r64 in sloppy storage
template <bool HALF>
static int cg_mixed(sfh_ctx* c, sfh_gauge* gauge, const sfh_cg_params* prm, const double2* phi, double2* x,
                    double phi_sq, sfh_cg_stats* stats) {
  using T = typename std::conditional<HALF, Half, float>::type;
  using S = typename Prec<T>::S;
  const Geom& g = gauge->g;
  const size_t n = 3 * (size_t)g.V;
  if (int e = ensure_lp(c, gauge, HALF ? SFH_PREC_F16 : SFH_PREC_F32)) return e;
  double2* r64 = buf<double2>(c, "cg_r", n);
  double2* t64 = buf<double2>(c, "cg_t", n);
  const bool fused = mix_fused(HALF ? SFH_PREC_F16 : SFH_PREC_F32, g);
  MixBufs<T> b{};
  if (int e = mix_bufs<T>(c, g, fused, b)) return e;
  if (!r64 || !t64 || !b.x) return SFH_E_NOMEM;
  // iteration 0 reads r_{-1} = phi from buffer 1, p_{-1} = v_{-1} = 0, alpha = beta = 0
  const size_t ns = (size_t)kSpinorElems<T> * g.V;
  hipMemsetAsync(b.x, 0, n * sizeof(float2), c->stream);
  hipMemsetAsync(b.p[1], 0, ns * sizeof(S), c->stream);
  hipMemsetAsync(b.v, 0, ns * sizeof(S), c->stream);
  if (HALF) {
    f64_to_f16(c->stream, g, phi, reinterpret_cast<H16*>(b.r[1]));
  } else {
    f64_to_f32(c->stream, phi, reinterpret_cast<float2*>(b.r[1]), n);
  }
  CGState s{};
  s.rsq = phi_sq;
  s.target = prm->tol * prm->tol * phi_sq;
  s.max_iter = prm->max_iter;
  s.mixed = 1;
  s.delta = prm->reliable_delta > 0 ? prm->reliable_delta : (HALF ? kDeltaF16 : kDeltaF32);
  s.rmax = std::sqrt(phi_sq);
  s.rsq_true = phi_sq;
  s.pbuf = 1;
  if (int e = upload_state(c, s)) return e;
  CGState* st = c->d_state;
  double* part = c->d_partials;
  double* part_ru = part + 3 * (size_t)kPartialSlot;
  const float m = (float)prm->mass;
  const typename Prec<T>::L* U = reinterpret_cast<const typename Prec<T>::L*>(HALF ? (const void*)gauge->lp.u16
                                                                                  : (const void*)gauge->lp.u32);
  // deferred dots: iteration j of a chunk reads state copy j & 1 and writes
  // copy (j + 1) & 1; the chunk length is even, so the chunk ends in copy 0
  const int dm = dots_mode(fused, g);
  const bool defer = dm != kDotsFinish;
  static_assert(kChunkMixed<HALF> % 2 == 0, "state copies alternate per iteration");
  CGState* sts[2] = {st, defer ? c->d_state + 2 : st};
  auto iter = [&](int j) {
    mix_step_a<T>(c, gauge, fused, b, m, sts[j & 1], sts[(j + 1) & 1], part, j & 1);
    KTimer kt(c, HALF ? "mix_b_f16" : "mix_b_f32");
    mix_b<T>(c->stream, g, U, b, m, sts[(j + 1) & 1], part, dm, gauge->lp_su3, j & 1);
  };
  auto tail = [&]() {
    KTimer kt(c, "reliable_update");
    // with deferred dots the update works on the second state copy (mix_acc
    // finishes the chunk's last B into it) and mix_apply copies it back
    CGState* sru = sts[1];
    mix_acc<T>(c->stream, g, x, b, st, sru, part, dm);
    hop_f64(c->stream, g, gauge->U, x, t64, prm->mass, -1.0, nullptr, sru, nullptr, 1, gauge->f64_su3);
    ru_resid(c->stream, g, gauge->U, t64, phi, r64, prm->mass, sru, part_ru, gauge->f64_su3);
    mix_apply<T>(c->stream, g, r64, b, sru, st);
  };
  auto direct = [&]() {
    for (int j = 0; j < kChunkMixed<HALF>; ++j) iter(j);
    tail();
  };
  const std::string key = graph_key(HALF ? (fused ? "mix16f" : "mix16s") : (fused ? "mix32f" : "mix32s"), g,
                                    {gauge->U, (const void*)U, x, r64, phi, t64, b.r[0], b.r[1], b.p[0], b.p[1],
                                     b.v, b.t, b.rn[0], b.rn[1], b.pn[0], b.pn[1], b.vn, b.tn, b.x, st, sts[1], part},
                                    prm->mass,
                                    ((get_hop_variant(0, g) * 4 + get_hop_variant(HALF ? 2 : 1, g)) * 2 +
                                     (gauge->lp_su3 ? 1 : 0)) * 2 + (gauge->f64_su3 ? 1 : 0));
  CGState fin{};
  if (int e = run_chunks(c, [&]() { return launch_chunk(c, key, direct); }, kChunkMixed<HALF>, s, fin)) return e;
  if (stats) {
    stats->iterations = fin.iter;
    stats->reliable_updates = fin.rupd;
    stats->rsq = fin.rsq;
    stats->true_rel_resid = std::sqrt(fin.rsq_true / phi_sq);
    // per iteration: kernel A (or the update and the hop) reads links, r, v,
    // p, x and writes r, p, x, t; kernel B reads links, t, p, r and writes v
    // (compulsory bytes per site)
    const double per = HALF ? 496.0 : 864.0;
    stats->hbm_bytes = (double)fin.iter * per * g.V + fin.rupd * 1700.0 * g.V;
  }
  return map_status(c, fin, prm->max_iter);
}

int cg_solve(sfh_ctx* c, sfh_gauge* gauge, const sfh_cg_params* prm, const double2* phi, double2* x,
             sfh_cg_stats* stats) {
  const Geom& g = gauge->g;
  const size_t n = 3 * (size_t)g.V;
  if (stats) std::memset(stats, 0, sizeof *stats);
  hipEventRecord(c->ev0, c->stream);
  hipMemsetAsync(x, 0, n * sizeof(double2), c->stream);
  const double phi_sq = dev_dot(c, phi, phi, n);
  if (!std::isfinite(phi_sq) && phi_sq != phi_sq) {
    // NaN source: let the iteration report the breakdown the reference raises
  }
  int rc = SFH_OK;
  if (phi_sq < 1e-28) {
    rc = SFH_OK;  // x = 0 (Fermions.cpp:562-564)
  } else if (prm->precision == SFH_PREC_F64) {
    rc = cg_f64(c, gauge, prm, phi, x, phi_sq, stats);
  } else if (prm->precision == SFH_PREC_F32) {
    rc = cg_mixed<false>(c, gauge, prm, phi, x, phi_sq, stats);
  } else if (prm->precision == SFH_PREC_F16) {
    rc = cg_mixed<true>(c, gauge, prm, phi, x, phi_sq, stats);
  } else {
    return set_err(c, SFH_E_INVALID, "unknown CG precision");
  }
  hipEventRecord(c->ev1, c->stream);
  hipEventSynchronize(c->ev1);
  if (stats) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    stats->seconds = ms * 1e-3;
    stats->status = rc;
    if (rc == SFH_OK && prm->precision == SFH_PREC_F64 && prm->compute_true_residual == 1 && phi_sq >= 1e-28) {
      double2* ax = buf<double2>(c, "cg_t", n);
      double2* rr = buf<double2>(c, "cg_r", n);
      if (ax && rr && apply_ddag_dev(c, gauge, prm->mass, x, ax) == SFH_OK) {
        sub_f64(c->stream, rr, phi, ax, n);
        stats->true_rel_resid = std::sqrt(dev_dot(c, rr, rr, n) / phi_sq);
      }
    }
  }
  return rc;
}

// Times the two steps of the mixed iteration on inputs of ones with a live
// (never finishing) solve state, n repetitions of each between HIP events.
template <class T>
static int time_mixed_t(sfh_ctx* c, sfh_gauge* gauge, int n, double* us_a, double* us_b) {
  constexpr bool half = std::is_same<T, Half>::value;
  const Geom& g = gauge->g;
  const size_t n3 = 3 * (size_t)g.V;
  if (int e = ensure_lp(c, gauge, half ? SFH_PREC_F16 : SFH_PREC_F32)) return e;
  const bool fused = mix_fused(half ? SFH_PREC_F16 : SFH_PREC_F32, g);
  MixBufs<T> b{};
  if (int e = mix_bufs<T>(c, g, fused, b)) return e;
  for (int k = 0; k < 2; ++k) {
    for (void* v : {(void*)b.r[k], (void*)b.p[k]}) {
      if (half) h16_ones(c->stream, static_cast<H16*>(v), g.V);
      else hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(v), 0x3F800000u, 2 * n3, c->stream);  // 1.0f
    }
  }
  hipMemsetAsync(b.v, 0, (size_t)kSpinorElems<T> * g.V * sizeof(typename Prec<T>::S), c->stream);
  hipMemsetAsync(b.x, 0, n3 * sizeof(float2), c->stream);
  // alpha = beta = 0 keeps kernel A idempotent; target 0 and no budget keep B running
  CGState* z = c->h_state;  // pinned staging
  *z = CGState{};
  z->rsq = 1.0;
  z->max_iter = 0x7fffffff;
  z->ctr = c->d_ctr + kCtrKinds * kCtrPerKind;
  CGState* st = &c->d_state[1];
  hipMemcpyAsync(st, z, sizeof(CGState), hipMemcpyHostToDevice, c->stream);
  if (int e = hip_check(c, hipStreamSynchronize(c->stream), "CG kernel timing state")) return e;
  double* part = c->d_partials + 4 * (size_t)kPartialSlot;
  const float m = 0.1f;
  const auto* U = reinterpret_cast<const typename Prec<T>::L*>(half ? (const void*)gauge->lp.u16
                                                                    : (const void*)gauge->lp.u32);
  auto timed = [&](auto launch, double* us) {
    launch();
    hipEventRecord(c->ev0, c->stream);
    for (int k = 0; k < n; ++k) launch();
    hipEventRecord(c->ev1, c->stream);
    if (int e = hip_check(c, hipEventSynchronize(c->ev1), "CG kernel timing")) return e;
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    *us = 1e3 * ms / n;
    return (int)SFH_OK;
  };
  // A reads st and (deferred dots) writes a scratch copy, so every launch
  // sees the same state; B changes alpha and beta or only leaves partials
  const int dm = dots_mode(fused, g);
  CGState* st_out = dm != kDotsFinish ? &c->d_state[3] : st;
  int rc = timed([&]() { mix_step_a<T>(c, gauge, fused, b, m, st, st_out, part, 0); }, us_a);
  if (!rc) rc = timed([&]() { mix_b<T>(c->stream, g, U, b, m, st, part, dm, gauge->lp_su3, 0); }, us_b);
  reset_aux_state(c);
  return rc;
}

void set_mix_form(int precision, int form) {
  if (precision == SFH_PREC_F32 || precision == SFH_PREC_F16) g_mix_form[precision] = form;
}
int get_mix_form(int precision, const Geom& g) { return mix_fused(precision, g) ? 0 : 1; }

int time_mixed(sfh_ctx* c, sfh_gauge* gauge, int precision, int n, double* us_a, double* us_b) {
  return precision == SFH_PREC_F16 ? time_mixed_t<Half>(c, gauge, n, us_a, us_b)
                                   : time_mixed_t<float>(c, gauge, n, us_a, us_b);
}

}  // namespace sfh
