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

// Low-precision links are derived lazily from the FP64 links and invalidated
// whenever those change.
int ensure_lp(sfh_ctx* c, sfh_gauge* gauge, int precision) {
  const Geom& g = gauge->g;
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

constexpr int kChunk = 16;

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

// A chunk (kChunk iterations and, in mixed precision, the gated reliable
// update) captured once into a HIP graph and replayed: one launch per chunk
// instead of 4 kChunk + 5.  Graphs are cached per context by the exact
// buffers, lattice and parameters they were captured with.  With kernel
// profiling on, chunks are launched directly so every kernel can be timed.
template <class F> int launch_chunk(sfh_ctx* c, const std::string& key, F direct) {
  if (c->profiling) {
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

// Runs chunks until the device state says stop; issue_chunk launches one.
// The state is read back one chunk behind, so the GPU never waits for the
// host.  Returns the final state through out.
template <class F> int run_chunks(sfh_ctx* c, F issue_chunk, CGState& out) {
  CGState* ring[2] = {c->h_state + 1, c->h_state + 2};
  hipEvent_t ev[2];
  hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
  hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
  int k = 0, inflight = 0, rc = SFH_OK;
  auto issue = [&](int slot) {
    if (int e = issue_chunk()) rc = e;
    hipMemcpyAsync(ring[slot], c->d_state, sizeof(CGState), hipMemcpyDeviceToHost, c->stream);
    hipEventRecord(ev[slot], c->stream);
  };
  issue(0);
  inflight = 1;
  for (;;) {
    const int cur = k & 1;
    issue(cur ^ 1);
    ++inflight;
    if (rc) break;
    rc = hip_check(c, hipEventSynchronize(ev[cur]), "CG chunk");
    --inflight;
    if (!rc) rc = hip_check(c, hipGetLastError(), "CG launch");
    if (rc) break;
    if (ring[cur]->done) break;
    ++k;
    // every iteration advances the device counter or stops the solve, so the
    // budget is exceeded only if kernels are not running at all
    if ((long)k * kChunk > (long)ring[cur]->max_iter + 4L * kChunk) {
      rc = set_err(c, SFH_E_HIP, "CG state did not advance");
      break;
    }
  }
  hipStreamSynchronize(c->stream);
  hipMemcpy(&out, c->d_state, sizeof(CGState), hipMemcpyDeviceToHost);
  hipEventDestroy(ev[0]);
  hipEventDestroy(ev[1]);
  (void)inflight;
  return rc ? rc : hip_check(c, hipGetLastError(), "CG launches");
}

}  // namespace

int apply_ddag_dev(sfh_ctx* c, const sfh_gauge* gauge, double mass, const double2* in, double2* out) {
  const Geom& g = gauge->g;
  double2* t = buf<double2>(c, "ddag_tmp", 3 * (size_t)g.V);
  if (!t) return SFH_E_NOMEM;
  KTimer kt(c, "hop_f64");
  hop_f64(c->stream, g, gauge->U, in, t, mass, -1.0, nullptr, nullptr, nullptr);
  hop_f64(c->stream, g, gauge->U, t, out, mass, +1.0, nullptr, nullptr, nullptr);
  return hip_check(c, hipGetLastError(), "DDdag");
}

static int cg_f64(sfh_ctx* c, const sfh_gauge* gauge, const sfh_cg_params* prm, const double2* phi, double2* x,
                  double phi_sq, sfh_cg_stats* stats) {
  const Geom& g = gauge->g;
  const size_t n = 3 * (size_t)g.V;
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
        hop_f64(c->stream, g, gauge->U, b.p, b.t, prm->mass, -1.0, nullptr, st, nullptr);
      }
      {
        KTimer kt(c, "hop_dot_f64");
        hop_f64(c->stream, g, gauge->U, b.t, b.v, prm->mass, +1.0, b.p, st, part);
      }
      KTimer kt(c, "cg_blas");
      cg_xr_f64(c->stream, g, x, b.r, b.p, b.v, st, part + kPartialSlot);
      cg_p_f64(c->stream, g, b.p, b.r, st);
    };
    auto direct = [&]() {
      for (int j = 0; j < kChunk; ++j) iter();
    };
    const std::string key = graph_key("f64", g, {gauge->U, x, b.r, b.p, b.t, b.v, st, part}, prm->mass,
                                      get_hop_variant(0, g));
    if (int e = run_chunks(c, [&]() { return launch_chunk(c, key, direct); }, fin)) return e;
  }
  if (stats) {
    stats->iterations = fin.iter;
    stats->rsq = fin.rsq;
    stats->hbm_bytes = (double)fin.iter * 1824.0 * g.V;
  }
  return map_status(c, fin, prm->max_iter);
}

template <bool HALF>
static int cg_mixed(sfh_ctx* c, sfh_gauge* gauge, const sfh_cg_params* prm, const double2* phi, double2* x,
                    double phi_sq, sfh_cg_stats* stats) {
  const Geom& g = gauge->g;
  const size_t n = 3 * (size_t)g.V;
  double2* r64 = buf<double2>(c, "cg_r", n);
  float2* xl = buf<float2>(c, "cgl_x", n);
  if (!r64 || !xl) return SFH_E_NOMEM;
  if (int e = ensure_lp(c, gauge, HALF ? SFH_PREC_F16 : SFH_PREC_F32)) return e;
  // sloppy vectors
  float2 *r32 = nullptr, *p32 = nullptr, *t32 = nullptr, *v32 = nullptr;
  __half2 *r16 = nullptr, *p16 = nullptr, *t16 = nullptr, *v16 = nullptr;
  float *rn = nullptr, *pn = nullptr, *tn = nullptr, *vn = nullptr;
  if (HALF) {
    r16 = buf<__half2>(c, "cgh_r", n); p16 = buf<__half2>(c, "cgh_p", n);
    t16 = buf<__half2>(c, "cgh_t", n); v16 = buf<__half2>(c, "cgh_v", n);
    rn = buf<float>(c, "cgh_rn", g.V); pn = buf<float>(c, "cgh_pn", g.V);
    tn = buf<float>(c, "cgh_tn", g.V); vn = buf<float>(c, "cgh_vn", g.V);
    if (!r16 || !p16 || !t16 || !v16 || !rn || !pn || !tn || !vn) return SFH_E_NOMEM;
  } else {
    r32 = buf<float2>(c, "cgl_r", n); p32 = buf<float2>(c, "cgl_p", n);
    t32 = buf<float2>(c, "cgl_t", n); v32 = buf<float2>(c, "cgl_v", n);
    if (!r32 || !p32 || !t32 || !v32) return SFH_E_NOMEM;
  }
  hipMemsetAsync(xl, 0, n * sizeof(float2), c->stream);
  hipMemcpyAsync(r64, phi, n * sizeof(double2), hipMemcpyDeviceToDevice, c->stream);
  auto to_sloppy_r = [&]() {
    if (HALF) f64_to_f16(c->stream, g, r64, r16, rn);
    else f64_to_f32(c->stream, r64, r32, n);
  };
  to_sloppy_r();
  if (HALF) {
    hipMemcpyAsync(p16, r16, n * sizeof(__half2), hipMemcpyDeviceToDevice, c->stream);
    hipMemcpyAsync(pn, rn, g.V * sizeof(float), hipMemcpyDeviceToDevice, c->stream);
  } else {
    hipMemcpyAsync(p32, r32, n * sizeof(float2), hipMemcpyDeviceToDevice, c->stream);
  }
  CGState s{};
  s.rsq = phi_sq;
  s.target = prm->tol * prm->tol * phi_sq;
  s.max_iter = prm->max_iter;
  s.mixed = 1;
  s.delta = prm->reliable_delta > 0 ? prm->reliable_delta : 0.1;
  s.rmax = std::sqrt(phi_sq);
  s.rsq_true = phi_sq;
  if (int e = upload_state(c, s)) return e;
  CGState* st = c->d_state;
  double* part = c->d_partials;
  const float m = (float)prm->mass;
  auto iter = [&]() {
    if (HALF) {
      {
        KTimer kt(c, "hop_f16");
        hop_f16(c->stream, g, gauge->lp.u16, p16, pn, t16, tn, m, -1.f, nullptr, nullptr, st, nullptr);
      }
      {
        KTimer kt(c, "hop_dot_f16");
        hop_f16(c->stream, g, gauge->lp.u16, t16, tn, v16, vn, m, +1.f, p16, pn, st, part);
      }
      KTimer kt(c, "cg_blas");
      cg_xr_f16(c->stream, g, xl, r16, rn, p16, pn, v16, vn, st, part + kPartialSlot);
      cg_p_f16(c->stream, g, p16, pn, r16, rn, st);
    } else {
      {
        KTimer kt(c, "hop_f32");
        hop_f32(c->stream, g, gauge->lp.u32, p32, t32, m, -1.f, nullptr, st, nullptr);
      }
      {
        KTimer kt(c, "hop_dot_f32");
        hop_f32(c->stream, g, gauge->lp.u32, t32, v32, m, +1.f, p32, st, part);
      }
      KTimer kt(c, "cg_blas");
      cg_xr_f32(c->stream, g, xl, r32, p32, v32, st, part + kPartialSlot);
      cg_p_f32(c->stream, g, p32, r32, st);
    }
  };
  // The reliable update runs on the device at the end of every chunk, gated
  // on the state: x += x_lo, r = phi - DDdag x in FP64, and either stop
  // (true residual met, or budget spent) or restart the sloppy r and p.
  double2* t64 = buf<double2>(c, "cg_t", n);
  double2* ax = buf<double2>(c, "cg_v", n);
  if (!t64 || !ax) return SFH_E_NOMEM;
  double* part_ru = part + 3 * (size_t)kPartialSlot;
  auto tail = [&]() {
    KTimer kt(c, "reliable_update");
    ru_acc(c->stream, x, xl, n, st);
    hop_f64(c->stream, g, gauge->U, x, t64, prm->mass, -1.0, nullptr, st, nullptr, 1);
    hop_f64(c->stream, g, gauge->U, t64, ax, prm->mass, +1.0, nullptr, st, nullptr, 1);
    ru_sub(c->stream, g, r64, phi, ax, st, part_ru);
    if (HALF) ru_apply16(c->stream, g, r64, r16, rn, p16, pn, st);
    else ru_apply32(c->stream, g, r64, r32, p32, st);
  };
  auto direct = [&]() {
    for (int j = 0; j < kChunk; ++j) iter();
    tail();
  };
  const std::string key =
      HALF ? graph_key("f16", g, {gauge->U, gauge->lp.u16, x, xl, r64, phi, t64, ax, r16, p16, t16, v16, rn, pn, tn,
                                  vn, st, part},
                       prm->mass, get_hop_variant(2, g) * 2 + get_hop_variant(0, g))
           : graph_key("f32", g, {gauge->U, gauge->lp.u32, x, xl, r64, phi, t64, ax, r32, p32, t32, v32, st, part},
                       prm->mass, get_hop_variant(1, g) * 2 + get_hop_variant(0, g));
  CGState fin{};
  if (int e = run_chunks(c, [&]() { return launch_chunk(c, key, direct); }, fin)) return e;
  if (stats) {
    stats->iterations = fin.iter;
    stats->reliable_updates = fin.rupd;
    stats->rsq = fin.rsq;
    stats->true_rel_resid = std::sqrt(fin.rsq_true / phi_sq);
    const double per = HALF ? 528.0 : 912.0;
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
    if (rc == SFH_OK && prm->precision == SFH_PREC_F64 && prm->compute_true_residual && phi_sq >= 1e-28) {
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

}  // namespace sfh
