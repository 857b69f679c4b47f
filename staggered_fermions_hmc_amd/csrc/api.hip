// C ABI of libsfh (include/sfh.h): contexts, device-resident gauge and spinor
// fields, host-layout entry points and the device-resident leapfrog / HMC
// trajectory.  Host buffers are in the reference's AoS layouts and are
// converted on the device; every status is returned, never thrown.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "internal.h"

using namespace sfh;

namespace sfh {

int set_err(sfh_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_check(sfh_ctx* c, hipError_t e, const char* what) {
  if (e == hipSuccess) return SFH_OK;
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return set_err(c, e == hipErrorOutOfMemory ? SFH_E_NOMEM : SFH_E_HIP, m);
}

void drop_graphs(sfh_ctx* c) {
  if (c->graphs.empty()) return;
  hipStreamSynchronize(c->stream);
  for (auto& kv : c->graphs) {
    hipGraphExecDestroy(kv.second.second);
    hipGraphDestroy(kv.second.first);
  }
  c->graphs.clear();
}

void reset_aux_state(sfh_ctx* c) {
  CGState z{};
  z.ctr = c->d_ctr + kCtrKinds * kCtrPerKind;
  CGState* h = c->h_state;  // pinned staging; the copy completes before return
  *h = z;
  hipMemcpyAsync(&c->d_state[1], h, sizeof(CGState), hipMemcpyHostToDevice, c->stream);
  hipStreamSynchronize(c->stream);
}

void* scratch(sfh_ctx* c, const std::string& key, size_t bytes) {
  auto it = c->scratch.find(key);
  if (it != c->scratch.end() && it->second.second >= bytes) return it->second.first;
  if (it != c->scratch.end()) {
    drop_graphs(c);
    hipStreamSynchronize(c->stream);
    hipFree(it->second.first);
    c->scratch.erase(it);
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
    set_err(c, SFH_E_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes for " + key);
    return nullptr;
  }
  c->scratch[key] = {p, bytes};
  return p;
}

static hipEvent_t pool_get(sfh_ctx* c) {
  if (c->ev_pool.empty()) {
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
  }
  hipEvent_t e = c->ev_pool.back();
  c->ev_pool.pop_back();
  return e;
}

KTimer::KTimer(sfh_ctx* c_, const char* n) : c(c_), name(n) {
  if (!c->profiling) return;
  if (c->prof_every > 1 && (c->prof_seen[n]++ % c->prof_every) != 0) return;
  a = pool_get(c);
  hipEventRecord(a, c->stream);
}

KTimer::~KTimer() {
  if (!a) return;
  hipEvent_t b = pool_get(c);
  hipEventRecord(b, c->stream);
  c->pending.push_back({name, a, b});
  if (c->pending.size() > 4096) resolve_timers(c);
}

void resolve_timers(sfh_ctx* c) {
  if (c->pending.empty()) return;
  hipEventSynchronize(c->pending.back().b);
  for (auto& p : c->pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, p.a, p.b);
    auto& e = c->ktime[p.name];
    e.first += ms;
    e.second += 1;
    c->ev_pool.push_back(p.a);
    c->ev_pool.push_back(p.b);
  }
  c->pending.clear();
}

}  // namespace sfh

namespace {

struct Guard {
  sfh_ctx* c;
  std::lock_guard<std::recursive_mutex> lk;
  explicit Guard(sfh_ctx* c_) : c(c_), lk(c_->mu) { hipSetDevice(c->device); }
};

bool dims_ok(const int* d) {
  if (!d) return false;
  long v = 1;
  for (int k = 0; k < 4; ++k) {
    if (d[k] < 2 || (d[k] & 1)) return false;
    v *= d[k];
  }
  return v <= (1L << 24);  // keeps every reduction grid within one partials slot
}

double* staging(sfh_ctx* c, const char* key, size_t doubles) {
  return static_cast<double*>(scratch(c, key, doubles * sizeof(double)));
}

int upload_spinor(sfh_ctx* c, const Geom& g, const double* host, double2* dev) {
  double* st = staging(c, "stage_s", 6 * (size_t)g.V);
  if (!st) return SFH_E_NOMEM;
  hipMemcpyAsync(st, host, 6 * (size_t)g.V * sizeof(double), hipMemcpyHostToDevice, c->stream);
  spinor_aos_to_soa(c->stream, g, st, dev);
  return hip_check(c, hipGetLastError(), "spinor upload");
}

int download_spinor(sfh_ctx* c, const Geom& g, const double2* dev, double* host) {
  double* st = staging(c, "stage_s", 6 * (size_t)g.V);
  if (!st) return SFH_E_NOMEM;
  spinor_soa_to_aos(c->stream, g, dev, st);
  hipMemcpyAsync(host, st, 6 * (size_t)g.V * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  return hip_check(c, hipStreamSynchronize(c->stream), "spinor download");
}

int upload_links(sfh_ctx* c, const Geom& g, const double* host, double2* dev) {
  double* st = staging(c, "stage_l", 72 * (size_t)g.V);
  if (!st) return SFH_E_NOMEM;
  hipMemcpyAsync(st, host, 72 * (size_t)g.V * sizeof(double), hipMemcpyHostToDevice, c->stream);
  links_aos_to_soa(c->stream, g, st, dev);
  return hip_check(c, hipGetLastError(), "link upload");
}

int download_links(sfh_ctx* c, const Geom& g, const double2* dev, double* host) {
  double* st = staging(c, "stage_l", 72 * (size_t)g.V);
  if (!st) return SFH_E_NOMEM;
  links_soa_to_aos(c->stream, g, dev, st);
  hipMemcpyAsync(host, st, 72 * (size_t)g.V * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  return hip_check(c, hipStreamSynchronize(c->stream), "link download");
}

double2* dbuf(sfh_ctx* c, const char* key, size_t n) {
  return static_cast<double2*>(scratch(c, key, n * sizeof(double2)));
}

void invalidate_lp(sfh_gauge* g) {
  g->lp32_valid = false;
  g->lp16_valid = false;
  g->su3_known = false;
}

// One force evaluation and momentum kick P -= coef F on the device; X (may be
// unused) receives the solve.
int md_kick(sfh_ctx* c, sfh_gauge* gauge, const sfh_md_params* p, const double2* phi, double2* P, double coef,
            double2* X) {
  const Geom& g = gauge->g;
  double2* Y = nullptr;
  if (p->which & SFH_FORCE_FERMION) {
    sfh_cg_params cp{};
    cp.mass = p->mass;
    cp.tol = p->tol;
    cp.max_iter = p->max_iter;
    cp.precision = p->precision;
    if (int e = cg_solve(c, gauge, &cp, phi, X, nullptr)) return e;
    Y = dbuf(c, "md_Y", 3 * (size_t)g.V);
    if (!Y) return SFH_E_NOMEM;
    if (int e = ensure_su3(c, gauge)) return e;
    KTimer kt(c, "hop_f64");
    hop_f64(c->stream, g, gauge->U, X, Y, p->mass, -1.0, nullptr, nullptr, nullptr, 0, gauge->f64_su3);
  }
  if (int e = ensure_su3(c, gauge)) return e;
  return force_field(c, g, gauge->U, X, Y, p->beta, p->c11, p->c12, p->which, nullptr, P, coef, gauge->f64_su3);
}

int md_leapfrog(sfh_ctx* c, sfh_gauge* gauge, const sfh_md_params* p, const double2* phi, double2* P,
                double2* X) {
  const Geom& g = gauge->g;
  if (int e = md_kick(c, gauge, p, phi, P, p->eps * 0.5, X)) return e;
  for (int n = 1; n < p->nsteps; ++n) {
    if (int e = link_update(c, g, gauge->U, P, p->eps, p->check_hermitian)) return e;
    invalidate_lp(gauge);
    if (int e = md_kick(c, gauge, p, phi, P, p->eps, X)) return e;
  }
  if (int e = link_update(c, g, gauge->U, P, p->eps, p->check_hermitian)) return e;
  invalidate_lp(gauge);
  return md_kick(c, gauge, p, phi, P, p->eps * 0.5, X);
}

bool md_ok(const sfh_md_params* p) {
  return p && p->nsteps >= 1 && std::isfinite(p->eps) && (p->which & 3u) != 0 && (p->which & ~3u) == 0;
}

}  // namespace

extern "C" {

int sfh_abi_version(void) { return SFH_ABI_VERSION; }

const char* sfh_status_string(int s) {
  switch (s) {
    case SFH_OK: return "ok";
    case SFH_E_INVALID: return "invalid argument";
    case SFH_E_NOMEM: return "out of device memory";
    case SFH_E_HIP: return "HIP runtime error";
    case SFH_E_NODEVICE: return "no HIP device";
    case SFH_E_BREAKDOWN: return "CG breakdown";
    case SFH_E_NONFINITE_ALPHA: return "CG non-finite alpha";
    case SFH_E_NONFINITE_BETA: return "CG non-finite beta";
    case SFH_E_NOCONV: return "CG did not converge";
    case SFH_E_NOT_HERMITIAN: return "momentum not hermitian traceless";
    case SFH_E_EXP_QCUBED: return "exp: Im Tr Q^3 non-zero";
    case SFH_E_EXP_NEG_TRACE: return "exp: Tr Q^2 negative";
    case SFH_E_SAMPLER: return "heatbath sampler found no sample in 2^20 tries";
    case SFH_E_IO: return "file could not be opened, read or written completely";
    default: return "unknown status";
  }
}

int sfh_create(int device, sfh_ctx** out) {
  if (!out) return SFH_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SFH_E_NODEVICE;
  if (device < 0 || device >= n) return SFH_E_INVALID;
  if (hipSetDevice(device) != hipSuccess) return SFH_E_HIP;
  sfh_ctx* c = new sfh_ctx();
  c->device = device;
  bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipMalloc(&c->d_state, 4 * sizeof(CGState)) == hipSuccess;
  ok = ok && hipMalloc(&c->d_partials, 8 * (size_t)kPartialSlot * sizeof(double)) == hipSuccess;
  ok = ok && hipHostMalloc(&c->h_state, 3 * sizeof(CGState), hipHostMallocDefault) == hipSuccess;
  ok = ok && hipHostMalloc(&c->h_scalar, 8 * sizeof(double), hipHostMallocDefault) == hipSuccess;
  ok = ok && hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&c->chunk_ev[0], hipEventDisableTiming) == hipSuccess &&
       hipEventCreateWithFlags(&c->chunk_ev[1], hipEventDisableTiming) == hipSuccess;
  ok = ok && hipMalloc(&c->d_ctr, 2 * kCtrBytes) == hipSuccess;
  if (ok) {
    CGState z[4] = {};
    z[0].ctr = c->d_ctr;
    z[1].ctr = c->d_ctr + kCtrKinds * kCtrPerKind;
    z[2].ctr = c->d_ctr;  // second copy of the mixed solve's state (shares state 0's counters)
    z[3].ctr = z[1].ctr;
    ok = hipMemsetAsync(c->d_ctr, 0, 2 * kCtrBytes, c->stream) == hipSuccess &&
         hipMemcpyAsync(c->d_state, z, sizeof z, hipMemcpyHostToDevice, c->stream) == hipSuccess &&
         hipStreamSynchronize(c->stream) == hipSuccess;
  }
  // device address of the allocation-free entry points' work area (once per device)
  ok = ok && noalloc_init(device) == SFH_OK;
  if (!ok) {
    sfh_destroy(c);
    return SFH_E_HIP;
  }
  *out = c;
  return SFH_OK;
}

int sfh_destroy(sfh_ctx* c) {
  if (!c) return SFH_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  resolve_timers(c);
  for (auto e : c->ev_pool) hipEventDestroy(e);
  for (auto& kv : c->graphs) {
    hipGraphExecDestroy(kv.second.second);
    hipGraphDestroy(kv.second.first);
  }
  for (auto& kv : c->scratch) hipFree(kv.second.first);
  if (c->d_state) hipFree(c->d_state);
  if (c->d_partials) hipFree(c->d_partials);
  if (c->d_ctr) hipFree(c->d_ctr);
  if (c->h_state) hipHostFree(c->h_state);
  if (c->h_scalar) hipHostFree(c->h_scalar);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  for (hipEvent_t e : c->chunk_ev)
    if (e) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return SFH_OK;
}

const char* sfh_last_error(const sfh_ctx* c) { return c ? c->err.c_str() : "no context"; }

int sfh_synchronize(sfh_ctx* c) {
  if (!c) return SFH_E_INVALID;
  Guard gd(c);
  return hip_check(c, hipStreamSynchronize(c->stream), "synchronize");
}

int sfh_set_profiling(sfh_ctx* c, int on) {
  if (!c) return SFH_E_INVALID;
  Guard gd(c);
  resolve_timers(c);
  c->profiling = on != 0;
  c->prof_every = on > 1 ? on : 1;
  c->prof_seen.clear();
  if (on) c->ktime.clear();
  return SFH_OK;
}

int sfh_set_graph_replay(sfh_ctx* c, int on) {
  if (!c) return SFH_E_INVALID;
  Guard gd(c);
  c->graph_replay = on != 0;
  return SFH_OK;
}

int sfh_kernel_time(sfh_ctx* c, const char* name, double* total_ms, long* launches) {
  if (!c || !name) return SFH_E_INVALID;
  Guard gd(c);
  resolve_timers(c);
  auto it = c->ktime.find(name);
  if (total_ms) *total_ms = it == c->ktime.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == c->ktime.end() ? 0 : it->second.second;
  return SFH_OK;
}

int sfh_set_hop_variant(int precision, int variant) {
  if (precision < 0 || precision > 2 || variant < -1 || variant > 3) return SFH_E_INVALID;
  set_hop_variant(precision, variant);
  return SFH_OK;
}

int sfh_tune_hop(sfh_ctx* c, const sfh_gauge* g, int precision, int n, int* variant, double* us) {
  if (!c || !g || precision < 0 || precision > 2 || n < 1) return SFH_E_INVALID;
  Guard gd(c);
  const Geom& gg = g->g;
  const size_t nc = 3 * (size_t)gg.V;
  void* a = scratch(c, "tune_a", nc * sizeof(double2));
  void* b = scratch(c, "tune_b", nc * sizeof(double2));
  if (!a || !b) return SFH_E_NOMEM;
  if (int e = ensure_lp(c, const_cast<sfh_gauge*>(g), precision)) return e;
  hipMemsetAsync(a, 0, nc * sizeof(double2), c->stream);
  double best_us = 1e30;
  int best = 0;
  for (int v = 0; v < 4; ++v) {
    set_hop_variant(precision, v);
    auto launch = [&]() {
      if (precision == 0)
        hop_f64(c->stream, gg, g->U, (const double2*)a, (double2*)b, 0.1, 1.0, nullptr, nullptr, nullptr, 0,
                g->f64_su3);
      else if (precision == 1)
        hop_f32(c->stream, gg, g->lp.u32, (const float2*)a, (float2*)b, 0.1f, 1.0f, nullptr, nullptr, nullptr,
                g->lp_su3);
      else
        hop_f16(c->stream, gg, g->lp.u16, (const H16*)a, (H16*)b, 0.1f, 1.0f, nullptr, nullptr, nullptr, g->lp_su3);
    };
    launch();
    hipEventRecord(c->ev0, c->stream);
    for (int k = 0; k < n; ++k) launch();
    hipEventRecord(c->ev1, c->stream);
    if (int e = hip_check(c, hipEventSynchronize(c->ev1), "hop tuning")) return e;
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    const double t = 1e3 * ms / n;
    if (t < best_us) {
      best_us = t;
      best = v;
    }
  }
  set_hop_variant(precision, best);
  if (variant) *variant = best;
  if (us) *us = best_us;
  return SFH_OK;
}

int sfh_time_hop(sfh_ctx* c, const sfh_gauge* g, int precision, int with_dot, int n, double* us) {
  if (!c || !g || precision < 0 || precision > 2 || n < 1 || !us) return SFH_E_INVALID;
  Guard gd(c);
  const Geom& gg = g->g;
  const size_t nc = 3 * (size_t)gg.V;
  void* a = scratch(c, "tune_a", nc * sizeof(double2));
  void* b = scratch(c, "tune_b", nc * sizeof(double2));
  if (!a || !b) return SFH_E_NOMEM;
  if (int e = ensure_lp(c, const_cast<sfh_gauge*>(g), precision)) return e;
  // input of ones (a non-zero <p,v> keeps the timed solve state running)
  if (precision == 0) {
    std::vector<double> ones(2 * nc, 1.0);
    hipMemcpyAsync(a, ones.data(), ones.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);
  } else if (precision == 1) {
    std::vector<float> ones(2 * nc, 1.0f);
    hipMemcpyAsync(a, ones.data(), ones.size() * sizeof(float), hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);
  } else {
    h16_ones(c->stream, static_cast<H16*>(a), gg.V);
  }
  // a live solve state so the reduction runs exactly as in the CG
  CGState* st = &c->d_state[1];
  CGState* z = c->h_state;  // pinned staging
  *z = CGState{};
  z->rsq = 1.0;
  z->ctr = c->d_ctr + kCtrKinds * kCtrPerKind;
  hipMemcpyAsync(st, z, sizeof(CGState), hipMemcpyHostToDevice, c->stream);
  if (int e = hip_check(c, hipStreamSynchronize(c->stream), "hop timing state")) return e;
  double* part = c->d_partials + 4 * (size_t)kPartialSlot;
  auto launch = [&]() {
    if (precision == 0)
      hop_f64(c->stream, gg, g->U, (const double2*)a, (double2*)b, 0.1, 1.0, with_dot ? (const double2*)a : nullptr,
              with_dot ? st : nullptr, part, 0, g->f64_su3);
    else if (precision == 1)
      hop_f32(c->stream, gg, g->lp.u32, (const float2*)a, (float2*)b, 0.1f, 1.0f,
              with_dot ? (const float2*)a : nullptr, with_dot ? st : nullptr, part, g->lp_su3);
    else
      hop_f16(c->stream, gg, g->lp.u16, (const H16*)a, (H16*)b, 0.1f, 1.0f, with_dot ? (const H16*)a : nullptr,
              with_dot ? st : nullptr, part, g->lp_su3);
  };
  launch();
  hipEventRecord(c->ev0, c->stream);
  for (int k = 0; k < n; ++k) launch();
  hipEventRecord(c->ev1, c->stream);
  if (int e = hip_check(c, hipEventSynchronize(c->ev1), "hop timing")) return e;
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  *us = 1e3 * ms / n;
  reset_aux_state(c);
  return SFH_OK;
}

int sfh_time_force(sfh_ctx* c, const sfh_gauge* g, double beta, double c11, double c12, unsigned which, int n,
                   double* us) {
  if (!c || !g || n < 1 || !us || !which || (which & ~3u)) return SFH_E_INVALID;
  Guard gd(c);
  const Geom& gg = g->g;
  double2* X = dbuf(c, "tf_x", 3 * (size_t)gg.V);
  double2* Y = dbuf(c, "tf_y", 3 * (size_t)gg.V);
  double2* P = dbuf(c, "tf_p", 36 * (size_t)gg.V);
  if (!X || !Y || !P) return SFH_E_NOMEM;
  hipMemsetAsync(X, 0, 3 * (size_t)gg.V * sizeof(double2), c->stream);
  hipMemsetAsync(Y, 0, 3 * (size_t)gg.V * sizeof(double2), c->stream);
  hipMemsetAsync(P, 0, 36 * (size_t)gg.V * sizeof(double2), c->stream);
  // as the leapfrog runs it: no force field stored, P -= coef F fused
  if (int e = ensure_su3(c, const_cast<sfh_gauge*>(g))) return e;
  auto launch = [&]() {
    return force_field(c, gg, g->U, X, Y, beta, c11, c12, which, nullptr, P, 1e-3, g->f64_su3);
  };
  if (int e = launch()) return e;
  hipEventRecord(c->ev0, c->stream);
  for (int k = 0; k < n; ++k)
    if (int e = launch()) return e;
  hipEventRecord(c->ev1, c->stream);
  if (int e = hip_check(c, hipEventSynchronize(c->ev1), "force timing")) return e;
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  *us = 1e3 * ms / n;
  return SFH_OK;
}

int sfh_time_cg_kernels(sfh_ctx* c, const sfh_gauge* g, int precision, int n, double* us_a, double* us_b) {
  if (!c || !g || (precision != SFH_PREC_F32 && precision != SFH_PREC_F16) || n < 1 || !us_a || !us_b)
    return SFH_E_INVALID;
  Guard gd(c);
  return time_mixed(c, const_cast<sfh_gauge*>(g), precision, n, us_a, us_b);
}

int sfh_set_cg_form(int precision, int form) {
  if ((precision != SFH_PREC_F32 && precision != SFH_PREC_F16) || form < -1 || form > 1) return SFH_E_INVALID;
  set_mix_form(precision, form);
  return SFH_OK;
}

int sfh_tune_cg_form(sfh_ctx* c, const sfh_gauge* g, int precision, int n, int* form, double* us) {
  if (!c || !g || (precision != SFH_PREC_F32 && precision != SFH_PREC_F16) || n < 1) return SFH_E_INVALID;
  Guard gd(c);
  double best_us = 1e30;
  int best = 0;
  for (int f = 0; f < 2; ++f) {
    set_mix_form(precision, f);
    double ua = 0.0, ub = 0.0;
    if (int e = time_mixed(c, const_cast<sfh_gauge*>(g), precision, n, &ua, &ub)) {
      set_mix_form(precision, -1);
      return e;
    }
    if (ua + ub < best_us) {
      best_us = ua + ub;
      best = f;
    }
  }
  set_mix_form(precision, best);
  if (form) *form = best;
  if (us) *us = best_us;
  return SFH_OK;
}

// ---------------------------------------------------------------- gauge
int sfh_gauge_create(sfh_ctx* c, const int dims[4], sfh_gauge** out) {
  if (!c || !out) return SFH_E_INVALID;
  Guard gd(c);
  if (!dims_ok(dims)) return set_err(c, SFH_E_INVALID, "lattice extents must be even and >= 2");
  sfh_gauge* g = new sfh_gauge();
  g->g = make_geom(dims);
  if (hipMalloc(&g->U, 36 * (size_t)g->g.V * sizeof(double2)) != hipSuccess) {
    delete g;
    return set_err(c, SFH_E_NOMEM, "gauge field allocation");
  }
  *out = g;
  return SFH_OK;
}

int sfh_gauge_destroy(sfh_ctx* c, sfh_gauge* g) {
  if (!c || !g) return SFH_E_INVALID;
  Guard gd(c);
  drop_graphs(c);
  hipStreamSynchronize(c->stream);
  hipFree(g->U);
  if (g->lp.u32) hipFree(g->lp.u32);
  if (g->lp.u16) hipFree(g->lp.u16);
  delete g;
  return SFH_OK;
}

int sfh_gauge_upload(sfh_ctx* c, sfh_gauge* g, const double* links) {
  if (!c || !g || !links) return SFH_E_INVALID;
  Guard gd(c);
  invalidate_lp(g);
  if (int e = upload_links(c, g->g, links, g->U)) return e;
  return hip_check(c, hipStreamSynchronize(c->stream), "gauge upload");
}

int sfh_gauge_download(sfh_ctx* c, const sfh_gauge* g, double* links) {
  if (!c || !g || !links) return SFH_E_INVALID;
  Guard gd(c);
  return download_links(c, g->g, g->U, links);
}

int sfh_gauge_save(sfh_ctx* c, const sfh_gauge* const* gauges, int n, const char* path) {
  if (!c || !gauges || n < 1 || !path) return SFH_E_INVALID;
  Guard gd(c);
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_err(c, SFH_E_INVALID, std::string("Error opening file for writing: ") + path);
  int rc = SFH_OK;
  std::vector<double> host;
  for (int k = 0; k < n && rc == SFH_OK; ++k) {
    if (!gauges[k]) {
      rc = SFH_E_INVALID;
      break;
    }
    host.resize((size_t)gauges[k]->g.V * 72);
    if ((rc = download_links(c, gauges[k]->g, gauges[k]->U, host.data()))) break;
    if (std::fwrite(host.data(), sizeof(double), host.size(), f) != host.size())
      rc = set_err(c, SFH_E_INVALID, std::string("Error writing complete binary data to: ") + path);
  }
  std::fclose(f);
  return rc;
}

int sfh_gauge_load(sfh_ctx* c, sfh_gauge* const* gauges, int n, const char* path) {
  if (!c || !gauges || n < 1 || !path) return SFH_E_INVALID;
  Guard gd(c);
  FILE* f = std::fopen(path, "rb");
  if (!f) return set_err(c, SFH_E_INVALID, std::string("Error opening file for reading: ") + path);
  int rc = SFH_OK;
  std::vector<double> host;
  for (int k = 0; k < n && rc == SFH_OK; ++k) {
    if (!gauges[k]) {
      rc = SFH_E_INVALID;
      break;
    }
    host.resize((size_t)gauges[k]->g.V * 72);
    if (std::fread(host.data(), sizeof(double), host.size(), f) != host.size()) {
      rc = set_err(c, SFH_E_INVALID, std::string("Error reading complete binary data from: ") + path);
      break;
    }
    invalidate_lp(gauges[k]);
    if ((rc = upload_links(c, gauges[k]->g, host.data(), gauges[k]->U))) break;
    rc = hip_check(c, hipStreamSynchronize(c->stream), "gauge load");
  }
  std::fclose(f);
  return rc;
}

int sfh_random_links(const int dims[4], double eps, unsigned long long seed, double* links) {
  if (!dims_ok(dims) || !links || !std::isfinite(eps)) return SFH_E_INVALID;
  gen_links(dims, eps, seed, links);
  return SFH_OK;
}

int sfh_random_spinor(const int dims[4], double sigma, unsigned long long seed, double* out) {
  if (!dims_ok(dims) || !out) return SFH_E_INVALID;
  gen_spinor(dims, sigma, seed, out);
  return SFH_OK;
}

int sfh_random_momenta(const int dims[4], unsigned long long seed, double* out) {
  if (!dims_ok(dims) || !out) return SFH_E_INVALID;
  gen_momenta(dims, seed, out);
  return SFH_OK;
}

// ---------------------------------------------------------------- fields
int sfh_field_create(sfh_ctx* c, const sfh_gauge* g, sfh_field** out) {
  if (!c || !g || !out) return SFH_E_INVALID;
  Guard gd(c);
  sfh_field* f = new sfh_field();
  f->g = g->g;
  if (hipMalloc(&f->d, 3 * (size_t)f->g.V * sizeof(double2)) != hipSuccess) {
    delete f;
    return set_err(c, SFH_E_NOMEM, "field allocation");
  }
  hipMemsetAsync(f->d, 0, 3 * (size_t)f->g.V * sizeof(double2), c->stream);
  *out = f;
  return hip_check(c, hipStreamSynchronize(c->stream), "field create");
}

int sfh_field_destroy(sfh_ctx* c, sfh_field* f) {
  if (!c || !f) return SFH_E_INVALID;
  Guard gd(c);
  drop_graphs(c);
  hipStreamSynchronize(c->stream);
  hipFree(f->d);
  delete f;
  return SFH_OK;
}

int sfh_field_upload(sfh_ctx* c, sfh_field* f, const double* host) {
  if (!c || !f || !host) return SFH_E_INVALID;
  Guard gd(c);
  if (int e = upload_spinor(c, f->g, host, f->d)) return e;
  return hip_check(c, hipStreamSynchronize(c->stream), "field upload");
}

int sfh_field_download(sfh_ctx* c, const sfh_field* f, double* host) {
  if (!c || !f || !host) return SFH_E_INVALID;
  Guard gd(c);
  return download_spinor(c, f->g, f->d, host);
}

// ---------------------------------------------------------------- fermion ops
int sfh_apply_d(sfh_ctx* c, const sfh_gauge* g, double mass, int dagger, const double* in, double* out) {
  if (!c || !g || !in || !out) return SFH_E_INVALID;
  Guard gd(c);
  const size_t n = 3 * (size_t)g->g.V;
  double2 *a = dbuf(c, "io_a", n), *b = dbuf(c, "io_b", n);
  if (!a || !b) return SFH_E_NOMEM;
  if (int e = upload_spinor(c, g->g, in, a)) return e;
  if (int e = ensure_su3(c, const_cast<sfh_gauge*>(g))) return e;
  {
    KTimer kt(c, "hop_f64");
    hop_f64(c->stream, g->g, g->U, a, b, mass, dagger ? -1.0 : 1.0, nullptr, nullptr, nullptr, 0, g->f64_su3);
  }
  return download_spinor(c, g->g, b, out);
}

int sfh_apply_d_prec(sfh_ctx* c, const sfh_gauge* g, double mass, int dagger, int precision, const double* in,
                     double* out) {
  if (!c || !g || !in || !out) return SFH_E_INVALID;
  if (precision == SFH_PREC_F64) return sfh_apply_d(c, g, mass, dagger, in, out);
  if (precision != SFH_PREC_F32 && precision != SFH_PREC_F16) return set_err(c, SFH_E_INVALID, "unknown precision");
  Guard gd(c);
  const Geom& gg = g->g;
  const size_t n = 3 * (size_t)gg.V;
  double2 *a = dbuf(c, "io_a", n), *b = dbuf(c, "io_b", n);
  if (!a || !b) return SFH_E_NOMEM;
  sfh_gauge* gm = const_cast<sfh_gauge*>(g);
  if (int e = upload_spinor(c, gg, in, a)) return e;
  if (int e = ensure_lp(c, gm, precision)) return e;
  const float fb = dagger ? -1.0f : 1.0f;
  if (precision == SFH_PREC_F32) {
    float2* s0 = static_cast<float2*>(scratch(c, "lp_io_a", n * sizeof(float2)));
    float2* s1 = static_cast<float2*>(scratch(c, "lp_io_b", n * sizeof(float2)));
    if (!s0 || !s1) return SFH_E_NOMEM;
    f64_to_f32(c->stream, a, s0, n);
    KTimer kt(c, "hop_f32");
    hop_f32(c->stream, gg, gm->lp.u32, s0, s1, (float)mass, fb, nullptr, nullptr, nullptr, gm->lp_su3);
    f32_to_f64(c->stream, s1, b, n);
  } else {
    H16* s0 = static_cast<H16*>(scratch(c, "lp_io_a16", (size_t)gg.V * sizeof(H16)));
    H16* s1 = static_cast<H16*>(scratch(c, "lp_io_b16", (size_t)gg.V * sizeof(H16)));
    if (!s0 || !s1) return SFH_E_NOMEM;
    f64_to_f16(c->stream, gg, a, s0);
    KTimer kt(c, "hop_f16");
    hop_f16(c->stream, gg, gm->lp.u16, s0, s1, (float)mass, fb, nullptr, nullptr, nullptr, gm->lp_su3);
    f16_to_f64(c->stream, gg, s1, b);
  }
  if (int e = hip_check(c, hipGetLastError(), "low-precision hop")) return e;
  return download_spinor(c, gg, b, out);
}

int sfh_apply_ddag(sfh_ctx* c, const sfh_gauge* g, double mass, const double* in, double* out) {
  if (!c || !g || !in || !out) return SFH_E_INVALID;
  Guard gd(c);
  const size_t n = 3 * (size_t)g->g.V;
  double2 *a = dbuf(c, "io_a", n), *b = dbuf(c, "io_b", n);
  if (!a || !b) return SFH_E_NOMEM;
  if (int e = upload_spinor(c, g->g, in, a)) return e;
  if (int e = apply_ddag_dev(c, g, mass, a, b)) return e;
  return download_spinor(c, g->g, b, out);
}

int sfh_apply_d_field(sfh_ctx* c, const sfh_gauge* g, double mass, int dagger, const sfh_field* in,
                      sfh_field* out) {
  if (!c || !g || !in || !out || in == out) return SFH_E_INVALID;
  Guard gd(c);
  if (in->g.V != g->g.V || out->g.V != g->g.V) return set_err(c, SFH_E_INVALID, "field / gauge size mismatch");
  if (int e = ensure_su3(c, const_cast<sfh_gauge*>(g))) return e;
  {
    KTimer kt(c, "hop_f64");
    hop_f64(c->stream, g->g, g->U, in->d, out->d, mass, dagger ? -1.0 : 1.0, nullptr, nullptr, nullptr, 0,
            g->f64_su3);
  }
  return hip_check(c, hipStreamSynchronize(c->stream), "apply_d_field");
}

int sfh_inner(sfh_ctx* c, size_t n, const double* a, const double* b, double* out) {
  if (!c || !a || !b || !out) return SFH_E_INVALID;
  Guard gd(c);
  if (n == 0) {
    *out = 0.0;
    return SFH_OK;
  }
  double* da = staging(c, "inner_a", n);
  double* db = staging(c, "inner_b", n);
  if (!da || !db) return SFH_E_NOMEM;
  hipMemcpyAsync(da, a, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(db, b, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
  dot_f64(c->stream, da, db, n, c->d_partials + 2 * kPartialSlot, c->aux_ctr(),
          &c->d_state[1].alpha);
  hipMemcpyAsync(c->h_scalar, &c->d_state[1].alpha, sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (int e = hip_check(c, hipStreamSynchronize(c->stream), "inner")) return e;
  *out = *c->h_scalar;
  return SFH_OK;
}

static bool cg_params_ok(const sfh_cg_params* p) {
  return p && std::isfinite(p->mass) && std::isfinite(p->tol) && p->tol >= 0 && p->precision >= 0 &&
         p->precision <= 2;
}

int sfh_cg(sfh_ctx* c, const sfh_gauge* g, const sfh_cg_params* p, const double* phi, double* x,
           sfh_cg_stats* stats) {
  if (!c || !g || !phi || !x || !cg_params_ok(p)) return SFH_E_INVALID;
  Guard gd(c);
  const size_t n = 3 * (size_t)g->g.V;
  double2 *a = dbuf(c, "io_a", n), *b = dbuf(c, "io_b", n);
  if (!a || !b) return SFH_E_NOMEM;
  if (int e = upload_spinor(c, g->g, phi, a)) return e;
  const int rc = cg_solve(c, const_cast<sfh_gauge*>(g), p, a, b, stats);
  if (rc != SFH_OK) return rc;
  return download_spinor(c, g->g, b, x);
}

int sfh_cg_field(sfh_ctx* c, const sfh_gauge* g, const sfh_cg_params* p, const sfh_field* phi, sfh_field* x,
                 sfh_cg_stats* stats) {
  if (!c || !g || !phi || !x || phi == x || !cg_params_ok(p)) return SFH_E_INVALID;
  Guard gd(c);
  if (phi->g.V != g->g.V || x->g.V != g->g.V) return set_err(c, SFH_E_INVALID, "field / gauge size mismatch");
  return cg_solve(c, const_cast<sfh_gauge*>(g), p, phi->d, x->d, stats);
}

namespace {

// Validates the caller's batch and lowers it to device pointers.
int lower_batch(sfh_ctx* c, int nsys, const sfh_batch_system* systems, std::vector<BatchSys>& bs,
                std::vector<std::vector<const double2*>>& rhs, std::vector<std::vector<double2*>>& xs) {
  bs.resize(nsys);
  rhs.assign(nsys, {});
  xs.assign(nsys, {});
  for (int k = 0; k < nsys; ++k) {
    const sfh_batch_system& s = systems[k];
    if (!s.gauge || s.nrhs < 1 || s.nrhs > 8 || !s.rhs || !s.x || !std::isfinite(s.mass))
      return set_err(c, SFH_E_INVALID, "batched system " + std::to_string(k) + ": bad gauge, mass or nrhs");
    for (int j = 0; j < s.nrhs; ++j) {
      if (!s.rhs[j] || !s.x[j] || s.rhs[j] == s.x[j] || s.rhs[j]->g.V != s.gauge->g.V ||
          s.x[j]->g.V != s.gauge->g.V)
        return set_err(c, SFH_E_INVALID, "batched system " + std::to_string(k) + ": bad field");
      rhs[k].push_back(s.rhs[j]->d);
      xs[k].push_back(s.x[j]->d);
    }
    bs[k] = BatchSys{s.gauge, s.mass, s.nrhs, rhs[k].data(), xs[k].data()};
  }
  return SFH_OK;
}

}  // namespace

int sfh_cg_batch(sfh_ctx* c, int nsys, const sfh_batch_system* systems, const sfh_cg_params* p,
                 sfh_cg_stats* stats) {
  if (!c || nsys < 1 || !systems || !cg_params_ok(p)) return SFH_E_INVALID;
  Guard gd(c);
  std::vector<BatchSys> bs;
  std::vector<std::vector<const double2*>> rhs;
  std::vector<std::vector<double2*>> xs;
  if (int e = lower_batch(c, nsys, systems, bs, rhs, xs)) return e;
  return cg_batch(c, nsys, bs.data(), p, stats);
}

int sfh_tune_batch(sfh_ctx* c, int nsys, const sfh_batch_system* systems, const sfh_cg_params* p, int* hop_threads,
                   int* rhs_tile, double* us) {
  if (!c || nsys < 1 || !systems || !cg_params_ok(p)) return SFH_E_INVALID;
  Guard gd(c);
  std::vector<BatchSys> bs;
  std::vector<std::vector<const double2*>> rhs;
  std::vector<std::vector<double2*>> xs;
  if (int e = lower_batch(c, nsys, systems, bs, rhs, xs)) return e;
  return tune_batch(c, nsys, bs.data(), p, hop_threads, rhs_tile, us);
}

// ---------------------------------------------------------------- forces, MD
int sfh_force(sfh_ctx* c, const sfh_gauge* g, double beta, double c11, double c12, const double* X,
              const double* Y, unsigned which, double* F) {
  if (!c || !g || !F || (which & ~3u) || !which) return SFH_E_INVALID;
  if ((which & SFH_FORCE_FERMION) && (!X || !Y)) return SFH_E_INVALID;
  Guard gd(c);
  const Geom& gg = g->g;
  const size_t n = 3 * (size_t)gg.V;
  double2 *dx = nullptr, *dy = nullptr;
  if (which & SFH_FORCE_FERMION) {
    dx = dbuf(c, "io_a", n);
    dy = dbuf(c, "io_b", n);
    if (!dx || !dy) return SFH_E_NOMEM;
    if (int e = upload_spinor(c, gg, X, dx)) return e;
    if (int e = upload_spinor(c, gg, Y, dy)) return e;
  }
  double2* dF = dbuf(c, "force", 36 * (size_t)gg.V);
  if (!dF) return SFH_E_NOMEM;
  if (int e = ensure_su3(c, const_cast<sfh_gauge*>(g))) return e;
  if (int e = force_field(c, gg, g->U, dx, dy, beta, c11, c12, which, dF, nullptr, 0.0, g->f64_su3)) return e;
  return download_links(c, gg, dF, F);
}

int sfh_action(sfh_ctx* c, const sfh_gauge* g, double beta, double c11, double c12, double* out) {
  if (!c || !g || !out) return SFH_E_INVALID;
  Guard gd(c);
  if (!staple_dims_ok(g->g)) return set_err(c, SFH_E_INVALID, "action: lattice extents above 255");
  if (int e = ensure_su3(c, const_cast<sfh_gauge*>(g))) return e;
  *out = gauge_action(c, g->g, g->U, beta, c11, c12, g->f64_su3);
  return std::isnan(*out) ? SFH_E_HIP : SFH_OK;
}

int sfh_avg_plaq(sfh_ctx* c, const sfh_gauge* g, double* out) {
  if (!c || !g || !out) return SFH_E_INVALID;
  Guard gd(c);
  if (!staple_dims_ok(g->g)) return set_err(c, SFH_E_INVALID, "plaquette: lattice extents above 255");
  if (int e = ensure_su3(c, const_cast<sfh_gauge*>(g))) return e;
  *out = avg_plaq(c, g->g, g->U, g->f64_su3);
  return std::isnan(*out) ? SFH_E_HIP : SFH_OK;
}

int sfh_polyakov_grid(sfh_ctx* c, const sfh_gauge* g, double* out) {
  if (!c || !g || !out) return SFH_E_INVALID;
  Guard gd(c);
  return polyakov_grid(c, g->g, g->U, out);
}

int sfh_polyakov_correlators(sfh_ctx* c, const sfh_gauge* g, int r2_first, int count, double* out) {
  if (!c || !g || !out || r2_first < 1 || count < 1 || count > 65535) return SFH_E_INVALID;
  Guard gd(c);
  return polyakov_correlators(c, g->g, g->U, r2_first, count, out);
}

int sfh_polyakov_correlators_grid(sfh_ctx* c, const int spatial[3], const double* grid, int r2_first, int count,
                                  double* out) {
  if (!c || !spatial || !grid || !out || r2_first < 1 || count < 1 || count > 65535) return SFH_E_INVALID;
  long ns = 1;
  for (int k = 0; k < 3; ++k) {
    if (spatial[k] < 1) return SFH_E_INVALID;
    ns *= spatial[k];
  }
  if (ns > (1L << 24)) return SFH_E_INVALID;
  Guard gd(c);
  return polyakov_correlators_grid(c, spatial, grid, r2_first, count, out);
}

int sfh_links_save(const double* const* arrays, int n, size_t doubles_each, const char* path) {
  if (!arrays || n < 1 || !path) return SFH_E_INVALID;
  FILE* f = std::fopen(path, "wb");
  if (!f) return SFH_E_IO;
  int rc = SFH_OK;
  for (int k = 0; k < n && rc == SFH_OK; ++k)
    if (!arrays[k] || std::fwrite(arrays[k], sizeof(double), doubles_each, f) != doubles_each) rc = SFH_E_IO;
  if (std::fclose(f) != 0) rc = SFH_E_IO;
  return rc;
}

int sfh_links_load(double* const* arrays, int n, size_t doubles_each, const char* path) {
  if (!arrays || n < 1 || !path) return SFH_E_INVALID;
  FILE* f = std::fopen(path, "rb");
  if (!f) return SFH_E_IO;
  int rc = SFH_OK;
  for (int k = 0; k < n && rc == SFH_OK; ++k)
    if (!arrays[k] || std::fread(arrays[k], sizeof(double), doubles_each, f) != doubles_each) rc = SFH_E_IO;
  std::fclose(f);
  return rc;
}

int sfh_heatbath(sfh_ctx* c, sfh_gauge* g, double beta, double c11, double c12, unsigned long long seed,
                 unsigned sweep0, int nsweeps, double* avg_plaq, double* action_density) {
  if (!c || !g || nsweeps < 0 || !(beta > 0) || !std::isfinite(beta) || !std::isfinite(c11) ||
      !std::isfinite(c12))
    return SFH_E_INVALID;
  if (!heatbath_dims_ok(g->g))
    return set_err(c, SFH_E_INVALID, "heatbath needs every lattice extent to be a multiple of 4, at most 252");
  Guard gd(c);
  invalidate_lp(g);
  return heatbath(c, g->g, g->U, beta, c11, c12, seed, sweep0, nsweeps, avg_plaq, action_density);
}

int sfh_hex_smear(sfh_ctx* c, sfh_gauge* g, const double alpha[3], int nsteps) {
  if (!c || !g || !alpha || nsteps < 0) return SFH_E_INVALID;
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(alpha[k])) return SFH_E_INVALID;
  Guard gd(c);
  invalidate_lp(g);
  return hex_smear(c, g->g, g->U, alpha, nsteps);
}

int sfh_tr_p2(sfh_ctx* c, const sfh_gauge* g, const double* P, double* out) {
  if (!c || !g || !P || !out) return SFH_E_INVALID;
  Guard gd(c);
  double2* dP = dbuf(c, "mom", 36 * (size_t)g->g.V);
  if (!dP) return SFH_E_NOMEM;
  if (int e = upload_links(c, g->g, P, dP)) return e;
  *out = tr_p2(c, g->g, dP);
  return std::isnan(*out) ? SFH_E_HIP : SFH_OK;
}

int sfh_link_update(sfh_ctx* c, sfh_gauge* g, double eps, const double* P, int check_hermitian) {
  if (!c || !g || !P) return SFH_E_INVALID;
  Guard gd(c);
  double2* dP = dbuf(c, "mom", 36 * (size_t)g->g.V);
  if (!dP) return SFH_E_NOMEM;
  if (int e = upload_links(c, g->g, P, dP)) return e;
  invalidate_lp(g);
  return link_update(c, g->g, g->U, dP, eps, check_hermitian);
}

int sfh_exp_iq(sfh_ctx* c, const double* Q18, double* out18) {
  if (!c || !Q18 || !out18) return SFH_E_INVALID;
  Guard gd(c);
  return exp_iq_one(c, Q18, out18);
}

int sfh_leapfrog(sfh_ctx* c, sfh_gauge* g, const sfh_md_params* p, const double* phi, double* P_inout,
                 double* X_out) {
  if (!c || !g || !P_inout || !md_ok(p)) return SFH_E_INVALID;
  if ((p->which & SFH_FORCE_FERMION) && !phi) return SFH_E_INVALID;
  Guard gd(c);
  const Geom& gg = g->g;
  const size_t n = 3 * (size_t)gg.V;
  double2* dphi = dbuf(c, "md_phi", n);
  double2* dX = dbuf(c, "md_X", n);
  double2* dP = dbuf(c, "mom", 36 * (size_t)gg.V);
  if (!dphi || !dX || !dP) return SFH_E_NOMEM;
  hipMemsetAsync(dX, 0, n * sizeof(double2), c->stream);
  if (p->which & SFH_FORCE_FERMION)
    if (int e = upload_spinor(c, gg, phi, dphi)) return e;
  if (int e = upload_links(c, gg, P_inout, dP)) return e;
  if (int e = md_leapfrog(c, g, p, dphi, dP, dX)) return e;
  if (int e = download_links(c, gg, dP, P_inout)) return e;
  if (X_out) return download_spinor(c, gg, dX, X_out);
  return SFH_OK;
}

int sfh_hmc_trajectory(sfh_ctx* c, sfh_gauge* g, const sfh_md_params* p, const double* chi, const double* P0,
                       double r_uniform, double* H, int* accepted) {
  if (!c || !g || !chi || !P0 || !H || !accepted || !md_ok(p)) return SFH_E_INVALID;
  Guard gd(c);
  const Geom& gg = g->g;
  const size_t n = 3 * (size_t)gg.V;
  double2* dchi = dbuf(c, "md_chi", n);
  double2* dphi = dbuf(c, "md_phi", n);
  double2* dX = dbuf(c, "md_X", n);
  double2* dP = dbuf(c, "mom", 36 * (size_t)gg.V);
  double2* dU0 = dbuf(c, "u_save", 36 * (size_t)gg.V);
  if (!dchi || !dphi || !dX || !dP || !dU0) return SFH_E_NOMEM;
  // Fermions.cpp:812-828
  if (int e = ensure_su3(c, g)) return e;
  const double sg_i = gauge_action(c, gg, g->U, p->beta, p->c11, p->c12, g->f64_su3);
  if (int e = upload_spinor(c, gg, chi, dchi)) return e;
  hop_f64(c->stream, gg, g->U, dchi, dphi, p->mass, 1.0, nullptr, nullptr, nullptr, 0, g->f64_su3);
  if (int e = upload_links(c, gg, P0, dP)) return e;
  const double tp_i = tr_p2(c, gg, dP);
  sfh_cg_params cp{};
  cp.mass = p->mass;
  cp.tol = p->tol;
  cp.max_iter = p->max_iter;
  cp.precision = p->precision;
  if (int e = cg_solve(c, g, &cp, dphi, dX, nullptr)) return e;
  double sf_i = 0.0;
  {
    double2* tmp = dX;
    dot_f64(c->stream, reinterpret_cast<const double*>(dphi), reinterpret_cast<const double*>(tmp), 2 * n,
            c->d_partials + 2 * kPartialSlot, c->aux_ctr(), &c->d_state[1].alpha);
    hipMemcpyAsync(c->h_scalar, &c->d_state[1].alpha, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (int e = hip_check(c, hipStreamSynchronize(c->stream), "S_f")) return e;
    sf_i = *c->h_scalar;
  }
  hipMemcpyAsync(dU0, g->U, 36 * (size_t)gg.V * sizeof(double2), hipMemcpyDeviceToDevice, c->stream);
  sfh_md_params mp = *p;
  mp.which = SFH_FORCE_GAUGE | SFH_FORCE_FERMION;
  int rc = md_leapfrog(c, g, &mp, dphi, dP, dX);
  if (rc != SFH_OK) {
    hipMemcpyAsync(g->U, dU0, 36 * (size_t)gg.V * sizeof(double2), hipMemcpyDeviceToDevice, c->stream);
    invalidate_lp(g);
    hipStreamSynchronize(c->stream);
    return rc;
  }
  const double tp_f = tr_p2(c, gg, dP);
  if (int e = ensure_su3(c, g)) return e;
  const double sg_f = gauge_action(c, gg, g->U, p->beta, p->c11, p->c12, g->f64_su3);
  dot_f64(c->stream, reinterpret_cast<const double*>(dphi), reinterpret_cast<const double*>(dX), 2 * n,
          c->d_partials + 2 * kPartialSlot, c->aux_ctr(), &c->d_state[1].alpha);
  hipMemcpyAsync(c->h_scalar, &c->d_state[1].alpha, sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (int e = hip_check(c, hipStreamSynchronize(c->stream), "S_f final")) return e;
  const double sf_f = *c->h_scalar;
  const double comp = tp_i - tp_f + sg_i - sg_f + sf_i - sf_f;
  H[0] = tp_i; H[1] = sg_i; H[2] = sf_i; H[3] = tp_f; H[4] = sg_f; H[5] = sf_f; H[6] = comp;
  *accepted = r_uniform < std::exp(comp) ? 1 : 0;
  if (!*accepted) {
    hipMemcpyAsync(g->U, dU0, 36 * (size_t)gg.V * sizeof(double2), hipMemcpyDeviceToDevice, c->stream);
    invalidate_lp(g);
  }
  return hip_check(c, hipStreamSynchronize(c->stream), "trajectory");
}

}  // extern "C"
