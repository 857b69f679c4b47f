// Device-resident solve service for concurrent HMC chains.
//
// Each chain registers its device-resident links and mass once, with a cap
// on the solves it may have in flight (up to 8).  Requests (a right-hand
// side and an output field, both device-resident) go into a FIFO queue; a
// worker thread admits queued requests into solve slots: a free slot of the
// chain, else a new slot when the chain is below its cap and the memory
// allows it (the service budget, and free HBM as hipMemGetInfo reports it
// at that moment, less a headroom); otherwise the request waits until a
// slot frees up.  Slots stay with their chain for later requests and pass
// to other chains when a chain is removed.  The worker advances every solve
// in flight by one chunk of batched CG iterations (continuous batching:
// requests that arrive while others run join at the next chunk boundary),
// runs the FP64 reliable updates, and completes finished requests.  Latency
// is measured from submit to completion.  Each result is the same as the
// solve alone: a chain's kernels and reduction order do not depend on the
// rest of the batch.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "batch.h"

struct sfh_service {
  sfh_ctx* ctx = nullptr;  // the service's own context (stream) used only by the worker
  std::unique_ptr<sfh::Engine<float>> eng;
  size_t budget = 0;
  std::thread worker;
  std::mutex emu;  // the engine
  std::mutex mu;   // queue, results, status
  std::condition_variable cv_work, cv_done;
  struct Req {
    long long ticket;
    int chain;
    const double2* b;
    double2* x;
    std::chrono::steady_clock::time_point t_submit;
  };
  struct Res {
    sfh_cg_stats stats;
    std::string err;
  };
  std::deque<Req> queue;
  std::map<long long, std::chrono::steady_clock::time_point> submitted;
  std::map<long long, Res> results;
  std::vector<int> chain_V;        // lattice volume per chain, -1 once removed
  std::vector<int> chain_pending;  // requests queued or in flight per chain
  long long next_ticket = 1;
  bool stop = false;
  int worker_status = SFH_OK;
  std::string worker_err;
  long completed = 0;
  long slots_refused = 0;  // admissions deferred because memory refused a new slot
  // callers waiting for the engine lock: the worker lets them in between
  // chunks instead of taking the lock straight back
  std::atomic<int> want_engine{0};
};

namespace {

// The engine lock for an API call: announced to the worker first, so a
// busy worker hands it over at its next chunk boundary.
struct EngineLock {
  sfh_service* s;
  std::unique_lock<std::mutex> lk;
  explicit EngineLock(sfh_service* s_) : s(s_) {
    ++s->want_engine;
    lk = std::unique_lock<std::mutex>(s->emu);
    --s->want_engine;
  }
};

}  // namespace

namespace {

// Whether a new solve slot for a lattice of V sites may be allocated now: a
// spare slot fits, or the slot stays within the service budget and free HBM
// (hipMemGetInfo, at this moment) keeps a headroom of max(256 MiB, 1/64 of
// the device) after it.
bool slot_memory_ok(sfh_service* s, int V) {
  if (s->eng->has_spare(V)) return true;
  const size_t need = sfh::Engine<float>::slot_bytes(V);
  if (s->budget && s->eng->bytes() + need > s->budget) return false;
  size_t free_b = 0, total = 0;
  if (hipMemGetInfo(&free_b, &total) != hipSuccess) return false;
  const size_t headroom = std::max<size_t>(size_t(256) << 20, total / 64);
  return free_b >= need + headroom;
}

// Lock order: emu (engine) before mu (queue, results).
void worker_loop(sfh_service* s) {
  hipSetDevice(s->ctx->device);
  for (;;) {
    while (s->want_engine.load() > 0) std::this_thread::yield();
    std::unique_lock<std::mutex> le(s->emu);
    std::vector<sfh_service::Req> admit;
    std::vector<sfh_service::Req> refused;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      // FIFO per chain: a request waits only for its own chain's slots; a
      // chain below its cap gets a new slot if the memory allows it now
      std::map<int, int> room;
      std::map<int, bool> blocked;  // memory refused a slot for this chain in this pass
      for (auto it = s->queue.begin(); it != s->queue.end();) {
        auto r = room.find(it->chain);
        if (r == room.end()) r = room.emplace(it->chain, s->eng->free_slots(it->chain)).first;
        if (r->second == 0 && !blocked[it->chain] && s->eng->slots_of(it->chain) < s->eng->cap_of(it->chain)) {
          if (slot_memory_ok(s, s->chain_V[it->chain]) && s->eng->grow(it->chain) >= 0) {
            ++r->second;
          } else {
            blocked[it->chain] = true;
            ++s->slots_refused;
          }
        }
        if (r->second > 0) {
          --r->second;
          admit.push_back(*it);
          it = s->queue.erase(it);
        } else if (s->eng->active() == 0 && admit.empty() && s->eng->slots_of(it->chain) == 0) {
          // nothing runs that could free memory, and the chain has no slot:
          // the request fails instead of waiting for ever
          refused.push_back(*it);
          it = s->queue.erase(it);
        } else {
          ++it;
        }
      }
      for (auto& q : refused) {
        sfh_service::Res res{};
        res.stats.status = SFH_E_NOMEM;
        res.err = "no device memory for a solve slot of this chain";
        s->submitted.erase(q.ticket);
        s->results[q.ticket] = res;
        --s->chain_pending[q.chain];
      }
      if (!refused.empty()) s->cv_done.notify_all();
    }
    for (auto& q : admit) {
      if (s->eng->admit(q.chain, q.b, q.x, q.ticket) < 0) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->worker_status = SFH_E_HIP;
        s->worker_err = s->ctx->err;
      }
    }
    int rc = SFH_OK;
    if (s->eng->active() > 0) rc = s->eng->step();
    auto done = s->eng->take_done();
    const bool idle = s->eng->active() == 0;
    std::unique_lock<std::mutex> lk(s->mu);
    for (auto& d : done) {
      sfh_service::Res r;
      r.stats = d.stats;
      const auto t0 = s->submitted[d.ticket];
      r.stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (d.stats.status != SFH_OK) {
        sfh::map_status(s->ctx, d.state, d.state.max_iter);
        r.err = s->ctx->err;
      }
      s->submitted.erase(d.ticket);
      s->results[d.ticket] = r;
      ++s->completed;
      --s->chain_pending[d.sys];
    }
    if (rc != SFH_OK) {
      s->worker_status = rc;
      s->worker_err = s->ctx->err;
    }
    s->cv_done.notify_all();
    if (rc != SFH_OK) break;
    le.unlock();
    if (idle) {
      if (s->stop && s->queue.empty()) break;
      if (s->queue.empty()) s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
      if (s->stop && s->queue.empty()) break;
    }
  }
}

}  // namespace

extern "C" {

int sfh_service_create(int device, size_t mem_budget, double tol, int max_iter, sfh_service** out) {
  if (!out || tol <= 0 || max_iter < 1) return SFH_E_INVALID;
  *out = nullptr;
  auto* s = new sfh_service();
  if (int e = sfh_create(device, &s->ctx)) {
    delete s;
    return e;
  }
  s->budget = mem_budget;
  // chunks of kBatchChunk iterations replayed as HIP graphs (reliable-update
  // threshold: the engine's FP32 default): a graph is recaptured
  // when requests are admitted (or many have finished), and replayed while
  // the set of solves in flight holds; one launch per chunk keeps a lightly
  // loaded service from being bound by kernel launches on the host
  s->eng.reset(new sfh::Engine<float>(s->ctx, tol, max_iter, 0.0, sfh::kBatchChunk, true));
  s->worker = std::thread(worker_loop, s);
  *out = s;
  return SFH_OK;
}

int sfh_service_add_chain(sfh_service* s, const sfh_gauge* g, double mass, int max_inflight, int* chain) {
  if (!s || !g || !chain || max_inflight < 1 || max_inflight > sfh::kMaxRhs) return SFH_E_INVALID;
  EngineLock le(s);
  // a chain one solve of which can never fit the budget is refused now;
  // its slots are allocated as its requests are admitted
  const size_t need = sfh::Engine<float>::slot_bytes(g->g.V);
  if (s->budget && need > s->budget)
    return sfh::set_err(s->ctx, SFH_E_NOMEM, "chain does not fit the service memory budget");
  const int id = s->eng->add_system(g, mass, max_inflight, 0);
  if (id < 0) return -id;
  std::lock_guard<std::mutex> lk(s->mu);
  *chain = id;
  if ((int)s->chain_V.size() <= id) {
    s->chain_V.resize(id + 1, -1);
    s->chain_pending.resize(id + 1, 0);
  }
  s->chain_V[id] = g->g.V;
  s->chain_pending[id] = 0;
  return SFH_OK;
}

// New links (same lattice) or mass for a chain with nothing queued or in
// flight, e.g. after each molecular-dynamics step of its trajectory.
int sfh_service_update_chain(sfh_service* s, int chain, const sfh_gauge* g, double mass) {
  if (!s || !g) return SFH_E_INVALID;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (chain >= 0 && chain < (int)s->chain_pending.size() && s->chain_pending[chain] != 0)
      return sfh::set_err(s->ctx, SFH_E_INVALID, "chain has requests queued or in flight");
  }
  EngineLock le(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (chain < 0 || chain >= (int)s->chain_V.size() || s->chain_V[chain] < 0)
      return sfh::set_err(s->ctx, SFH_E_INVALID, "no such chain");
    if (s->chain_pending[chain] != 0)
      return sfh::set_err(s->ctx, SFH_E_INVALID, "chain has requests queued or in flight");
    if (g->g.V != s->chain_V[chain]) return sfh::set_err(s->ctx, SFH_E_INVALID, "chain lattice changed size");
  }
  return s->eng->update_system(chain, g, mass);
}

// Drops a chain with nothing queued or in flight; its solve slots are kept
// for other chains.  The chain id is not reused.
int sfh_service_remove_chain(sfh_service* s, int chain) {
  if (!s) return SFH_E_INVALID;
  EngineLock le(s);
  std::lock_guard<std::mutex> lk(s->mu);
  if (chain < 0 || chain >= (int)s->chain_V.size() || s->chain_V[chain] < 0)
    return sfh::set_err(s->ctx, SFH_E_INVALID, "no such chain");
  if (s->chain_pending[chain] != 0)
    return sfh::set_err(s->ctx, SFH_E_INVALID, "chain has requests queued or in flight");
  if (int e = s->eng->remove_system(chain)) return e;
  s->chain_V[chain] = -1;
  return SFH_OK;
}

// Solve-slot memory of the service and how often a slot was refused.
int sfh_service_memory(sfh_service* s, size_t* slot_bytes, long* refusals) {
  if (!s) return SFH_E_INVALID;
  EngineLock le(s);
  std::lock_guard<std::mutex> lk(s->mu);
  if (slot_bytes) *slot_bytes = s->eng->bytes();
  if (refusals) *refusals = s->slots_refused;
  return SFH_OK;
}

int sfh_service_submit(sfh_service* s, int chain, const sfh_field* rhs, sfh_field* x, long long* ticket) {
  if (!s || !rhs || !x || !ticket || rhs == x) return SFH_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (chain < 0 || chain >= (int)s->chain_V.size() || s->chain_V[chain] < 0) return SFH_E_INVALID;
  if (rhs->g.V != s->chain_V[chain] || x->g.V != s->chain_V[chain]) return SFH_E_INVALID;
  if (s->worker_status != SFH_OK) return s->worker_status;
  ++s->chain_pending[chain];
  const long long t = s->next_ticket++;
  const auto now = std::chrono::steady_clock::now();
  s->queue.push_back({t, chain, rhs->d, x->d, now});
  s->submitted[t] = now;
  *ticket = t;
  s->cv_work.notify_one();
  return SFH_OK;
}

int sfh_service_poll(sfh_service* s, double timeout_s, int max, long long* tickets, sfh_cg_stats* stats, int* n) {
  if (!s || !tickets || !n || max < 1) return SFH_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
  while (s->results.empty() && s->worker_status == SFH_OK) {
    if (s->cv_done.wait_until(lk, until) == std::cv_status::timeout) break;
  }
  int k = 0;
  for (auto it = s->results.begin(); it != s->results.end() && k < max; ++k) {
    tickets[k] = it->first;
    if (stats) stats[k] = it->second.stats;
    it = s->results.erase(it);
  }
  *n = k;
  return s->worker_status;
}

int sfh_service_wait(sfh_service* s, long long ticket, double timeout_s, sfh_cg_stats* stats) {
  if (!s) return SFH_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
  for (;;) {
    auto it = s->results.find(ticket);
    if (it != s->results.end()) {
      if (stats) *stats = it->second.stats;
      const int st = it->second.stats.status;
      s->results.erase(it);
      return st;
    }
    if (s->worker_status != SFH_OK) return s->worker_status;
    if (s->cv_done.wait_until(lk, until) == std::cv_status::timeout) return SFH_E_HIP;
  }
}

const char* sfh_service_last_error(const sfh_service* s) {
  return s ? (s->worker_err.empty() ? s->ctx->err.c_str() : s->worker_err.c_str()) : "no service";
}

int sfh_service_destroy(sfh_service* s) {
  if (!s) return SFH_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  if (s->worker.joinable()) s->worker.join();
  s->eng.reset();
  sfh_destroy(s->ctx);
  delete s;
  return SFH_OK;
}

}  // extern "C"
