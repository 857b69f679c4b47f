// Device-resident solve service for concurrent HMC chains.
//
// Each chain registers its device-resident links and mass once; the service
// reserves a fixed pool of solve memory for it (up to 8 solves in flight per
// chain) and refuses a chain whose pool would exceed the memory budget.
// Requests (a right-hand side and an output field, both device-resident)
// go into a FIFO queue; a worker thread admits queued requests whenever
// their chain has a free slot, advances every solve in flight by one chunk
// of batched CG iterations (continuous batching: requests that arrive
// while others run join at the next chunk boundary), runs the FP64
// reliable updates, and completes finished requests.  Latency is measured
// from submit to completion.  Each result is the same as the solve alone:
// a chain's kernels and reduction order do not depend on the rest of the
// batch.
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
  std::vector<int> chain_V;
  long long next_ticket = 1;
  bool stop = false;
  int worker_status = SFH_OK;
  std::string worker_err;
  long completed = 0;
};

namespace {

// Lock order: emu (engine) before mu (queue, results).
void worker_loop(sfh_service* s) {
  hipSetDevice(s->ctx->device);
  for (;;) {
    std::unique_lock<std::mutex> le(s->emu);
    std::vector<sfh_service::Req> admit;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      // FIFO per chain: a request waits only for its own chain's free slots
      std::map<int, int> room;
      for (auto it = s->queue.begin(); it != s->queue.end();) {
        auto r = room.find(it->chain);
        if (r == room.end()) r = room.emplace(it->chain, s->eng->free_slots(it->chain)).first;
        if (r->second > 0) {
          --r->second;
          admit.push_back(*it);
          it = s->queue.erase(it);
        } else {
          ++it;
        }
      }
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
  // chunks of 8 iterations replayed as HIP graphs: a graph is recaptured
  // when requests are admitted (or many have finished), and replayed while
  // the set of solves in flight holds; one launch per chunk keeps a lightly
  // loaded service from being bound by kernel launches on the host
  s->eng.reset(new sfh::Engine<float>(s->ctx, tol, max_iter, 0.1, sfh::kBatchChunk, true));
  s->worker = std::thread(worker_loop, s);
  *out = s;
  return SFH_OK;
}

int sfh_service_add_chain(sfh_service* s, const sfh_gauge* g, double mass, int max_inflight, int* chain) {
  if (!s || !g || !chain || max_inflight < 1 || max_inflight > sfh::kMaxRhs) return SFH_E_INVALID;
  std::lock_guard<std::mutex> le(s->emu);
  const size_t need = sfh::Engine<float>::slot_bytes(g->g.V) * max_inflight;
  if (s->budget && s->eng->bytes() + need > s->budget)
    return sfh::set_err(s->ctx, SFH_E_NOMEM, "chain does not fit the service memory budget");
  const int id = s->eng->add_system(g, mass, max_inflight);
  if (id < 0) return -id;
  std::lock_guard<std::mutex> lk(s->mu);
  *chain = id;
  s->chain_V.push_back(g->g.V);
  return SFH_OK;
}

int sfh_service_submit(sfh_service* s, int chain, const sfh_field* rhs, sfh_field* x, long long* ticket) {
  if (!s || !rhs || !x || !ticket || rhs == x) return SFH_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (chain < 0 || chain >= (int)s->chain_V.size()) return SFH_E_INVALID;
  if (rhs->g.V != s->chain_V[chain] || x->g.V != s->chain_V[chain]) return SFH_E_INVALID;
  if (s->worker_status != SFH_OK) return s->worker_status;
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
