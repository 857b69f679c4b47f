// Batched CG engine (batch.hip): device tables and the host-side engine
// shared by sfh_cg_batch and the solve service.
#pragma once

#include <chrono>
#include <vector>

#include "hopcore.h"
#include "internal.h"

namespace sfh {

constexpr int kMaxRhs = 8;  // right-hand sides in flight per system

// Iterations per chunk of the mixed engine, for batched solves and the solve
// service alike (a result is the same alone or in any mix only under one
// chunk length): a reliable update that falls due runs at the chunk's end.
constexpr int kBatchChunk = 16;

// Launch shape of the batched hop, chosen by the autotuner.  hop_threads
// (128, 256 or 512 lanes, one site each, half per parity) applies to the
// first hop of an iteration, t = D^dag p, which has no reduction; rhs_tile
// (1, 2, 4 or 8) is how many right-hand sides of a system one lane applies
// each loaded link to (their accumulators held in registers), for both hops.  Neither changes any solve's arithmetic or
// reduction order: the second hop (with the reduction) always runs 256 lanes
// per block.
struct BatchConfig {
  int hop_threads = 256;
  int rhs_tile = kMaxRhs;
};

template <class T> struct BSys {
  Geom g;
  const typename Prec<T>::L* U;
  const double2* U64;  // FP64 links, for the reliable updates of mixed solves
  typename Prec<T>::R mass;
  double mass64;
  int nslot;
  int nb_hop;
  int slot[kMaxRhs];
};

template <class T> struct BSlot {
  CGState* st;
  typename Prec<T>::S *x, *r, *p, *t, *v;
  double* part;
  int V;
  // mixed precision: right-hand side, FP64 solution (the caller's x), and the
  // FP64 work vectors of the reliable update
  const double2* b;
  double2 *x64, *r64, *t64, *ax;
};

// T = float: FP32 inner iterations with FP64 reliable updates; T = double:
// the reference recurrence in FP64.
template <class T> class Engine {
 public:
  using S = typename Prec<T>::S;
  using R = typename Prec<T>::R;
  static constexpr bool kMixed = std::is_same<T, float>::value;

  struct Done {
    long long ticket;
    int slot;
    int sys;
    sfh_cg_stats stats;
    CGState state;
  };

  // use_graph: replay each chunk as a captured HIP graph (best when the set
  // of solves is stable for many chunks, as in a fixed batch); otherwise
  // launch the chunk's kernels directly.  A chunk is `chunk` iterations of
  // every active solve, then (mixed precision) the reliable update of every
  // solve that has one due, then one copy of all solve states to the host.
  Engine(sfh_ctx* c, double tol, int max_iter, double delta, int chunk, bool use_graph = true);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  // Registers a system (device links, mass) that may hold up to `cap`
  // solves at once, with `prealloc` slots of solve memory allocated now (the
  // rest on demand by grow()); returns its id, or minus a status.
  int add_system(const sfh_gauge* gauge, double mass, int cap, int prealloc);
  int add_system(const sfh_gauge* gauge, double mass, int pool) { return add_system(gauge, mass, pool, pool); }
  // One more slot for system sys (a spare slot of a removed system if one is
  // large enough, else a new allocation): its id, or minus a status.
  int grow(int sys);
  // New links (same lattice) or mass for a system with no solve in flight:
  // the SU(3) check and low-precision copy are redone.
  int update_system(int sys, const sfh_gauge* gauge, double mass);
  // Drops a system with no solve in flight; its slots become spares.
  int remove_system(int sys);
  int free_slots(int sys) const;
  int slots_of(int sys) const;
  int cap_of(int sys) const;
  // A spare slot large enough for a lattice of V sites exists.
  bool has_spare(int V) const;
  // Starts a solve of (D D^dag) x = b on system sys in a free slot: returns
  // the slot or -1 when the pool is full.  b and x are device spinors.
  int admit(int sys, const double2* b, double2* x, long long ticket);
  // One chunk of iterations for every active slot, then reliable updates
  // and completions.
  int step();
  int active() const;
  std::vector<Done> take_done();
  size_t bytes() const { return bytes_; }
  void set_config(const BatchConfig& cfg) {
    cfg_ = cfg;
    dirty_ = true;
  }
  const BatchConfig& config() const { return cfg_; }
  // Average microseconds of the first hop over the current tables with
  // launch shape cfg (n launches between HIP events).  Writes only t.
  int time_hop(const BatchConfig& cfg, int n, double* us);
  long chunks() const { return chunks_; }
  static size_t slot_bytes(int V);

 private:
  struct SysH {
    const sfh_gauge* gauge = nullptr;  // not dereferenced once removed
    Geom g{};
    double mass = 0.0;
    const typename Prec<T>::L* U = nullptr;
    const double2* U64 = nullptr;
    int cap = 0;
    bool removed = false;
    bool su3_lp = false, su3_64 = false;  // links pass the two-row checks
    std::vector<int> pool;
  };
  struct SlotH {
    void* mem = nullptr;  // this slot's device memory (owned_)
    size_t capacity = 0;
    int sys = 0;
    bool used = false, finished = false;
    int status = SFH_OK, rupd = 0;
    double phi_sq = 0.0, rsq_true = 0.0;
    const double2* b = nullptr;
    double2* xh = nullptr;
    double2* r64 = nullptr;
    unsigned* ctr = nullptr;  // this slot's reduction counters
    long long ticket = 0;
    std::chrono::steady_clock::time_point t_admit;
    CGState last{};
    BSlot<T> dev{};
  };
  int rebuild();
  int bind_system(SysH& s, const sfh_gauge* gauge, double mass);
  int layout_slot(SlotH& h, int V);
  void launch_tail();
  void hop_map(int threads, int tile, std::vector<int4>& out) const;

  sfh_ctx* c_;
  double tol_;
  int max_iter_;
  double delta_;
  int chunk_;
  bool use_graph_;
  void launch_chunk();
  std::vector<SysH> sys_;
  // every system's links are SU(3): the iteration's hops read two rows (to
  // kSu3Tol for the FP32 copies, kSu3TolF64 for FP64 links), and so do the
  // FP64 hops of the reliable updates (r12_64_)
  bool r12_ = true;
  bool r12_64_ = true;
  std::vector<SlotH> slots_;
  std::vector<int> spare_;  // slots of removed systems, kept for reuse
  std::vector<int> done_;
  std::vector<void*> owned_;
  size_t bytes_ = 0;
  bool dirty_ = true;
  int nbh0_ = 0, nbh_ = 0, nbb_ = 0;  // blocks of the first hop, the second hop, the vector kernels
  long chunks_ = 0;
  BatchConfig cfg_;
  BSys<T>* dsys_ = nullptr;
  BSlot<T>* dslot_ = nullptr;
  int4 *dmh0_ = nullptr, *dmh_ = nullptr;
  int2* dmb_ = nullptr;
  // slots in the current launch tables, their gathered states (device and
  // pinned host copies), and how many of them have finished
  std::vector<BSys<T>> hsys_;  // host copy of the system table
  std::vector<int> mapped_;
  int* dids_ = nullptr;
  CGState* dstates_ = nullptr;
  CGState* hstates_ = nullptr;
  size_t hstates_cap_ = 0;
  int finished_mapped_ = 0;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t gexec_ = nullptr;
};

}  // namespace sfh
