// Internal host-side structures and kernel entry points of libsfh.
#pragma once

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sfh.h"
#include "common.h"
#include "hopcore.h"

namespace sfh {

// Reliable-update threshold when the caller passes none: a drop of |r| by
// this factor since the last update.  FP16 storage: 0.1 (smaller values let
// the FP16 recursion drift, 32^4: 251 iterations at 0.05 against 229); FP32
// storage: 0.003 (the FP32 recursion holds: the same iteration counts from
// 16^4 to 32^4 as at 0.1 with half the updates, 16^4 2.63 -> 2.53 ms).
constexpr double kDeltaF16 = 0.1, kDeltaF32 = 0.003;

// Per-solve device scalars.  Written only by the last workgroup of a grid
// reduction (or by the host between launches), read by later launches, so
// kernels of one stream never race on them.
struct CGState {
  double rsq;       // iterated |r|^2
  double rsq_prev;  // |r|^2 before the last update (for beta after a reliable update)
  double alpha, beta, pv;
  double target;    // tol^2 |phi|^2
  double rmax;      // mixed precision: max |r| since the last reliable update
  double delta;     // mixed precision: reliable-update factor
  double rsq_true;  // mixed precision: FP64 true |r|^2 at the last reliable update
  int iter, done, status, max_iter;
  int pause_reason;  // mixed precision: 1 reliable update due, 2 iterated residual converged, 3 budget spent
  int mixed;
  int ru_pending;    // mixed precision: the reliable-update kernels run
  int ru_apply;      // mixed precision: restart the sloppy residual and direction
  int rupd;          // reliable updates done
  int converged;     // mixed precision: the FP64 true residual met the target
  int fin_pending;   // mixed iteration, deferred dots: the last B's dots still await their finisher
  int pbuf;          // mixed iteration, fused form: buffer of the latest r_k and p_k
  unsigned* ctr;     // grid-reduction arrival counters (kCtrKinds x kCtrPerKind, device)
};

enum { CNT_PV = 0, CNT_RR = 1, CNT_AUX = 2, CNT_RU = 3 };

// Arrival counting of the grid reductions is split over kRedGroups counters
// (each on its own 128-byte line) and one top counter, so no single address
// takes more than kRedGroups atomics per reduction.
constexpr int kRedGroups = 64;
constexpr int kCtrStride = 32;
constexpr int kCtrPerKind = (kRedGroups + 1) * kCtrStride;
constexpr int kCtrKinds = 4;
constexpr size_t kCtrBytes = sizeof(unsigned) * kCtrKinds * kCtrPerKind;
SFH_HD unsigned* ctr_of(unsigned* arena, int kind) { return arena + kind * kCtrPerKind; }

// Low-precision link / spinor storage (FP16 uses a per-site FP32 norm).
struct LinksLP {
  float2* u32 = nullptr;  // 4*9*V
  __half2* u16 = nullptr; // 4*9*V, packed per direction (see Prec<Half>)
};

}  // namespace sfh

struct sfh_field {
  sfh::Geom g;
  double2* d = nullptr;  // 3*V complex, device layout
};

struct sfh_gauge {
  sfh::Geom g;
  double2* U = nullptr;     // 4*9*V complex, device layout
  sfh::LinksLP lp;
  bool lp32_valid = false, lp16_valid = false;
  // whether the FP64 links are SU(3) (third row = conj(row0 x row1)), so
  // kernels may read two rows and rebuild the third: to kSu3Tol for the
  // low-precision copies (lp_su3), to kSu3TolF64 for the FP64 kernels
  // (f64_su3); measured once after the links change (su3_known)
  bool su3_known = false;
  bool lp_su3 = false;
  bool f64_su3 = false;
};

struct sfh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::recursive_mutex mu;
  // grow-only device scratch, keyed by purpose
  std::map<std::string, std::pair<void*, size_t>> scratch;
  // pinned host staging for scalars
  sfh::CGState* h_state = nullptr;
  // d_state[0]: solves; [1]: reductions, timings; [2]: second copy of a
  // mixed solve's state (iterations whose dots are finished by the next
  // kernel alternate copies);
  // [3]: scratch of the mixed-kernel timings
  sfh::CGState* d_state = nullptr;
  double* d_partials = nullptr;  // reduction partials (8 slots of kPartialSlot)
  unsigned* d_ctr = nullptr;     // reduction counters of d_state[0] and d_state[1] (2 x kCtrBytes)
  unsigned* aux_ctr() const { return sfh::ctr_of(d_ctr + sfh::kCtrKinds * sfh::kCtrPerKind, sfh::CNT_AUX); }
  double* h_scalar = nullptr;
  bool profiling = false;
  bool graph_replay = true;  // replay CG chunks as HIP graphs (false: direct launches)
  int prof_every = 1;   // time one launch in prof_every per kernel name
  std::map<std::string, long> prof_seen;
  std::map<std::string, std::pair<double, long>> ktime;
  // batched-hop launch shape (threads per block, right-hand sides per block)
  // chosen by the tuner, per batch shape
  std::map<std::string, std::pair<int, int>> batch_cfg;
  // event pool for kernel timing: pending (name, start, stop) resolved lazily
  std::vector<hipEvent_t> ev_pool;
  struct Pending { std::string name; hipEvent_t a, b; };
  std::vector<Pending> pending;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // whole-solve timing
  hipEvent_t chunk_ev[2] = {nullptr, nullptr};  // CG chunks in flight (no timing)
  // captured CG chunks, keyed by the buffers and parameters they replay
  std::map<std::string, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
};

namespace sfh {

// Doubles per reduction slot: one partial per workgroup and value plus one
// per arrival group and value.  The four-lane hop on the largest allowed
// lattice (2^24 sites) reduces one value over 2^24 / 64 = 2^18 workgroups;
// the two-kernel mixed iteration reduces four over 2^24 / 256 = 2^16.
constexpr int kPartialSlot = (1 << 18) + 4 * kRedGroups;

// ------------------------------------------------------------------ helpers
int set_err(sfh_ctx* c, int code, const std::string& msg);
int hip_check(sfh_ctx* c, hipError_t e, const char* what);
void* scratch(sfh_ctx* c, const std::string& key, size_t bytes);
// Destroys every cached CG chunk graph (after the stream drains): called
// whenever device memory that a graph may reference is freed, so no cached
// graph ever holds a dangling pointer.
void drop_graphs(sfh_ctx* c);
// Returns d_state[1] (reductions and kernel timings) to an idle state.
void reset_aux_state(sfh_ctx* c);

// Profiling: brackets the launches of its scope with a pair of events on the
// context stream when profiling is on (no synchronisation); the durations are
// accumulated by resolve_timers().
struct KTimer {
  sfh_ctx* c;
  const char* name;
  hipEvent_t a = nullptr;
  KTimer(sfh_ctx* c_, const char* n);
  ~KTimer();
};
void resolve_timers(sfh_ctx* c);

// ------------------------------------------------------------------ layout (hop.hip)
void links_aos_to_soa(hipStream_t s, const Geom& g, const double* aos, double2* soa);
void links_soa_to_aos(hipStream_t s, const Geom& g, const double2* soa, double* aos);
void spinor_aos_to_soa(hipStream_t s, const Geom& g, const double* aos, double2* soa);
void spinor_soa_to_aos(hipStream_t s, const Geom& g, const double2* soa, double* aos);
void links_to_f32(hipStream_t s, const Geom& g, const double2* U, float2* u32);
void links_to_f16(hipStream_t s, const Geom& g, const double2* U, __half2* u16);
// max over links of |row2 - conj(row0 x row1)| as the bits of a float, into
// *dev (which must start at 0; non-finite entries give +inf).  Below kSu3Tol
// the FP32 and FP16 kernels read two rows and rebuild the third; below
// kSu3TolF64 (the rounding of a few FP64 operations) the FP64 kernels do.
void su3_deviation(hipStream_t s, const Geom& g, const double2* U, unsigned* dev);
constexpr float kSu3Tol = 1e-10f;
constexpr float kSu3TolF64 = 1e-14f;

// ------------------------------------------------------------------ hop (hop.hip)
// out = a*in + (b/2) * sum_mu eta [bup U psi(n+mu) - bdn U^dag psi(n-mu)]
// If dot_with != nullptr, also reduces Re<dot_with, out> into st (pv, alpha).
// gate 0: skip when st->done (a CG step); gate 1: run only while
// st->ru_pending (a reliable-update step).
// r12: the links are SU(3); read two rows and rebuild the third.
void hop_f64(hipStream_t s, const Geom& g, const double2* U, const double2* in, double2* out, double a, double b,
             const double2* dot_with, CGState* st, double* partials, int gate, bool r12);
void hop_f32(hipStream_t s, const Geom& g, const float2* U, const float2* in, float2* out, float a, float b,
             const float2* dot_with, CGState* st, double* partials, bool r12);
void hop_f16(hipStream_t s, const Geom& g, const __half2* U, const H16* in, H16* out, float a, float b,
             const H16* dot_with, CGState* st, double* partials, bool r12);

// Hop kernel form: 0 one lane per site, 1 four lanes per site (one per
// direction); -1 restores the size-based default.  prec: 0 f64, 1 f32, 2 f16.
void set_hop_variant(int prec, int variant);
int get_hop_variant(int prec, const Geom& g);

// ------------------------------------------------------------------ BLAS (blas.hip)
void dot_f64(hipStream_t s, const double* a, const double* b, size_t n, double* partials, unsigned* cnt,
             double* result);
// CG update kernels for each storage precision.
void cg_xr_f64(hipStream_t s, const Geom& g, double2* x, double2* r, const double2* p, const double2* v,
               CGState* st, double* partials);
void cg_p_f64(hipStream_t s, const Geom& g, double2* p, const double2* r, CGState* st);
void cg_p_f32(hipStream_t s, const Geom& g, float2* p, const float2* r, CGState* st);
// mixed-precision support
void f64_to_f32(hipStream_t s, const double2* in, float2* out, size_t n);
void f64_to_f16(hipStream_t s, const Geom& g, const double2* in, H16* out);
void f32_to_f64(hipStream_t s, const float2* in, double2* out, size_t n);
void f16_to_f64(hipStream_t s, const Geom& g, const H16* in, double2* out);  // halves times the record's norm
void h16_ones(hipStream_t s, H16* f, int V);  // every colour (1 + i)
// *h_state (pinned host memory) = *d_state, on the stream
void state_out(hipStream_t s, const CGState* d_state, CGState* h_state);
void acc_f32_into_f64(hipStream_t s, double2* xh, float2* xl, size_t n);  // xh += xl; xl = 0
void xpay_f64(hipStream_t s, double2* y, const double2* x, double a, size_t n);  // y = x + a*y
void axpy_f64(hipStream_t s, double2* y, const double2* x, double a, size_t n);  // y += a*x
void sub_f64(hipStream_t s, double2* out, const double2* a, const double2* b, size_t n);  // out = a - b

// ------------------------------------------------------------------ gauge (gauge.hip)
// r12: the links are SU(3) to kSu3TolF64; the gauge force reads two rows.
int force_field(sfh_ctx* c, const Geom& g, const double2* U, const double2* X, const double2* Y, double beta,
                double c11, double c12, unsigned which, double2* F, double2* P, double coef, bool r12);
// Every extent below 256: the staple walks (force, action, plaquette, HEX)
// pack each coordinate into 8 bits.
bool staple_dims_ok(const Geom& g);
// r12: the links pass the FP64 SU(3) check (ensure_su3), so the staple walks read two rows
double gauge_action(sfh_ctx* c, const Geom& g, const double2* U, double beta, double c11, double c12, bool r12);
double avg_plaq(sfh_ctx* c, const Geom& g, const double2* U, bool r12);
double tr_p2(sfh_ctx* c, const Geom& g, const double2* P);
int link_update(sfh_ctx* c, const Geom& g, double2* U, const double2* P, double eps, int check);
int exp_iq_one(sfh_ctx* c, const double* Q18, double* out18);
int polyakov_grid(sfh_ctx* c, const Geom& g, const double2* U, double* out);
// Polyakov-loop correlators C(r^2) for r^2 = r2_first .. r2_first + count - 1
// into out (host, count complex), from the device links or a host grid.
int polyakov_correlators(sfh_ctx* c, const Geom& g, const double2* U, int r2_first, int count, double* out);
int polyakov_correlators_grid(sfh_ctx* c, const int L[3], const double* grid, int r2_first, int count, double* out);
// Pieces of the correlators for callers that bring their own device buffers:
// the displacement tables, the grid from device links, the block count of
// the first kernel, and the run (see gauge.hip).
void corr_tables(const int L[3], int r2_first, int count, std::vector<int4>& disp, std::vector<int>& off);
void polyakov_grid_into(sfh_ctx* c, const Geom& g, const double2* U, double2* P);
int poly_corr_blocks(const int L[3]);
int poly_corr_run(sfh_ctx* c, const int L[3], const double2* P, const std::vector<int4>& disp,
                  const std::vector<int>& off, int4* d_disp, int* d_off, double* d_part, double* d_out, double* out);
// Allocation-free entry points (noalloc.hip): per-device address of the
// library's static device work area, looked up once per device by the
// context's creation.
int noalloc_init(int device);

// ------------------------------------------------------------------ heatbath (heatbath.hip)
bool heatbath_dims_ok(const Geom& g);
int heatbath(sfh_ctx* c, const Geom& g, double2* U, double beta, double c11, double c12, unsigned long long seed,
             unsigned sweep0, int nsweeps, double* plaq, double* act);

// ------------------------------------------------------------------ smearing (smear.hip)
int hex_smear(sfh_ctx* c, const Geom& g, double2* U, const double alpha[3], int nsteps);

// ------------------------------------------------------------------ solver (solver.cpp)
int cg_solve(sfh_ctx* c, sfh_gauge* gauge, const sfh_cg_params* p, const double2* phi, double2* x,
             sfh_cg_stats* stats);
int apply_ddag_dev(sfh_ctx* c, const sfh_gauge* gauge, double mass, const double2* in, double2* out);
// Average microseconds of the two steps of one mixed-precision iteration.
int time_mixed(sfh_ctx* c, sfh_gauge* gauge, int precision, int n, double* us_a, double* us_b);
// Mixed iteration form per precision: 0 fused, 1 split, -1 by lattice size.
void set_mix_form(int precision, int form);
int get_mix_form(int precision, const Geom& g);
// Maps a stopped solve's device status to the reference's error text.
int map_status(sfh_ctx* c, const CGState& s, int max_iter);
// FP64 dot of two device vectors of ncomplex complex numbers (synchronises).
double dev_dot(sfh_ctx* c, const double2* a, const double2* b, size_t ncomplex);
// Brings the FP32 / FP16 link copy of g up to date with its FP64 links.
int ensure_lp(sfh_ctx* c, sfh_gauge* g, int precision);
// Measures whether the links are SU(3) if they changed since the last time
// (one pass over the links and one host sync): sets lp_su3 and f64_su3.
int ensure_su3(sfh_ctx* c, sfh_gauge* g);

// ------------------------------------------------------------------ batched solves (batch.hip)
struct BatchSys {
  const sfh_gauge* gauge;
  double mass;
  int nrhs;
  const double2* const* rhs;  // device spinors, nrhs of them
  double2* const* x;
};
int cg_batch(sfh_ctx* c, int nsys, const BatchSys* sys, const sfh_cg_params* p, sfh_cg_stats* stats);
// Times the batched hop's launch shapes on this batch and records the
// fastest for batches of the same shape (lattices and right-hand-side counts).
int tune_batch(sfh_ctx* c, int nsys, const BatchSys* sys, const sfh_cg_params* p, int* hop_threads, int* rhs_tile,
               double* us);

// host generators (gen.cpp)
void gen_links(const int* dims, double eps, unsigned long long seed, double* out);
void gen_spinor(const int* dims, double sigma, unsigned long long seed, double* out);
void gen_momenta(const int* dims, unsigned long long seed, double* out);

}  // namespace sfh
