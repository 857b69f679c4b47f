/*
 * sfh.h - C ABI of the MI355X (gfx950) staggered-fermion HMC hot path.
 *
 * Device-resident replacement for the field-level functions of the
 * reference's Fermions.h / Lattice.h.  Each entry point below names the
 * reference function it replaces (file:line).  Host buffers use the
 * reference's array-of-structures layouts:
 *   links / momenta / forces : double[V*4*18], (s*4+mu)*18 + (3a+b)*2 + re/im
 *                              (Lattice.cpp:135-151, Parameters.h:11,20)
 *   spinors                  : double[V*6],    s*6 + 2c + re/im
 *                              (Fermions.cpp:99-116, Parameters.h:19)
 *   site s = ((n0*L1 + n1)*L2 + n2)*L3 + n3, time = direction 3 carrying the
 *   antiperiodic fermion boundary (Fermions.cpp:20-94).
 * On the device, fields live in a structure-of-arrays, site-parity-ordered
 * layout owned by the library (see sfh_gauge / sfh_field).
 *
 * Nothing throws across this ABI.  Every call returns an sfh_status; the
 * message of the last failure on a context is sfh_last_error().  A context
 * owns one HIP stream; calls on one context are serialised by an internal
 * lock, so one context per host thread gives concurrency.
 */
#ifndef SFH_H
#define SFH_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFH_ABI_VERSION 1

/* Status codes.  The failure cases mirror the reference's exceptions:
 * BREAKDOWN / NONFINITE_* / NOCONV  <- std::runtime_error in solve_CG_DDdagger
 *                                      (Fermions.cpp:573-600, :609-614)
 * NOT_HERMITIAN                      <- std::runtime_error in the link update
 *                                      (Fermions.cpp:712-716, :752-756)
 * EXP_QCUBED / EXP_NEG_TRACE         <- std::invalid_argument in
 *                                      numerically_stable_matrix_exponential
 *                                      (Fermions.cpp:848-857)
 * IO                                 <- the false return of
 *                                      save_all_lattice_configs /
 *                                      load_all_lattice_configs
 *                                      (Lattice.cpp:810-847)
 * SAMPLER                            <- no counterpart: a heatbath rejection
 *                                      loop (SU3_Sampling.cpp:17-29, :53-68,
 *                                      :175-191) that found no sample in 2^20
 *                                      tries (non-finite links, or a*beta so
 *                                      small that the reference would loop for
 *                                      ever)                                  */
typedef enum {
  SFH_OK = 0,
  SFH_E_INVALID = 1,
  SFH_E_NOMEM = 2,
  SFH_E_HIP = 3,
  SFH_E_NODEVICE = 4,
  SFH_E_BREAKDOWN = 5,
  SFH_E_NONFINITE_ALPHA = 6,
  SFH_E_NONFINITE_BETA = 7,
  SFH_E_NOCONV = 8,
  SFH_E_NOT_HERMITIAN = 9,
  SFH_E_EXP_QCUBED = 10,
  SFH_E_EXP_NEG_TRACE = 11,
  SFH_E_SAMPLER = 12,
  SFH_E_IO = 13
} sfh_status;

/* Arithmetic of the CG inner iterations.  F64 follows the reference's
 * recurrence (Fermions.cpp:553-617).  F32 and F16 run the inner iterations in
 * low precision with FP64 reliable updates (true residual recomputed in FP64
 * once the iterated residual has dropped by reliable_delta, at the end of the
 * chunk of iterations in which that happens: up to 23 FP32 or 9 FP16
 * iterations later), FP64 global reductions, and stop on the FP64 true
 * residual.  F16 stores spinors as
 * 16-bit floats with a per-site FP32 max-norm and links as 16-bit floats,
 * accumulating in FP32. */
typedef enum { SFH_PREC_F64 = 0, SFH_PREC_F32 = 1, SFH_PREC_F16 = 2 } sfh_precision;

/* Force selection for sfh_force / sfh_leapfrog (Fermions.cpp:640-659 full,
 * :1114-1134 fermion only, :1278-1294 gauge only). */
#define SFH_FORCE_GAUGE 1u
#define SFH_FORCE_FERMION 2u

typedef struct sfh_ctx sfh_ctx;
typedef struct sfh_gauge sfh_gauge;
typedef struct sfh_field sfh_field;

typedef struct {
  double mass;           /* fermion_mass, Parameters.h:18 (runtime here)      */
  double tol;            /* relative residual, Parameters.h:23                */
  int max_iter;          /* Parameters.h:24                                   */
  int precision;         /* sfh_precision                                     */
  double reliable_delta; /* mixed precision: reliable update due when |r|
                            drops by this factor since the last one, run at
                            the end of that chunk (0 -> 0.1 for F16, 0.003
                            for F32 and the batched solves)                   */
  int compute_true_residual; /* F64: 1 also report the FP64 true residual; 2
                                also stop on it (once the iterated residual
                                meets tol, r = phi - D D^dag x is formed and
                                the recurrence restarts from it, p = r, while
                                it misses tol).  Mixed precision always stops
                                on the FP64 true residual.                    */
} sfh_cg_params;

typedef struct {
  int iterations;
  int reliable_updates;
  int status;
  int _pad;
  double rsq;             /* final iterated |r|^2                              */
  double true_rel_resid;  /* ||phi - DDdag x|| / ||phi|| in FP64 (if computed) */
  double seconds;         /* device time of the solve                          */
  double hbm_bytes;       /* algorithmic HBM bytes of the solve                */
} sfh_cg_stats;

typedef struct {
  double mass, tol;
  int max_iter, precision;
  double beta, c11, c12; /* Parameters.h:51-52; Wilson is c11=1, c12=0        */
  double eps;            /* MD step, Parameters.h:21                          */
  int nsteps;            /* number_of_steps_fermion, Parameters.h:22          */
  unsigned which;        /* SFH_FORCE_GAUGE | SFH_FORCE_FERMION              */
  int check_hermitian;   /* 1: production guard of Fermions.cpp:712-716       */
  int _pad;
} sfh_md_params;

/* ---------------------------------------------------------------- context */
int sfh_abi_version(void);
const char* sfh_status_string(int status);
int sfh_create(int device, sfh_ctx** out);
int sfh_destroy(sfh_ctx* ctx);
const char* sfh_last_error(const sfh_ctx* ctx);
int sfh_synchronize(sfh_ctx* ctx);
/* Kernel timing with HIP events on the context stream, by kernel family:
 * "hop_f64", "hop_dot_f64" (the hop with the fused <p,v> reduction), and the
 * same for f32 / f16, "cg_blas", "force", "link_update".  on = 0 stops
 * timing, on = 1 times every launch, on = k > 1 times one launch in k per
 * family.  Setting it resets the totals.  No synchronisation is added to the
 * timed launches. */
int sfh_set_profiling(sfh_ctx* ctx, int on);
int sfh_kernel_time(sfh_ctx* ctx, const char* name, double* total_ms, long* launches);
/* on = 1 (the default): the CG solvers replay each chunk of iterations as a
 * HIP graph; on = 0: they launch the chunk's kernels directly (the same
 * kernels and results), e.g. under a kernel tracer.  No counterpart in the
 * reference. */
int sfh_set_graph_replay(sfh_ctx* ctx, int on);

/* Hop kernel form, per precision (0 f64, 1 f32, 2 f16): 0 = one lane per
 * site with all four directions' loads in flight, 1 = four lanes per site
 * (one per direction), 2 / 3 = one lane per site stepping through the
 * directions one / two at a time (fewer registers, more waves), -1 = the
 * default form.  Process-wide.  No counterpart in the reference: a tuning
 * knob of this build. */
int sfh_set_hop_variant(int precision, int variant);
/* Times each hop form on g's lattice (n launches each, HIP events on the
 * context stream), fixes the faster one for that precision and returns it
 * with its average launch time in microseconds.  No counterpart in the
 * reference. */
int sfh_tune_hop(sfh_ctx* ctx, const sfh_gauge* g, int precision, int n, int* variant, double* us);
/* Average duration in microseconds of the hop kernel the solver uses for
 * this lattice and precision (with the <p,v> reduction if with_dot),
 * n back-to-back launches between HIP events on the context stream.  No
 * counterpart in the reference. */
int sfh_time_hop(sfh_ctx* ctx, const sfh_gauge* g, int precision, int with_dot, int n, double* us);
/* Average duration in microseconds of the force kernel as the leapfrog
 * runs it on g's links (force selection `which`, Symanzik coefficients, the
 * momentum kick P -= eps F fused, no force field stored), n back-to-back
 * launches between HIP events on the context stream.  No counterpart in the
 * reference. */
int sfh_time_force(sfh_ctx* ctx, const sfh_gauge* g, double beta, double c11, double c12, unsigned which, int n,
                   double* us);
/* Form of the mixed-precision CG iteration, per precision (1 f32, 2 f16):
 * 0 = fused (the first hop recomputes the vector updates at its neighbours,
 * two kernels per iteration), 1 = split (in-place update kernel, then the
 * plain hop: three kernels), -1 = choose by lattice size (the default).
 * Process-wide.  No counterpart in the reference: a tuning knob. */
int sfh_set_cg_form(int precision, int form);
/* Times both forms on g's lattice (sfh_time_cg_kernels, n repetitions),
 * fixes the faster one for that precision and returns it with its time per
 * iteration in microseconds.  No counterpart in the reference. */
int sfh_tune_cg_form(sfh_ctx* ctx, const sfh_gauge* g, int precision, int n, int* form, double* us);
/* Average durations in microseconds of the two steps of one mixed-precision
 * CG iteration (precision SFH_PREC_F32 or SFH_PREC_F16) as the solver runs
 * them for this lattice: us_a = the update and first hop (one kernel, or the
 * in-place update and the plain hop on large lattices), us_b = the second hop
 * with its four dot products.  n back-to-back repetitions of each between
 * HIP events on the context stream.  No counterpart in the reference. */
int sfh_time_cg_kernels(sfh_ctx* ctx, const sfh_gauge* g, int precision, int n, double* us_a, double* us_b);

/* ---------------------------------------------------------------- gauge field
 * Device-resident links for one lattice.  dims = {L0, L1, L2, L3}, all even,
 * L3 = time.  Replaces Link_array (Lattice.h:11) held in host memory. */
int sfh_gauge_create(sfh_ctx* ctx, const int dims[4], sfh_gauge** out);
int sfh_gauge_destroy(sfh_ctx* ctx, sfh_gauge* g);
int sfh_gauge_upload(sfh_ctx* ctx, sfh_gauge* g, const double* links);
int sfh_gauge_download(sfh_ctx* ctx, const sfh_gauge* g, double* links);
/* Checkpoint of device-resident links in the reference's byte format: the n
 * link arrays, each V*4*18 little-endian doubles in the reference AoS order,
 * written back to back (save_all_lattice_configs / load_all_lattice_configs,
 * Lattice.cpp:810-847, with CONFIG = n).  Load fails with SFH_E_INVALID if
 * the file is shorter than n arrays. */
int sfh_gauge_save(sfh_ctx* ctx, const sfh_gauge* const* gauges, int n, const char* path);
int sfh_gauge_load(sfh_ctx* ctx, sfh_gauge* const* gauges, int n, const char* path);
/* Seeded links U = exp(i eps H), H = sum_a r_a lambda_a/2, r_a ~ N(0,1)
 * (Philox4x32-10 + Box-Muller, counter = link index), generated on the host
 * so the same seed gives the same bytes everywhere.  eps = 0 gives a cold
 * start (Lattice.cpp:172-186). */
int sfh_random_links(const int dims[4], double eps, unsigned long long seed, double* links);
/* Seeded complex Gaussian field, sigma per real part (Fermions.cpp:207-212
 * uses 1/sqrt(2)), and seeded momenta P = sum_a p_a lambda_a/2
 * (Fermions.cpp:342-360). */
int sfh_random_spinor(const int dims[4], double sigma, unsigned long long seed, double* out);
int sfh_random_momenta(const int dims[4], unsigned long long seed, double* out);

/* ---------------------------------------------------------------- spinor fields
 * Device-resident spinor for the lattice of g (6V doubles). */
int sfh_field_create(sfh_ctx* ctx, const sfh_gauge* g, sfh_field** out);
int sfh_field_destroy(sfh_ctx* ctx, sfh_field* f);
int sfh_field_upload(sfh_ctx* ctx, sfh_field* f, const double* host);
int sfh_field_download(sfh_ctx* ctx, const sfh_field* f, double* host);

/* ---------------------------------------------------------------- fermion ops */
/* out = D in (dagger=0) or D^dagger in (dagger=1), host buffers.
 * Replaces the per-site loops of apply_dirac_operator_at_site_s /
 * apply_dirac_dagger_at_site_s (Fermions.cpp:218-288) as used by
 * generate_phi_field (:302-314) and solve_Y_field (:623-633). */
int sfh_apply_d(sfh_ctx* ctx, const sfh_gauge* g, double mass, int dagger, const double* in, double* out);
/* out = D in or D^dagger in applied in a storage precision of the mixed CG's
 * inner iteration: SFH_PREC_F32 (FP32 spinors and links) or SFH_PREC_F16
 * (FP16 records with per-site norms, FP16 links; FP32 arithmetic), the same
 * kernels the inner iterations run, with in converted from FP64 and out
 * converted back; SFH_PREC_F64 is sfh_apply_d.  No counterpart in the
 * reference (which works in FP64 only): it exposes the low-precision
 * operator for checks against the FP64 one and for callers that want it. */
int sfh_apply_d_prec(sfh_ctx* ctx, const sfh_gauge* g, double mass, int dagger, int precision, const double* in,
                     double* out);
/* out = D (D^dagger in).  Replaces apply_DDdagger (Fermions.cpp:530-547). */
int sfh_apply_ddag(sfh_ctx* ctx, const sfh_gauge* g, double mass, const double* in, double* out);
/* Device-resident variants of the two above. */
int sfh_apply_d_field(sfh_ctx* ctx, const sfh_gauge* g, double mass, int dagger, const sfh_field* in,
                      sfh_field* out);
/* Re<a,b> over n doubles, FP64 accumulation, fixed reduction order.
 * Replaces field_inner_product_real (Fermions.cpp:521-528). */
int sfh_inner(sfh_ctx* ctx, size_t n, const double* a, const double* b, double* out);

/* Solve (D D^dagger) x = phi.  Replaces solve_CG_DDdagger / solve_X_field
 * (Fermions.cpp:553-621).  stats may be NULL. */
int sfh_cg(sfh_ctx* ctx, const sfh_gauge* g, const sfh_cg_params* p, const double* phi, double* x,
           sfh_cg_stats* stats);
int sfh_cg_field(sfh_ctx* ctx, const sfh_gauge* g, const sfh_cg_params* p, const sfh_field* phi,
                 sfh_field* x, sfh_cg_stats* stats);

/* Batched solves: many independent systems (each its own lattice, mass and
 * 1-8 right-hand sides sharing its links) advanced together, one launch per
 * CG phase for the whole batch, chunks of iterations replayed as a HIP graph.
 * Each system's result does not depend on what else is in the batch.
 * precision: SFH_PREC_F32 (FP32 inner + FP64 reliable updates, stops on the
 * FP64 true residual) or SFH_PREC_F64 (the reference recurrence).  stats
 * (may be NULL) receives one entry per right-hand side, systems in order.
 * No counterpart in the reference (it solves one system at a time,
 * Fermions.cpp:553-617). */
typedef struct {
  const sfh_gauge* gauge;
  double mass;
  int nrhs;
  int _pad;
  const sfh_field* const* rhs; /* nrhs device fields on gauge's lattice */
  sfh_field* const* x;         /* nrhs device fields receiving the solutions */
} sfh_batch_system;
int sfh_cg_batch(sfh_ctx* ctx, int nsys, const sfh_batch_system* systems, const sfh_cg_params* p,
                 sfh_cg_stats* stats);
/* Autotuner of the batched engine: times the batched hop on this batch for
 * every launch shape (128, 256 or 512 sites per block; 1, 2, 4 or 8
 * right-hand sides per block) and records the fastest in the context for
 * later sfh_cg_batch calls on batches of the same shape (lattices and
 * right-hand-side counts, in order, and precision).  If *hop_threads and
 * *rhs_tile are nonzero on input, that shape is recorded instead (and timed).
 * The shape changes no solve's arithmetic: results are bit-identical tuned
 * or not.  The x fields are used as work space (left zeroed).  No
 * counterpart in the reference. */
int sfh_tune_batch(sfh_ctx* ctx, int nsys, const sfh_batch_system* systems, const sfh_cg_params* p,
                   int* hop_threads, int* rhs_tile, double* us);

/* ---------------------------------------------------------------- solve service
 * Continuous-batching CG service for concurrent HMC chains (mixed precision:
 * FP32 inner, FP64 reliable updates, stop on the FP64 true residual).
 * A chain registers its device-resident links once, with a cap of up to 8
 * solves in flight; a chain one solve of which exceeds mem_budget
 * (0 = unlimited) is refused with SFH_E_NOMEM.  Submitted requests queue
 * FIFO per chain and are admitted by memory as they arrive: into a free
 * slot of the chain, else into a new slot while the chain is below its cap,
 * the service's slot memory stays within mem_budget and free HBM
 * (hipMemGetInfo at that moment) keeps a headroom of max(256 MiB, 1/64 of
 * the device); otherwise the request waits for a slot to free up (a
 * request of a chain with no slot, while nothing runs, completes with
 * SFH_E_NOMEM).  A worker thread advances every solve in flight together, a
 * chunk of iterations at a time, so requests that arrive while others run
 * join at the next chunk boundary.  Each result equals the solve alone
 * (bit-identical).  stats.seconds is the latency from submit to completion.
 * The links of a chain must not change while it has requests queued or in
 * flight; after they change, sfh_service_update_chain refreshes the chain.
 * No counterpart in the reference. */
typedef struct sfh_service sfh_service;
int sfh_service_create(int device, size_t mem_budget, double tol, int max_iter, sfh_service** out);
int sfh_service_add_chain(sfh_service* s, const sfh_gauge* g, double mass, int max_inflight, int* chain);
/* New links for a chain with no request queued or in flight (g: the same
 * lattice shape, e.g. the chain's own gauge after a link update) and its
 * mass: the SU(3) check and the FP32 link copy are redone.  SFH_E_INVALID if
 * requests are pending. */
int sfh_service_update_chain(sfh_service* s, int chain, const sfh_gauge* g, double mass);
/* Drops a chain with no request queued or in flight; its solve slots stay
 * allocated for other chains' requests.  The id is not reused. */
int sfh_service_remove_chain(sfh_service* s, int chain);
/* Device memory held by solve slots, and how many times a new slot was
 * refused for memory (the request then waited). */
int sfh_service_memory(sfh_service* s, size_t* slot_bytes, long* refusals);
int sfh_service_submit(sfh_service* s, int chain, const sfh_field* rhs, sfh_field* x, long long* ticket);
/* Waits for one ticket (returns its status; SFH_E_HIP on timeout). */
int sfh_service_wait(sfh_service* s, long long ticket, double timeout_s, sfh_cg_stats* stats);
/* Collects up to max completed requests (waiting up to timeout_s for one). */
int sfh_service_poll(sfh_service* s, double timeout_s, int max, long long* tickets, sfh_cg_stats* stats, int* n);
const char* sfh_service_last_error(const sfh_service* s);
int sfh_service_destroy(sfh_service* s);

/* ---------------------------------------------------------------- forces, MD */
/* F = F_gauge(beta, c11, c12) + F_fermion(X, Y) per link, selected by which.
 * Replaces compute_full_force_field (Fermions.cpp:640-659) and the
 * fermion-only / gauge-only fields (:1114-1134, :1278-1294).  X and Y may be
 * NULL when which excludes the fermion force.  The force (and every MD call
 * that evaluates it) takes lattice extents up to 255 (SFH_E_INVALID above). */
int sfh_force(sfh_ctx* ctx, const sfh_gauge* g, double beta, double c11, double c12, const double* X,
              const double* Y, unsigned which, double* F);
/* Symanzik gauge action summed per link.  Replaces compute_action
 * (Lattice.cpp:759-781). */
int sfh_action(sfh_ctx* ctx, const sfh_gauge* g, double beta, double c11, double c12, double* out);
/* Replaces compute_avg_plaq (Lattice.cpp:784-804), same normalisation. */
int sfh_avg_plaq(sfh_ctx* ctx, const sfh_gauge* g, double* out);
/* Polyakov loop trace Tr prod_t U_3(x,y,z,t) at every spatial site, complex
 * (re, im) pairs at [(x*L1 + y)*L2 + z].  Replaces precompute_polyakov_grid /
 * poly_loop_at_spat_coord (Lattice.cpp:849-884); the correlators of
 * Lattice.cpp:886-948 are sums over this grid. */
int sfh_polyakov_grid(sfh_ctx* ctx, const sfh_gauge* g, double* out);
/* Polyakov-loop correlators of the device-resident links:
 *   C(r^2) = (1/N) sum_m sum_{d in D(r^2)} P(m) conj(P(m + d)),
 * P the grid of sfh_polyakov_grid (computed on the device, never copied to
 * the host), D(r^2) the integer spatial displacements with d.d = r^2 and
 * |d_k| <= min(floor(sqrt(r^2)), L_k / 2), d_k = +L_k / 2 left out for even
 * L_k, N = (L0 L1 L2) |D(r^2)|, and C = 0 when D(r^2) is empty.
 * out[2k], out[2k+1] = Re, Im C(r2_first + k) for k < count.  Replaces the
 * measurement step precompute_polyakov_grid + correlator_over_fixed_distance_fast
 * (Lattice.cpp:869-948) as Main.cpp:219-228 calls it (r2_first = 1,
 * count = maxdistance). */
int sfh_polyakov_correlators(sfh_ctx* ctx, const sfh_gauge* g, int r2_first, int count, double* out);
/* The same correlators of a host grid (complex (re, im) pairs at
 * [(x*L1 + y)*L2 + z], spatial = {L0, L1, L2}): replaces
 * correlator_over_fixed_distance_fast (Lattice.cpp:886-948) on the grid
 * that precompute_polyakov_grid returns. */
int sfh_polyakov_correlators_grid(sfh_ctx* ctx, const int spatial[3], const double* grid, int r2_first, int count,
                                  double* out);
/* sfh_polyakov_correlators without any device allocation.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its work buffers are one static device array of the library
 * per GPU (9.3 MiB), so it serves lattices of at most 2^18 spatial sites
 * (L0 L1 L2), at most 256 values of r^2 per call and at most 65536
 * displacements in total; larger requests fail with SFH_E_INVALID.  Calls
 * on one GPU share that array and are serialised by a per-device lock in
 * this process; results equal sfh_polyakov_correlators' bit for bit. */
int sfh_polyakov_correlators_noalloc(sfh_ctx* ctx, const sfh_gauge* g, int r2_first, int count, double* out);
/* Host link arrays (n of them, doubles_each doubles each) written back to
 * back as raw little-endian doubles, and read back: the byte format of
 * save_all_lattice_configs / load_all_lattice_configs (Lattice.cpp:810-847),
 * the same as sfh_gauge_save / sfh_gauge_load.  SFH_E_IO where the reference
 * returns false (the file cannot be opened, or is shorter than n arrays). */
int sfh_links_save(const double* const* arrays, int n, size_t doubles_each, const char* path);
int sfh_links_load(double* const* arrays, int n, size_t doubles_each, const char* path);
/* nsweeps Cabibbo-Marinari heatbath sweeps of the Symanzik gauge action
 * (beta, c11, c12) on the device-resident links.  Replaces heatbath_update
 * (Lattice.cpp:725-757) with single_link_heatbath (:696-723), the subgroup
 * updates R, S, T (:570-694) and the SU(2) sampler (SU3_Sampling.cpp:17-199).
 * Every extent must be a multiple of 4 (at most 252) and beta > 0.  The sweep visits the
 * links direction by direction, each direction in four colour classes
 * (2 n_mu + sum of the other coordinates) mod 4 whose links are updated in
 * parallel; the reference visits sites lexicographically.  Uniforms are
 * Philox4x32-10 draws addressed by (seed; draw, link, sweep index), the
 * sweeps of one call taking indices sweep0, sweep0 + 1, ...: a result
 * depends on the seed and sweep indices only.  avg_plaq[k] and
 * action_density[k] (either may be NULL) receive sweep k's
 * (sum of ReTr(U A1)/3 / 4) / (6 V) and (sum of -(beta/3) ReTr(U A) / 4) / V,
 * each link's term taken right after its update, as the reference returns
 * them. */
int sfh_heatbath(sfh_ctx* ctx, sfh_gauge* g, double beta, double c11, double c12, unsigned long long seed,
                 unsigned sweep0, int nsweeps, double* avg_plaq, double* action_density);
/* nsteps HEX (hypercubic stout) smearing steps of the device-resident links
 * in place.  Replaces Apply_HEX_smearing (Smearing.cpp:304-311) and, with
 * nsteps = 4, Apply_4HEX_smearing (:313-318); the history variant (:320-331)
 * is nsteps = 1 four times with a download in between.  alpha = {alpha1,
 * alpha2, alpha3}, the reference's 0.95, 0.76, 0.38 (Smearing.h:7-9); each
 * level maps V = exp(proj(alpha/2 Gamma U^dagger)) U as the reference writes
 * it, proj the traceless anti-hermitian part.  The reference hands that
 * anti-hermitian matrix to its exp(iQ) routine, which rejects it on any
 * non-trivial links; here the exponential is exp(A) = exp(iQ), Q = -iA, and
 * the level-3 staple sum starts at zero.  Non-finite links fail the call
 * with SFH_E_INVALID (the links are then partly smeared). */
int sfh_hex_smear(sfh_ctx* ctx, sfh_gauge* g, const double alpha[3], int nsteps);
/* sum over links of ReTr(P^2).  Replaces compute_Tr_P2 (Fermions.cpp:662-678). */
int sfh_tr_p2(sfh_ctx* ctx, const sfh_gauge* g, const double* P, double* out);
/* U <- exp(i eps P) U on every device link (Fermions.cpp:703-723, exponential
 * :843-902).  check_hermitian=1 applies the reference's guard. */
int sfh_link_update(sfh_ctx* ctx, sfh_gauge* g, double eps, const double* P, int check_hermitian);
/* exp(iQ) of one 3x3 matrix on the device (Fermions.cpp:843-902). */
int sfh_exp_iq(sfh_ctx* ctx, const double* Q18, double* out18);

/* Leapfrog with the links kept on the device: P half-kick, (nsteps-1) x
 * {link update, kick}, link update, half-kick.  Replaces
 * apply_initial_step_to_conjugate_field, apply_intermediate_update_... and
 * apply_final_step (Fermions.cpp:684-776) and their eps/n_steps copies
 * (:967-1028).  phi may be NULL when which excludes the fermion force;
 * X_out (may be NULL) receives the final solve. */
int sfh_leapfrog(sfh_ctx* ctx, sfh_gauge* g, const sfh_md_params* p, const double* phi, double* P_inout,
                 double* X_out);
/* One HMC update (Fermions.cpp:778-828): phi = D chi, H_i, leapfrog, H_f,
 * Metropolis with the caller's uniform r.  H[0..6] = TrP2_i, Sg_i, Sf_i,
 * TrP2_f, Sg_f, Sf_f, H_i - H_f.  On accept the device links take the end
 * point; otherwise they are left unchanged. */
int sfh_hmc_trajectory(sfh_ctx* ctx, sfh_gauge* g, const sfh_md_params* p, const double* chi, const double* P0,
                       double r_uniform, double* H, int* accepted);

#ifdef __cplusplus
}
#endif

#endif /* SFH_H */
