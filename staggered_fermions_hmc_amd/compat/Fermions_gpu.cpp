// Drop-in field-level Fermions / Lattice functions on the MI355X library.
// See Fermions_gpu.h for the contract; each function cites the reference
// function it replaces.
#include "Fermions_gpu.h"

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>

#include "../../include/sfh.h"

// OpenMP thread number when the caller runs OpenMP (resolved at load time if
// libgomp is in the process; null otherwise).
extern "C" int omp_get_thread_num(void) __attribute__((weak));

namespace sfh_compat {

namespace {

Config g_cfg;
unsigned long long g_generation = 1;  // bumps when configure() changes the lattice
std::atomic<int> g_threads{0};        // threads registered outside OpenMP

// splitmix64 finaliser: decorrelates the (seed, stream, draw) words.
unsigned long long mix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Content fingerprint of a host link array: 64-bit words folded in four
// independent multiply-xorshift lanes per 1 MiB block, blocks hashed on
// several threads and combined in order.  Reading the array is several times
// cheaper than the host-to-device copy of it that a match saves.
unsigned long long fingerprint(const double* a, size_t n) {
  const uint64_t* w = reinterpret_cast<const uint64_t*>(a);
  constexpr size_t kBlock = 1 << 17;  // words per block (1 MiB)
  const size_t nb = (n + kBlock - 1) / kBlock;
  std::vector<uint64_t> bh(nb);
  auto work = [&](size_t b0, size_t b1) {
    for (size_t b = b0; b < b1; ++b) {
      uint64_t h[4] = {b, b ^ 0x1234567ULL, b ^ 0x89abcdefULL, ~b};
      const size_t lo = b * kBlock, hi = std::min(n, lo + kBlock);
      size_t i = lo;
      for (; i + 4 <= hi; i += 4)
        for (int k = 0; k < 4; ++k) h[k] = (h[k] ^ w[i + k]) * 0x100000001b3ULL + (h[k] >> 29);
      for (; i < hi; ++i) h[0] = (h[0] ^ w[i]) * 0x100000001b3ULL + (h[0] >> 29);
      bh[b] = mix64(h[0] ^ mix64(h[1] ^ mix64(h[2] ^ mix64(h[3]))));
    }
  };
  const size_t nt = std::min<size_t>(nb, std::max(1u, std::min(8u, std::thread::hardware_concurrency())));
  if (nt <= 1) {
    work(0, nb);
  } else {
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; ++t) th.emplace_back(work, nb * t / nt, nb * (t + 1) / nt);
    for (auto& t : th) t.join();
  }
  uint64_t h = mix64(n);
  for (uint64_t x : bh) h = mix64(h ^ x);
  return h;
}

struct ThreadState {
  sfh_ctx* ctx = nullptr;
  sfh_gauge* gauge = nullptr;
  unsigned long long generation = 0;
  int stream = -1;               // stable stream index of this thread
  unsigned long long draws = 0;  // Philox stream position of this thread
  unsigned hb_sweeps = 0;        // heatbath sweep index of this thread
  bool uniform_seeded = false;
  std::mt19937_64 uniform;
  // what the device links hold: the fingerprint of the host array they were
  // last uploaded from or downloaded to (valid only with `resident`)
  bool resident = false;
  unsigned long long links_fp = 0;
  ~ThreadState() {
    if (gauge) sfh_gauge_destroy(ctx, gauge);
    if (ctx) sfh_destroy(ctx);
  }
};

thread_local ThreadState ts;

// The thread's stream index: its OpenMP thread number when called from an
// OpenMP region (the reference seeds one generator per OpenMP thread,
// rng.cpp:8-16), else the order in which threads first drew, or the value
// set with set_thread_stream().
int stream_index() {
  if (ts.stream < 0) ts.stream = omp_get_thread_num ? omp_get_thread_num() : g_threads.fetch_add(1);
  return ts.stream;
}

[[noreturn]] void fail(int status, const char* what) {
  std::string msg = ts.ctx ? sfh_last_error(ts.ctx) : "";
  if (msg.empty()) msg = std::string(what) + ": " + sfh_status_string(status);
  if (status == SFH_E_EXP_QCUBED || status == SFH_E_EXP_NEG_TRACE || status == SFH_E_INVALID)
    throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

void check(int status, const char* what) {
  if (status != SFH_OK) fail(status, what);
}

int dims_[4];

const int* dims() {
  dims_[0] = dims_[1] = dims_[2] = g_cfg.Spatial_Size;
  dims_[3] = g_cfg.temporal_size;
  return dims_;
}

sfh_ctx* ctx() {
  if (!ts.ctx) check(sfh_create(g_cfg.device, &ts.ctx), "sfh_create");
  return ts.ctx;
}

// Device links for this thread's lattice, loaded from the caller's array.
sfh_gauge* gauge_from(const Link_array& arr) {
  sfh_ctx* c = ctx();
  if (ts.gauge && ts.generation != g_generation) {
    sfh_gauge_destroy(c, ts.gauge);
    ts.gauge = nullptr;
  }
  if (!ts.gauge) {
    check(sfh_gauge_create(c, dims(), &ts.gauge), "sfh_gauge_create");
    ts.generation = g_generation;
  }
  if ((int)arr.size() != volume() * 72) throw std::invalid_argument("link array has the wrong size");
  // consecutive calls on an unchanged array find its links already on the device
  const unsigned long long fp = fingerprint(arr.data(), arr.size());
  if (ts.resident && ts.links_fp == fp) return ts.gauge;
  ts.resident = false;
  check(sfh_gauge_upload(c, ts.gauge, arr.data()), "sfh_gauge_upload");
  ts.resident = true;
  ts.links_fp = fp;
  return ts.gauge;
}

// The device links were changed by the library: they are copied into arr,
// which the device copy then matches.
void links_to_host(sfh_gauge* g, Link_array& arr) {
  ts.resident = false;
  check(sfh_gauge_download(ctx(), g, arr.data()), "sfh_gauge_download");
  ts.links_fp = fingerprint(arr.data(), arr.size());
  ts.resident = true;
}

sfh_cg_params cg_params() {
  sfh_cg_params p{};
  p.mass = g_cfg.fermion_mass;
  p.tol = g_cfg.tol;
  p.max_iter = g_cfg.max_iter;
  p.precision = g_cfg.precision;
  p.reliable_delta = 0.0;  // the library default for the precision
  return p;
}

sfh_md_params md_params(double beta, double eps, int nsteps) {
  sfh_md_params p{};
  p.mass = g_cfg.fermion_mass;
  p.tol = g_cfg.tol;
  p.max_iter = g_cfg.max_iter;
  p.precision = g_cfg.precision;
  p.beta = beta;
  p.c11 = g_cfg.c_11;
  p.c12 = g_cfg.c_12;
  p.eps = eps;
  p.nsteps = nsteps;
  p.which = SFH_FORCE_GAUGE | SFH_FORCE_FERMION;
  p.check_hermitian = 1;
  return p;
}

unsigned long long next_seed() {
  // distinct Philox key per (configured seed, thread stream, draw): the same
  // seed replays the same fields in every process
  return mix64(g_cfg.seed ^ mix64((unsigned long long)stream_index() ^ mix64(++ts.draws)));
}

std::vector<double> spinor() { return std::vector<double>((size_t)volume() * 6, 0.0); }
std::vector<double> links() { return std::vector<double>((size_t)volume() * 72, 0.0); }

void check_spinor(const std::vector<double>& v) {
  if ((int)v.size() != volume() * 6) throw std::invalid_argument("spinor field has the wrong size");
}

void kick(std::vector<double>& P, double coef, const std::vector<double>& F) {
  for (size_t i = 0; i < P.size(); ++i) P[i] -= coef * F[i];
}

}  // namespace

void configure(const Config& c) {
  g_cfg = c;
  ++g_generation;
}
void set_thread_stream(int k) {
  ts.stream = k;
  ts.draws = 0;
  ts.hb_sweeps = 0;
  ts.uniform_seeded = false;
}
const Config& config() { return g_cfg; }
int volume() { return g_cfg.Spatial_Size * g_cfg.Spatial_Size * g_cfg.Spatial_Size * g_cfg.temporal_size; }

}  // namespace sfh_compat

using namespace sfh_compat;

// Fermions.cpp:302-314: phi = D chi, chi complex Gaussian (sigma 1/sqrt 2).
std::vector<double> generate_phi_field(const Link_array& arr) {
  sfh_gauge* g = gauge_from(arr);
  std::vector<double> chi = spinor(), phi = spinor();
  check(sfh_random_spinor(dims(), std::sqrt(0.5), next_seed(), chi.data()), "generate_chi_field");
  check(sfh_apply_d(ctx(), g, config().fermion_mass, 0, chi.data(), phi.data()), "generate_phi_field");
  return phi;
}

// Fermions.cpp:342-360.
std::vector<double> generate_conjugate_field_configuration() {
  std::vector<double> P = links();
  check(sfh_random_momenta(dims(), next_seed(), P.data()), "generate_conjugate_field_configuration");
  return P;
}

// Fermions.cpp:521-528.
double field_inner_product_real(const std::vector<double>& A, const std::vector<double>& B) {
  if (A.size() != B.size()) throw std::invalid_argument("inner product of fields of different sizes");
  double out = 0.0;
  check(sfh_inner(ctx(), A.size(), A.data(), B.data(), &out), "field_inner_product_real");
  return out;
}

// Fermions.cpp:530-547.
void apply_DDdagger(const std::vector<double>& in_field, const Link_array& arr, std::vector<double>& out_field) {
  check_spinor(in_field);
  check_spinor(out_field);
  sfh_gauge* g = gauge_from(arr);
  check(sfh_apply_ddag(ctx(), g, config().fermion_mass, in_field.data(), out_field.data()), "apply_DDdagger");
}

// Fermions.cpp:553-617.
std::vector<double> solve_CG_DDdagger(const std::vector<double>& phi_field, const Link_array& arr) {
  check_spinor(phi_field);
  sfh_gauge* g = gauge_from(arr);
  std::vector<double> X = spinor();
  const sfh_cg_params p = cg_params();
  check(sfh_cg(ctx(), g, &p, phi_field.data(), X.data(), nullptr), "solve_CG_DDdagger");
  return X;
}

// Fermions.cpp:619-621.
std::vector<double> solve_X_field(const std::vector<double>& phi_field, const Link_array& arr) {
  return solve_CG_DDdagger(phi_field, arr);
}

// Fermions.cpp:623-633.
std::vector<double> solve_Y_field(const std::vector<double>& X_field, const Link_array& arr) {
  check_spinor(X_field);
  sfh_gauge* g = gauge_from(arr);
  std::vector<double> Y = spinor();
  check(sfh_apply_d(ctx(), g, config().fermion_mass, 1, X_field.data(), Y.data()), "solve_Y_field");
  return Y;
}

static std::vector<double> force(const std::vector<double>* X, const std::vector<double>* Y, const Link_array& arr,
                                 double beta, unsigned which) {
  sfh_gauge* g = gauge_from(arr);
  std::vector<double> F = links();
  check(sfh_force(ctx(), g, beta, config().c_11, config().c_12, X ? X->data() : nullptr, Y ? Y->data() : nullptr,
                  which, F.data()),
        "force");
  return F;
}

// Fermions.cpp:640-659.
std::vector<double> compute_full_force_field(const std::vector<double>& X_field, const std::vector<double>& Y_field,
                                             const Link_array& arr, double beta) {
  check_spinor(X_field);
  check_spinor(Y_field);
  return force(&X_field, &Y_field, arr, beta, SFH_FORCE_GAUGE | SFH_FORCE_FERMION);
}

// Fermions.cpp:1114-1134.
std::vector<double> compute_fermion_only_force_field(const std::vector<double>& X_field,
                                                     const std::vector<double>& Y_field, const Link_array& arr) {
  check_spinor(X_field);
  check_spinor(Y_field);
  return force(&X_field, &Y_field, arr, 0.0, SFH_FORCE_FERMION);
}

// Fermions.cpp:1278-1294.
std::vector<double> compute_gauge_only_force_field(const Link_array& arr, double beta) {
  return force(nullptr, nullptr, arr, beta, SFH_FORCE_GAUGE);
}

// Fermions.cpp:662-678.
double compute_Tr_P2(const std::vector<double>& P_field) {
  if ((int)P_field.size() != volume() * 72) throw std::invalid_argument("momentum field has the wrong size");
  sfh_ctx* c = ctx();
  if (!ts.gauge || ts.generation != g_generation) {
    if (ts.gauge) sfh_gauge_destroy(c, ts.gauge);
    check(sfh_gauge_create(c, dims(), &ts.gauge), "sfh_gauge_create");
    ts.generation = g_generation;
  }
  double out = 0.0;
  check(sfh_tr_p2(c, ts.gauge, P_field.data(), &out), "compute_Tr_P2");
  return out;
}

// U <- exp(i eps P) U on every link, with the hermiticity guard
// (Fermions.cpp:703-723); updates arr in place.
static void link_step(Link_array& arr, const std::vector<double>& P, double eps, int check_h) {
  sfh_gauge* g = gauge_from(arr);
  ts.resident = false;
  check(sfh_link_update(ctx(), g, eps, P.data(), check_h), "link update");
  links_to_host(g, arr);
}

// Fermions.cpp:684-697 and the eps copy :967-976.
void local_initial_half_step(std::vector<double>& P_field, double beta, const std::vector<double>& phi_field,
                             const Link_array& arr, double eps) {
  const std::vector<double> X = solve_X_field(phi_field, arr);
  const std::vector<double> Y = solve_Y_field(X, arr);
  kick(P_field, eps * 0.5, compute_full_force_field(X, Y, arr, beta));
}

void apply_initial_step_to_conjugate_field(std::vector<double>& P_field, double beta,
                                           const std::vector<double>& phi_field, const Link_array& arr) {
  local_initial_half_step(P_field, beta, phi_field, arr, config().epsilon);
}

// Fermions.cpp:699-738 and :978-1004.
static void steps(std::vector<double>& P_field, double beta, const std::vector<double>& phi_field, Link_array& arr,
                  double eps, int n_steps, int check_h) {
  for (int n = 1; n < n_steps; ++n) {
    link_step(arr, P_field, eps, check_h);
    const std::vector<double> X = solve_X_field(phi_field, arr);
    const std::vector<double> Y = solve_Y_field(X, arr);
    kick(P_field, eps, compute_full_force_field(X, Y, arr, beta));
  }
}

void local_leapfrog_steps(std::vector<double>& P_field, double beta, const std::vector<double>& phi_field,
                          Link_array& arr, double eps, int n_steps) {
  steps(P_field, beta, phi_field, arr, eps, n_steps, 0);
}

void apply_intermediate_update_to_gauge_field_and_conjugate_field(std::vector<double>& P_field, double beta,
                                                                  const std::vector<double>& phi_field,
                                                                  Link_array& arr) {
  steps(P_field, beta, phi_field, arr, config().epsilon, config().number_of_steps_fermion, 1);
}

// Fermions.cpp:740-776 and :1006-1028.
static std::vector<double> final_step(std::vector<double>& P_field, double beta, const std::vector<double>& phi_field,
                                      Link_array& arr, double eps, int check_h) {
  link_step(arr, P_field, eps, check_h);
  std::vector<double> X = solve_X_field(phi_field, arr);
  const std::vector<double> Y = solve_Y_field(X, arr);
  kick(P_field, eps * 0.5, compute_full_force_field(X, Y, arr, beta));
  return X;
}

std::vector<double> local_final_step(std::vector<double>& P_field, double beta, const std::vector<double>& phi_field,
                                     Link_array& arr, double eps) {
  return final_step(P_field, beta, phi_field, arr, eps, 0);
}

std::vector<double> apply_final_step(std::vector<double>& P_field, double beta, const std::vector<double>& phi_field,
                                     Link_array& arr) {
  return final_step(P_field, beta, phi_field, arr, config().epsilon, 1);
}

// Lattice.cpp:759-781.
double compute_action(const Link_array& arr, double beta) {
  sfh_gauge* g = gauge_from(arr);
  double out = 0.0;
  check(sfh_action(ctx(), g, beta, config().c_11, config().c_12, &out), "compute_action");
  return out;
}

// Lattice.cpp:784-804.
double compute_avg_plaq(const Link_array& arr) {
  sfh_gauge* g = gauge_from(arr);
  double out = 0.0;
  check(sfh_avg_plaq(ctx(), g, &out), "compute_avg_plaq");
  return out;
}

// Lattice.cpp:725-757.
std::pair<double, double> heatbath_update(Link_array& arr, double beta) {
  sfh_gauge* g = gauge_from(arr);
  double plaq = 0.0, act = 0.0;
  const unsigned long long key = mix64(config().seed ^ mix64(0x4842ULL + (unsigned long long)stream_index()));
  ts.resident = false;
  check(sfh_heatbath(ctx(), g, beta, config().c_11, config().c_12, key, ts.hb_sweeps++, 1, &plaq, &act),
        "heatbath_update");
  links_to_host(g, arr);
  return {plaq, act};
}

static void hex_steps(Link_array& U_array, int n) {
  sfh_gauge* g = gauge_from(U_array);
  const double alpha[3] = {config().alpha1, config().alpha2, config().alpha3};
  ts.resident = false;
  check(sfh_hex_smear(ctx(), g, alpha, n), "Apply_HEX_smearing");
  links_to_host(g, U_array);
}

// Smearing.cpp:304-311.
void Apply_HEX_smearing(Link_array& U_array) { hex_steps(U_array, 1); }

// Smearing.cpp:313-318.
void Apply_4HEX_smearing(Link_array& U_array) { hex_steps(U_array, 4); }

// Smearing.cpp:320-331: the thin links and each of the four smearing levels.
std::vector<Link_array> Apply_4HEX_smearing_with_history(const Link_array& U_thin) {
  std::vector<Link_array> history{U_thin};
  sfh_gauge* g = gauge_from(U_thin);
  const double alpha[3] = {config().alpha1, config().alpha2, config().alpha3};
  ts.resident = false;  // the device links end smeared, not equal to U_thin
  for (int step = 0; step < 4; ++step) {
    check(sfh_hex_smear(ctx(), g, alpha, 1), "Apply_4HEX_smearing_with_history");
    history.emplace_back(U_thin.size());
    check(sfh_gauge_download(ctx(), g, history.back().data()), "sfh_gauge_download");
  }
  return history;
}

// Metropolis uniforms: one mt19937_64 per thread stream, seeded from the
// configured seed (the reference's per-thread generator, rng.cpp:8-32).
static double uniform01() {
  if (!ts.uniform_seeded) {
    ts.uniform.seed(mix64(config().seed ^ mix64(0x4d43ULL + (unsigned long long)stream_index())));
    ts.uniform_seeded = true;
  }
  return std::uniform_real_distribution<double>(0.0, 1.0)(ts.uniform);
}

// Fermions.cpp:778-810.
void Monte_carlo_step(double initial_Tr_P2, double initial_phi_field_action, const std::vector<double>& P_field_final,
                      const std::vector<double>& phi_field, double gauge_field_initial_action,
                      Link_array& gauge_field_initial, Link_array& gauge_field_final,
                      const std::vector<double>& X_field, double beta) {
  const double r = uniform01();
  const double final_Tr_P2 = compute_Tr_P2(P_field_final);
  const double final_gauge_action = compute_action(gauge_field_final, beta);
  const double final_phi_field_action = field_inner_product_real(phi_field, X_field);
  const double computation = initial_Tr_P2 - final_Tr_P2 + gauge_field_initial_action - final_gauge_action +
                             initial_phi_field_action - final_phi_field_action;
  if (config().verbose) {
    std::cout << "Initial Tr(P^2): " << initial_Tr_P2 << "\nFinal Tr(P^2): " << final_Tr_P2
              << "\nInitial Gauge Action: " << gauge_field_initial_action
              << "\nFinal Gauge Action: " << final_gauge_action
              << "\nInitial Fermion Action: " << initial_phi_field_action
              << "\nFinal Fermion Action: " << final_phi_field_action << "\n"
              << computation << std::endl;
  }
  if (r < std::exp(computation)) gauge_field_initial = gauge_field_final;
}

// Fermions.cpp:812-828, run device-resident for the whole trajectory: the
// links stay in HBM from the initial action to the Metropolis test.
void single_sweep_monte_carlo_update_WITH_FERMIONS(Link_array& gauge_field, double beta) {
  sfh_gauge* g = gauge_from(gauge_field);
  std::vector<double> chi = spinor(), P = links();
  check(sfh_random_spinor(dims(), std::sqrt(0.5), next_seed(), chi.data()), "generate_chi_field");
  check(sfh_random_momenta(dims(), next_seed(), P.data()), "generate_conjugate_field_configuration");
  const sfh_md_params p = md_params(beta, config().epsilon, config().number_of_steps_fermion);
  double H[7];
  int accepted = 0;
  const double r = uniform01();
  check(sfh_hmc_trajectory(ctx(), g, &p, chi.data(), P.data(), r, H, &accepted), "trajectory");
  if (config().verbose) {
    std::cout << "Initial Tr(P^2): " << H[0] << "\nFinal Tr(P^2): " << H[3] << "\nInitial Gauge Action: " << H[1]
              << "\nFinal Gauge Action: " << H[4] << "\nInitial Fermion Action: " << H[2]
              << "\nFinal Fermion Action: " << H[5] << "\n"
              << H[6] << std::endl;
  }
  // on reject the device links are the ones uploaded, which gauge_field still holds
  if (accepted) links_to_host(g, gauge_field);
}

// Lattice.cpp:868-884: Tr prod_t U_3 at every spatial site, computed on the
// device from the (resident) links.
std::vector<complex> precompute_polyakov_grid(const Link_array& arr) {
  sfh_gauge* g = gauge_from(arr);
  const int L = config().Spatial_Size;
  std::vector<complex> grid((size_t)L * L * L);
  check(sfh_polyakov_grid(ctx(), g, reinterpret_cast<double*>(grid.data())), "precompute_polyakov_grid");
  return grid;
}

// Lattice.cpp:886-948, summed on the device.
complex correlator_over_fixed_distance_fast(const std::vector<complex>& P_grid, int r_squared) {
  const int L = config().Spatial_Size;
  if ((long)P_grid.size() != (long)L * L * L) throw std::invalid_argument("Polyakov grid has the wrong size");
  if (r_squared < 1) return complex(0.0, 0.0);
  const int spatial[3] = {L, L, L};
  double out[2];
  check(sfh_polyakov_correlators_grid(ctx(), spatial, reinterpret_cast<const double*>(P_grid.data()), r_squared, 1,
                                      out),
        "correlator_over_fixed_distance_fast");
  return complex(out[0], out[1]);
}

namespace sfh_compat {

// Lattice.cpp:810-847: the arrays back to back as raw doubles, with the
// reference's messages on std::cerr and a false return on failure.
bool save_links(const Link_array* arrays, size_t n, const std::string& filename) {
  std::FILE* f = std::fopen(filename.c_str(), "wb");
  if (!f) {
    std::cerr << "Error opening file for writing: " << filename << "\n";
    return false;
  }
  std::fclose(f);
  std::vector<const double*> p(n);
  for (size_t k = 0; k < n; ++k) p[k] = arrays[k].data();
  if (sfh_links_save(p.data(), (int)n, n ? arrays[0].size() : 0, filename.c_str()) != SFH_OK) {
    std::cerr << "Error writing complete binary data to: " << filename << "\n";
    return false;
  }
  return true;
}

bool load_links(Link_array* arrays, size_t n, const std::string& filename) {
  std::vector<double*> p(n);
  for (size_t k = 0; k < n; ++k) p[k] = arrays[k].data();
  std::FILE* f = std::fopen(filename.c_str(), "rb");
  if (!f) {
    std::cerr << "Error opening file for reading: " << filename << "\n";
    return false;
  }
  std::fclose(f);
  if (sfh_links_load(p.data(), (int)n, n ? arrays[0].size() : 0, filename.c_str()) != SFH_OK) {
    std::cerr << "Error reading complete binary data from: " << filename << "\n";
    return false;
  }
  return true;
}

}  // namespace sfh_compat
