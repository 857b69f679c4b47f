// Reference-style checks of the drop-in layer, run through the same
// field-level calls the reference's own checks use (ai_check.cpp:172-333,
// Fermions.cpp:908-955, :1061-1108).  Prints PASS/FAIL per check and exits
// non-zero if any fails.  Needs a GPU.
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sfh.h"
#include "Fermions_gpu.h"

// the reference's checkpoint signatures (Lattice.h:43-44), CONFIG = 2
bool save_all_lattice_configs(const std::array<Link_array, 2>& links_at_coupling, const std::string& filename);
bool load_all_lattice_configs(std::array<Link_array, 2>& links_at_coupling, const std::string& filename);

static int failures = 0;

static void report(const char* name, bool ok, double value) {
  std::printf("%-48s %s  (%.3e)\n", name, ok ? "PASS" : "FAIL", value);
  if (!ok) ++failures;
}

static Link_array cold() {
  Link_array U((size_t)sfh_compat::volume() * 72, 0.0);
  for (size_t l = 0; l < U.size() / 18; ++l) U[18 * l] = U[18 * l + 8] = U[18 * l + 16] = 1.0;
  return U;
}

static Link_array hot(double eps, unsigned long long seed) {
  const auto& c = sfh_compat::config();
  const int d[4] = {c.Spatial_Size, c.Spatial_Size, c.Spatial_Size, c.temporal_size};
  Link_array U((size_t)sfh_compat::volume() * 72);
  sfh_random_links(d, eps, seed, U.data());
  return U;
}

static int run() {
  sfh_compat::Config cfg;
  cfg.Spatial_Size = 4;
  cfg.temporal_size = 4;
  cfg.fermion_mass = 0.1;
  cfg.epsilon = 0.05;
  cfg.number_of_steps_fermion = 10;
  cfg.verbose = false;
  sfh_compat::configure(cfg);
  const double V = sfh_compat::volume();
  const double beta = 5.6;

  // cold-start known answers (ai_check.cpp:227-333, Lattice.cpp:784-804)
  {
    const Link_array U = cold();
    const double S = compute_action(U, beta);
    const double expect = -beta * (6 * V * cfg.c_11 + 12 * V * cfg.c_12);
    report("cold Symanzik action = -beta(6V c11 + 12V c12)", std::abs(S - expect) < 1e-13 * std::abs(expect),
           S - expect);
    report("cold average plaquette = 6", std::abs(compute_avg_plaq(U) - 6.0) < 1e-13, compute_avg_plaq(U) - 6.0);
    const auto F = compute_gauge_only_force_field(U, beta);
    double fmax = 0;
    for (double f : F) fmax = std::max(fmax, std::abs(f));
    report("cold gauge force vanishes", fmax < 1e-14, fmax);
  }

  Link_array U = hot(0.3, 11);
  // pseudofermion solve identities (App. A.5): Y = D^dag X = chi, DDdag X = phi
  {
    const auto phi = generate_phi_field(U);
    const auto X = solve_X_field(phi, U);
    std::vector<double> DDX(phi.size());
    apply_DDdagger(X, U, DDX);
    double r2 = 0, p2 = 0;
    for (size_t i = 0; i < phi.size(); ++i) {
      r2 += (phi[i] - DDX[i]) * (phi[i] - DDX[i]);
      p2 += phi[i] * phi[i];
    }
    report("CG true residual <= 1e-11", std::sqrt(r2 / p2) < 1e-11, std::sqrt(r2 / p2));
    const double sf = field_inner_product_real(phi, X);
    const auto Y = solve_Y_field(X, U);
    report("S_f = <phi, X> = |D^dag X|^2", std::abs(sf - field_inner_product_real(Y, Y)) < 1e-9 * sf,
           sf - field_inner_product_real(Y, Y));
  }

  // gauge force is the gradient of compute_action: dS/de = 2 ReTr(F H)
  {
    const auto F = compute_gauge_only_force_field(U, beta);
    const size_t l = 4 * 37 + 2;  // site 37, direction 2
    double Q[18] = {0}, E[18], Em[18];
    // H = (lambda_1 + lambda_5 + lambda_8) / 2
    Q[2] = 0.5; Q[6] = 0.5;                        // lambda_1
    Q[5] = -0.5; Q[13] = 0.5;                      // lambda_5 (0,2)=-i, (2,0)=+i
    const double a = 0.5 / std::sqrt(3.0);
    Q[0] = a; Q[8] = a; Q[16] = -2 * a;            // lambda_8
    const double eps = 1e-4;
    double Qp[18], Qm[18];
    for (int k = 0; k < 18; ++k) { Qp[k] = eps * Q[k]; Qm[k] = -eps * Q[k]; }
    sfh_ctx* c = nullptr;
    sfh_create(0, &c);
    sfh_exp_iq(c, Qp, E);
    sfh_exp_iq(c, Qm, Em);
    sfh_destroy(c);
    auto apply = [&](const double* M) {
      Link_array W = U;
      std::complex<double> u[9], m[9], r[9];
      for (int k = 0; k < 9; ++k) {
        u[k] = {U[18 * l + 2 * k], U[18 * l + 2 * k + 1]};
        m[k] = {M[2 * k], M[2 * k + 1]};
      }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[3 * i + j] = m[3 * i] * u[j] + m[3 * i + 1] * u[3 + j] + m[3 * i + 2] * u[6 + j];
      for (int k = 0; k < 9; ++k) { W[18 * l + 2 * k] = r[k].real(); W[18 * l + 2 * k + 1] = r[k].imag(); }
      return compute_action(W, beta);
    };
    const double num = (apply(E) - apply(Em)) / (2 * eps);
    double ana = 0;  // 2 ReTr(F H)
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) {
        const std::complex<double> f(F[18 * l + 2 * (3 * i + k)], F[18 * l + 2 * (3 * i + k) + 1]);
        const std::complex<double> h(Q[2 * (3 * k + i)], Q[2 * (3 * k + i) + 1]);
        ana += 2 * (f * h).real();
      }
    report("gauge force = FD gradient of the action", std::abs(num - ana) < 1e-6 * std::max(1.0, std::abs(ana)),
           num - ana);
  }

  // reversibility (Fermions.cpp:908-955)
  {
    const Link_array U0 = U;
    Link_array Ut = U;
    const auto phi = generate_phi_field(U0);
    auto P = generate_conjugate_field_configuration();
    sfh_compat::Config c2 = cfg;
    c2.tol = 1e-13;
    sfh_compat::configure(c2);
    apply_initial_step_to_conjugate_field(P, beta, phi, Ut);
    apply_intermediate_update_to_gauge_field_and_conjugate_field(P, beta, phi, Ut);
    apply_final_step(P, beta, phi, Ut);
    for (double& p : P) p = -p;
    apply_initial_step_to_conjugate_field(P, beta, phi, Ut);
    apply_intermediate_update_to_gauge_field_and_conjugate_field(P, beta, phi, Ut);
    apply_final_step(P, beta, phi, Ut);
    double dmax = 0;
    for (size_t i = 0; i < Ut.size(); ++i) dmax = std::max(dmax, std::abs(Ut[i] - U0[i]));
    report("reversibility: max link violation < 1e-10", dmax < 1e-10, dmax);
    sfh_compat::configure(cfg);
  }

  // the reference's exceptions
  {
    sfh_compat::Config c3 = cfg;
    c3.max_iter = 2;
    sfh_compat::configure(c3);
    bool thrown = false;
    try {
      const auto phi = generate_phi_field(U);
      solve_CG_DDdagger(phi, U);
    } catch (const std::runtime_error& e) {
      thrown = std::string(e.what()).find("failed to converge after 2 iterations") != std::string::npos;
    }
    report("non-convergence throws std::runtime_error", thrown, 0);
    sfh_compat::configure(cfg);
    thrown = false;
    try {
      auto P = generate_conjugate_field_configuration();
      P[0] += 1.0;
      Link_array W = U;
      const auto phi = generate_phi_field(U);
      apply_intermediate_update_to_gauge_field_and_conjugate_field(P, beta, phi, W);
    } catch (const std::runtime_error& e) {
      thrown = std::string(e.what()).find("not Hermitian") != std::string::npos;
    }
    report("non-hermitian momentum throws std::runtime_error", thrown, 0);
  }

  // a production HMC update runs and keeps the links in SU(3)
  {
    Link_array W = U;
    single_sweep_monte_carlo_update_WITH_FERMIONS(W, beta);
    double err = 0;
    for (size_t l = 0; l < W.size() / 18; ++l) {
      std::complex<double> m[9];
      for (int k = 0; k < 9; ++k) m[k] = {W[18 * l + 2 * k], W[18 * l + 2 * k + 1]};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          std::complex<double> s = 0;
          for (int k = 0; k < 3; ++k) s += m[3 * i + k] * std::conj(m[3 * j + k]);
          err = std::max(err, std::abs(s - (i == j ? 1.0 : 0.0)));
        }
    }
    report("HMC update keeps links unitary", err < 1e-12, err);
  }

  // heatbath sweeps (Lattice.cpp:725-757) from a cold start thermalise the
  // links and keep them in SU(3); the returned averages are those of the
  // reference (plaquette normalised to 1 on a cold lattice)
  {
    Link_array W = cold();
    std::pair<double, double> r{};
    for (int k = 0; k < 10; ++k) r = heatbath_update(W, 4.1);
    double err = 0;
    for (size_t l = 0; l < W.size() / 18; ++l) {
      std::complex<double> m[9];
      for (int k = 0; k < 9; ++k) m[k] = {W[18 * l + 2 * k], W[18 * l + 2 * k + 1]};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          std::complex<double> s = 0;
          for (int k = 0; k < 3; ++k) s += m[3 * i + k] * std::conj(m[3 * j + k]);
          err = std::max(err, std::abs(s - (i == j ? 1.0 : 0.0)));
        }
    }
    report("heatbath keeps links unitary", err < 1e-12, err);
    report("heatbath plaquette in (0, 1) and below cold", r.first > 0.2 && r.first < 0.99, r.first);
    report("heatbath plaquette matches compute_avg_plaq / 6 (10%)",
           std::abs(compute_avg_plaq(W) / 6.0 - r.first) < 0.1 * r.first, compute_avg_plaq(W) / 6.0 - r.first);
    report("heatbath action density negative", r.second < 0, r.second);
  }

  // HEX smearing (Smearing.cpp:304-331): cold links are a fixed point, the
  // history holds the thin links and four levels, smeared links stay in SU(3)
  {
    Link_array C = cold();
    Apply_4HEX_smearing(C);
    double d = 0;
    const Link_array C0 = cold();
    for (size_t k = 0; k < C.size(); ++k) d = std::max(d, std::abs(C[k] - C0[k]));
    report("4HEX leaves cold links unchanged", d == 0.0, d);
    const Link_array H = hot(0.4, 3);
    const auto hist = Apply_4HEX_smearing_with_history(H);
    Link_array W = H;
    Apply_4HEX_smearing(W);
    double dh = 0;
    for (size_t k = 0; k < W.size(); ++k) dh = std::max(dh, std::abs(W[k] - hist[4][k]));
    report("4HEX history: 5 levels, last = Apply_4HEX_smearing", hist.size() == 5 && hist[0] == H && dh == 0.0, dh);
    double err = 0;
    for (size_t l = 0; l < W.size() / 18; ++l) {
      std::complex<double> m[9];
      for (int k = 0; k < 9; ++k) m[k] = {W[18 * l + 2 * k], W[18 * l + 2 * k + 1]};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          std::complex<double> s = 0;
          for (int k = 0; k < 3; ++k) s += m[3 * i + k] * std::conj(m[3 * j + k]);
          err = std::max(err, std::abs(s - (i == j ? 1.0 : 0.0)));
        }
    }
    report("4HEX keeps links unitary", err < 1e-12, err);
  }

  // several host threads on different lattices at once, as the reference's
  // Main.cpp runs its configurations under `omp parallel for` (Main.cpp:174):
  // each thread gets its own context and stream, and every result equals the
  // same calls made one after another on one thread, bit for bit
  {
    constexpr int kThreads = 4;
    struct Out {
      std::vector<double> X, F;
      double S = 0;
    };
    auto work = [&](int t, Out& o) {
      const Link_array U = hot(0.3, 500 + t);
      std::vector<double> phi((size_t)V * 6);
      for (size_t k = 0; k < phi.size(); ++k) phi[k] = std::sin(0.37 * (double)k + t);
      o.X = solve_X_field(phi, U);
      o.F = compute_full_force_field(o.X, solve_Y_field(o.X, U), U, beta);
      o.S = compute_action(U, beta);
    };
    std::vector<Out> serial(kThreads), conc(kThreads);
    for (int t = 0; t < kThreads; ++t) work(t, serial[t]);
    std::vector<std::thread> pool;
    std::vector<std::string> errors(kThreads);
    for (int t = 0; t < kThreads; ++t)
      pool.emplace_back([&, t]() {
        try {
          work(t, conc[t]);
        } catch (const std::exception& e) {
          errors[t] = e.what();
        }
      });
    for (auto& th : pool) th.join();
    bool same = true;
    for (int t = 0; t < kThreads; ++t)
      same = same && errors[t].empty() && conc[t].X == serial[t].X && conc[t].F == serial[t].F &&
             conc[t].S == serial[t].S;
    report("4 threads at once = the same calls serially", same, same ? 0.0 : 1.0);
  }

  // measurement step of Main.cpp:219-228 on cold links: P = 3, C(r^2) = 9
  {
    const Link_array U = cold();
    const std::vector<complex> P = precompute_polyakov_grid(U);
    double dp = 0.0, dc = 0.0;
    for (const complex& p : P) dp = std::max(dp, std::abs(p - 3.0));
    const int maxdistance = cfg.Spatial_Size / 2 - 1;  // Parameters.h:15
    for (int r2 = 1; r2 <= maxdistance; ++r2) dc = std::max(dc, std::abs(correlator_over_fixed_distance_fast(P, r2) - 9.0));
    report("cold Polyakov grid = 3, correlators = 9", P.size() == 64 && dp < 1e-14 && dc < 1e-13, std::max(dp, dc));
    // a hot grid: C(1) summed over the six axis neighbours by hand
    const Link_array H = hot(0.6, 77);
    const std::vector<complex> Q = precompute_polyakov_grid(H);
    const int L = cfg.Spatial_Size;
    complex acc = 0.0;
    for (int x = 0; x < L; ++x)
      for (int y = 0; y < L; ++y)
        for (int z = 0; z < L; ++z) {
          const complex pm = Q[(x * L + y) * L + z];
          const int nb[6][3] = {{x - 1, y, z}, {x + 1, y, z}, {x, y - 1, z}, {x, y + 1, z}, {x, y, z - 1}, {x, y, z + 1}};
          for (const auto& n : nb)
            acc += pm * std::conj(Q[(((n[0] + L) % L) * L + (n[1] + L) % L) * L + (n[2] + L) % L]);
        }
    acc /= 6.0 * L * L * L;
    const double e = std::abs(correlator_over_fixed_distance_fast(Q, 1) - acc);
    report("Polyakov correlator r^2 = 1 = mean over 6 axis neighbours", e < 1e-12 * std::max(1.0, std::abs(acc)), e);
  }

  // checkpoints (Lattice.cpp:810-847): round trip, missing and short files
  {
    // the reference's non-template signatures with CONFIG = 2 (lattice_io_dropin.cpp)
    std::array<Link_array, 2> saved{hot(0.4, 81), hot(0.4, 82)}, loaded{cold(), cold()};
    const std::string path = "/tmp/sfh_dropin_selftest_ckpt.bin";
    const bool ok_save = save_all_lattice_configs(saved, path);
    const bool ok_load = load_all_lattice_configs(loaded, path);
    report("checkpoint round trip is byte-exact", ok_save && ok_load && loaded == saved, ok_load ? 0.0 : 1.0);
    std::array<Link_array, 3> three{cold(), cold(), cold()};
    const bool short_fails = !load_all_lattice_configs(three, path);
    const bool missing_fails = !load_all_lattice_configs(loaded, "/tmp/sfh_dropin_selftest_missing/none.bin");
    report("short and missing checkpoints return false", short_fails && missing_fails, 0.0);
    std::remove(path.c_str());
  }

  // resident links: a call after a link-changing call, or on an unchanged
  // array, gives the same result as after a fresh upload
  {
    Link_array U = hot(0.3, 91);
    const double a0 = compute_action(U, beta);
    const double p0 = compute_avg_plaq(U);
    U[5] += 1e-3;  // the caller changes one entry: the library must see it
    const double p1 = compute_avg_plaq(U);
    U[5] -= 1e-3;
    const double p2 = compute_avg_plaq(U);
    report("changed link array is re-uploaded", p1 != p0 && p2 == p0 && compute_action(U, beta) == a0,
           std::abs(p1 - p0));
  }

  std::printf("%s: %d failure(s)\n", failures ? "FAIL" : "PASS", failures);
  return failures ? 1 : 0;
}

// --fields SEED: host-side field generation only (no GPU): prints checksums
// of two momentum fields drawn on the main thread and one drawn on a second
// thread, for the reproducibility test (same seed, two processes).
static int fields(unsigned long long seed) {
  sfh_compat::Config cfg;
  cfg.seed = seed;
  cfg.verbose = false;
  sfh_compat::configure(cfg);
  auto sum = [](const std::vector<double>& v) {
    double a = 0.0, b = 0.0;
    for (size_t i = 0; i < v.size(); ++i) {
      a += v[i];
      b += v[i] * (double)(i % 97 + 1);
    }
    std::printf("%a %a %zu\n", a, b, v.size());
  };
  sum(generate_conjugate_field_configuration());
  sum(generate_conjugate_field_configuration());
  std::thread t([&] { sum(generate_conjugate_field_configuration()); });
  t.join();
  return 0;
}

int main(int argc, char** argv) {
  try {
    if (argc == 3 && std::string(argv[1]) == "--fields") return fields(std::stoull(argv[2]));
    return run();
  } catch (const std::exception& e) {
    std::printf("FAIL: unexpected exception: %s\n", e.what());
    return 2;
  }
}
