// Runs every exported function of the oracle on small lattices, built with
// AddressSanitizer and UndefinedBehaviorSanitizer (`make -C oracle sanitize`):
// an out-of-bounds access, a use of freed memory, a signed overflow or any
// other undefined behaviour aborts the run with the sanitizer's report.  It
// also checks that every call returns status 0 and finite results.
//
// TEST INFRASTRUCTURE ONLY (run by tests/test_oracle_sanitizers.py).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" {
int oc_apply_d(const int* dims, double mass, int dagger, const double* U, const double* in, double* out);
int oc_apply_ddag(const int* dims, double mass, const double* U, const double* in, double* out);
double oc_inner(long n, const double* a, const double* b);
int oc_cg(const int* dims, double mass, double tol, int max_iter, const double* U, const double* phi, double* X,
          int* iters, double* rsq);
int oc_staples(const int* dims, const double* U, int s, int mu, double* A1, double* A2);
int oc_force(const int* dims, const double* U, const double* X, const double* Y, double beta, double c11, double c12,
             int which, double* F);
double oc_action(const int* dims, const double* U, double beta, double c11, double c12);
double oc_avg_plaq(const int* dims, const double* U);
double oc_tr_p2(const int* dims, const double* P);
int oc_exp_iq(const double* Q18, double* out18);
int oc_link_update(const int* dims, double eps, const double* P, double* U, int check);
int oc_leapfrog(const int* dims, double mass, double tol, int max_iter, double beta, double c11, double c12,
                double eps, int nsteps, int which, int check, const double* phi, double* U, double* P, double* X);
int oc_hmc_trajectory(const int* dims, double mass, double tol, int max_iter, double beta, double c11, double c12,
                      double eps, int nsteps, const double* chi, const double* P0, double r_uniform, double* U,
                      double* H, int* accepted);
int oc_set_threads(int n);
int oc_polyakov_grid(const int* dims, const double* U, double* out);
int oc_polyakov_correlator(const int* spatial, const double* grid, int r_squared, double* out);
int oc_philox(const unsigned* key, const unsigned* ctr, unsigned* out);
int oc_heatbath_x0(double a, double beta, unsigned long long seed, unsigned sweep, int n, double* out);
int oc_heatbath_subgroup(const double* U18, const double* A18, double beta, int type, unsigned long long seed,
                         unsigned link, unsigned sweep, double* out18, double* ax);
int oc_heatbath(const int* dims, double* U, double beta, double c11, double c12, unsigned long long seed,
                unsigned sweep0, int nsweeps, int order, double* plaq, double* act);
int oc_hex_smear(const int* dims, double* U, const double* alpha, int nsteps);
}

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++failures;
  }
}

bool finite(const std::vector<double>& v) {
  for (double x : v)
    if (!std::isfinite(x)) return false;
  return true;
}

// xorshift64* uniforms in (-1, 1)
struct Rng {
  unsigned long long s;
  double next() {
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    const unsigned long long r = s * 2685821657736338717ULL;
    return 2.0 * ((double)(r >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
  }
};

// a Hermitian traceless 3x3 matrix (18 doubles, row-major, re/im)
void hermitian(Rng& rng, double scale, double* q) {
  double a[18];
  for (double& x : a) x = scale * rng.next();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const int ij = 2 * (3 * i + j), ji = 2 * (3 * j + i);
      q[ij] = 0.5 * (a[ij] + a[ji]);
      q[ij + 1] = 0.5 * (a[ij + 1] - a[ji + 1]);
    }
  const double tr = (q[0] + q[8] + q[16]) / 3.0;
  q[0] -= tr;
  q[8] -= tr;
  q[16] -= tr;
}

void run(const int dims[4]) {
  const int V = dims[0] * dims[1] * dims[2] * dims[3];
  const size_t nl = size_t(V) * 72, ns = size_t(V) * 6;
  Rng rng{0x9E3779B97F4A7C15ULL ^ (unsigned long long)V};
  std::vector<double> U(nl), P(nl), F(nl), chi(ns), phi(ns), X(ns), Y(ns), out(ns);
  for (size_t l = 0; l < nl / 18; ++l) {
    double q[18];
    hermitian(rng, 0.3, q);
    expect(oc_exp_iq(q, &U[18 * l]) == 0, "exp_iq");
    hermitian(rng, 1.0, &P[18 * l]);
  }
  for (double& x : chi) x = rng.next();
  const double mass = 0.2;
  expect(oc_apply_d(dims, mass, 0, U.data(), chi.data(), phi.data()) == 0, "apply_d");
  expect(oc_apply_d(dims, mass, 1, U.data(), chi.data(), out.data()) == 0, "apply_d dagger");
  expect(oc_apply_ddag(dims, mass, U.data(), chi.data(), out.data()) == 0, "apply_ddag");
  expect(std::isfinite(oc_inner((long)ns, phi.data(), chi.data())), "inner");
  int it = 0;
  double rsq = 0.0;
  expect(oc_cg(dims, mass, 1e-10, 5000, U.data(), phi.data(), X.data(), &it, &rsq) == 0 && it > 0, "cg");
  expect(oc_apply_d(dims, mass, 1, U.data(), X.data(), Y.data()) == 0, "Y = D^dag X");
  std::vector<double> A1(18), A2(18);
  for (int mu = 0; mu < 4; ++mu)
    expect(oc_staples(dims, U.data(), V - 1, mu, A1.data(), A2.data()) == 0 && finite(A1) && finite(A2), "staples");
  for (int which = 1; which <= 3; ++which)
    expect(oc_force(dims, U.data(), X.data(), Y.data(), 4.0, 5.0 / 3, -1.0 / 12, which, F.data()) == 0 && finite(F),
           "force");
  expect(std::isfinite(oc_action(dims, U.data(), 4.0, 5.0 / 3, -1.0 / 12)), "action");
  expect(std::isfinite(oc_avg_plaq(dims, U.data())), "avg_plaq");
  expect(std::isfinite(oc_tr_p2(dims, P.data())), "tr_p2");
  std::vector<double> W = U;
  expect(oc_link_update(dims, 0.01, P.data(), W.data(), 1) == 0 && finite(W), "link_update");
  std::vector<double> Pm = P;
  W = U;
  expect(oc_leapfrog(dims, mass, 1e-10, 5000, 4.0, 5.0 / 3, -1.0 / 12, 0.02, 3, 3, 1, phi.data(), W.data(),
                     Pm.data(), X.data()) == 0 && finite(W) && finite(Pm),
         "leapfrog");
  std::vector<double> H(8, 0.0);
  int acc = -1;
  W = U;
  expect(oc_hmc_trajectory(dims, mass, 1e-10, 5000, 4.0, 5.0 / 3, -1.0 / 12, 0.02, 3, chi.data(), P.data(), 0.5,
                           W.data(), H.data(), &acc) == 0 && (acc == 0 || acc == 1),
         "hmc_trajectory");
  const int nsp = dims[0] * dims[1] * dims[2];
  std::vector<double> grid(2 * size_t(nsp)), corr(2);
  expect(oc_polyakov_grid(dims, U.data(), grid.data()) == 0 && finite(grid), "polyakov_grid");
  for (int r2 = 1; r2 <= 6; ++r2)
    expect(oc_polyakov_correlator(dims, grid.data(), r2, corr.data()) == 0 && finite(corr), "polyakov_correlator");
  std::vector<double> pl(2), ac(2);
  W = U;
  for (int order = 0; order <= 1; ++order)
    expect(oc_heatbath(dims, W.data(), 4.1, 5.0 / 3, -1.0 / 12, 7, 0, 2, order, pl.data(), ac.data()) == 0 &&
               finite(W) && finite(pl) && finite(ac),
           "heatbath");
  const double alpha[3] = {0.95, 0.76, 0.38};
  W = U;
  expect(oc_hex_smear(dims, W.data(), alpha, 2) == 0 && finite(W), "hex_smear");
}

}  // namespace

int main() {
  oc_set_threads(2);
  const unsigned key[2] = {1u, 2u}, ctr[4] = {3u, 4u, 5u, 6u};
  unsigned pk[4];
  expect(oc_philox(key, ctr, pk) == 0, "philox");
  std::vector<double> x0(64), u18(18, 0.0), a18(18, 0.0), o18(18), ax(2);
  expect(oc_heatbath_x0(0.8, 2.7, 11, 0, 64, x0.data()) == 0 && finite(x0), "heatbath_x0");
  u18[0] = u18[8] = u18[16] = 1.0;
  a18[0] = 2.0;
  a18[8] = 1.5;
  a18[16] = 1.0;
  for (int type = 0; type < 3; ++type)
    expect(oc_heatbath_subgroup(u18.data(), a18.data(), 4.0, type, 5, 0, 0, o18.data(), ax.data()) == 0,
           "heatbath_subgroup");
  const int d1[4] = {4, 4, 4, 4}, d2[4] = {4, 4, 4, 8};
  run(d1);
  run(d2);
  if (failures) {
    std::fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  std::printf("sanitized oracle run ok\n");
  return 0;
}
