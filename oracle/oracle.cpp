// FP64 CPU oracle for the staggered-fermion HMC hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load the shared library built from it.  The product path is the HIP
// library in staggered_fermions_hmc_amd/ and fails loudly without it.
//
// It restates, in plain C++17 + OpenMP (no Eigen), the algorithms of the
// reference files listed beside each function, with the reference's
// compile-time constants (lattice extents, mass, tol, max_iter, epsilon,
// number of MD steps, c_11/c_12) turned into runtime arguments.  Layouts are
// the reference's own array-of-structures layouts:
//   site     s = ((n0*L1 + n1)*L2 + n2)*L3 + n3        (Fermions.cpp:42)
//   link     (s*4 + mu)*18 + (3a+b)*2 + {re,im}          (Lattice.cpp:135-151)
//   spinor   s*6 + 2c + {re,im}                          (Fermions.cpp:99-116)
// Time is direction 3; it carries the antiperiodic fermion boundary.
//
// Status codes match include/sfh.h so tests can compare error paths.
//
// How each function is pinned (tests/test_oracle_*.py): cold-start staple
// and action known answers of ai_check.cpp:172-333; the free-field closed
// form of D D^dagger; hand-derived delta-source stencils at boundary sites;
// anti-hermiticity and epsilon-anticommutation of the hop; gauge covariance;
// dense numpy solves; brute-force enumeration of every plaquette and
// rectangle; finite-difference derivatives of the actions against the
// forces (dS/de = 2 ReTr(F H), ai_check.cpp:31-34); eigendecomposition for
// the exponential; reversibility and dH ~ eps^2 for the integrator.
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

using cd = std::complex<double>;

enum {
  OK = 0, E_INVALID = 1, E_BREAKDOWN = 5, E_NONFINITE_ALPHA = 6,
  E_NONFINITE_BETA = 7, E_NOCONV = 8, E_NOT_HERMITIAN = 9, E_EXP_QCUBED = 10,
  E_EXP_NEG_TRACE = 11, E_SAMPLER = 12
};

struct M3 { cd a[9]; };

inline M3 zero3() { M3 m; for (auto& x : m.a) x = 0.0; return m; }
inline M3 eye3() { M3 m = zero3(); m.a[0] = m.a[4] = m.a[8] = 1.0; return m; }
inline M3 mul(const M3& x, const M3& y) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cd s = 0.0;
      for (int k = 0; k < 3; ++k) s += x.a[3 * i + k] * y.a[3 * k + j];
      r.a[3 * i + j] = s;
    }
  return r;
}
inline M3 dag(const M3& x) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[3 * i + j] = std::conj(x.a[3 * j + i]);
  return r;
}
inline M3 add(const M3& x, const M3& y) { M3 r; for (int i = 0; i < 9; ++i) r.a[i] = x.a[i] + y.a[i]; return r; }
inline M3 sub(const M3& x, const M3& y) { M3 r; for (int i = 0; i < 9; ++i) r.a[i] = x.a[i] - y.a[i]; return r; }
inline M3 scale(cd c, const M3& x) { M3 r; for (int i = 0; i < 9; ++i) r.a[i] = c * x.a[i]; return r; }
inline cd trace(const M3& x) { return x.a[0] + x.a[4] + x.a[8]; }
inline double fnorm(const M3& x) { double s = 0; for (auto& v : x.a) s += std::norm(v); return std::sqrt(s); }

struct Geo {
  int L[4];
  int V;
  std::vector<int> up, dn, bup, bdn, eta;  // [s*4+mu]
  std::vector<int> coord;                  // [s*4+k]
  explicit Geo(const int* dims) {
    for (int k = 0; k < 4; ++k) L[k] = dims[k];
    V = L[0] * L[1] * L[2] * L[3];
    up.resize(4 * V); dn.resize(4 * V); bup.resize(4 * V); bdn.resize(4 * V);
    eta.resize(4 * V); coord.resize(4 * V);
    // Neighbour, boundary-sign and eta tables: Fermions.cpp:20-94.
    for (int s = 0; s < V; ++s) {
      int n[4], r = s;
      for (int k = 3; k >= 0; --k) { n[k] = r % L[k]; r /= L[k]; }
      for (int k = 0; k < 4; ++k) coord[4 * s + k] = n[k];
      for (int mu = 0; mu < 4; ++mu) {
        int e = 1;
        for (int nu = 0; nu < mu; ++nu) if (n[nu] & 1) e = -e;
        eta[4 * s + mu] = e;
        int m[4];
        std::memcpy(m, n, sizeof m);
        int b = 1;
        m[mu] += 1;
        if (m[mu] == L[mu]) { m[mu] = 0; if (mu == 3) b = -1; }
        up[4 * s + mu] = index(m); bup[4 * s + mu] = b;
        std::memcpy(m, n, sizeof m);
        b = 1;
        if (m[mu] == 0) { m[mu] = L[mu] - 1; if (mu == 3) b = -1; } else m[mu] -= 1;
        dn[4 * s + mu] = index(m); bdn[4 * s + mu] = b;
      }
    }
  }
  int index(const int* n) const { return ((n[0] * L[1] + n[1]) * L[2] + n[2]) * L[3] + n[3]; }
  // Periodic move used by the gauge-action walks (Lattice.cpp:213-235).
  int step(int s, int mu, int dir) const { return dir > 0 ? up[4 * s + mu] : dn[4 * s + mu]; }
};

inline M3 load_link(const double* U, int s, int mu) {
  M3 m; const double* p = U + (size_t(s) * 4 + mu) * 18;
  for (int i = 0; i < 9; ++i) m.a[i] = cd(p[2 * i], p[2 * i + 1]);
  return m;
}
inline void store_link(double* U, int s, int mu, const M3& m) {
  double* p = U + (size_t(s) * 4 + mu) * 18;
  for (int i = 0; i < 9; ++i) { p[2 * i] = m.a[i].real(); p[2 * i + 1] = m.a[i].imag(); }
}
inline void load_vec(const double* f, int s, cd v[3]) {
  for (int c = 0; c < 3; ++c) v[c] = cd(f[6 * s + 2 * c], f[6 * s + 2 * c + 1]);
}
inline void store_vec(double* f, int s, const cd v[3]) {
  for (int c = 0; c < 3; ++c) { f[6 * s + 2 * c] = v[c].real(); f[6 * s + 2 * c + 1] = v[c].imag(); }
}
inline void mv(const M3& m, const cd v[3], cd out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = m.a[3 * i] * v[0] + m.a[3 * i + 1] * v[1] + m.a[3 * i + 2] * v[2];
}
inline void mdv(const M3& m, const cd v[3], cd out[3]) {  // m^dagger v
  for (int i = 0; i < 3; ++i)
    out[i] = std::conj(m.a[i]) * v[0] + std::conj(m.a[3 + i]) * v[1] + std::conj(m.a[6 + i]) * v[2];
}

// D psi (sign=+1) or D^dagger psi (sign=-1) at one site: Fermions.cpp:218-247, :254-283.
inline void dirac_site(const Geo& g, double mass, int sign, const double* U, const double* psi, int s, cd out[3]) {
  cd x[3];
  load_vec(psi, s, x);
  for (int c = 0; c < 3; ++c) out[c] = mass * x[c];
  for (int mu = 0; mu < 4; ++mu) {
    const double e = g.eta[4 * s + mu];
    const int su = g.up[4 * s + mu], sd = g.dn[4 * s + mu];
    cd nb[3], t[3];
    load_vec(psi, su, nb);
    mv(load_link(U, s, mu), nb, t);
    const double cf = 0.5 * e * g.bup[4 * s + mu] * sign;
    for (int c = 0; c < 3; ++c) out[c] += cf * t[c];
    load_vec(psi, sd, nb);
    mdv(load_link(U, sd, mu), nb, t);
    const double cb = 0.5 * e * g.bdn[4 * s + mu] * sign;
    for (int c = 0; c < 3; ++c) out[c] -= cb * t[c];
  }
}

void dirac(const Geo& g, double mass, int sign, const double* U, const double* in, double* out) {
#pragma omp parallel for
  for (int s = 0; s < g.V; ++s) {
    cd r[3];
    dirac_site(g, mass, sign, U, in, s, r);
    store_vec(out, s, r);
  }
}

double dot(size_t n, const double* a, const double* b) {
  double sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
  for (long i = 0; i < (long)n; ++i) sum += a[i] * b[i];
  return sum;
}

// D D^dagger = D(D^dagger in): Fermions.cpp:530-547.
void ddag(const Geo& g, double mass, const double* U, const double* in, double* out, double* tmp) {
  dirac(g, mass, -1, U, in, tmp);
  dirac(g, mass, +1, U, tmp, out);
}

// Unpreconditioned CG on D D^dagger with the reference's recursive-residual
// stop and error checks: Fermions.cpp:553-617.
int cg(const Geo& g, double mass, double tol, int max_iter, const double* U, const double* phi,
       double* X, int* iters_out, double* rsq_out) {
  const size_t n = size_t(g.V) * 6;
  std::vector<double> r(phi, phi + n), p(phi, phi + n), v(n), tmp(n);
  std::fill(X, X + n, 0.0);
  double rsq = dot(n, r.data(), r.data());
  const double phisq = dot(n, phi, phi);
  if (iters_out) *iters_out = 0;
  if (rsq_out) *rsq_out = rsq;
  if (phisq < 1e-28) return OK;
  const double target = tol * tol * phisq;
  int it = 0;
  for (; it < max_iter; ++it) {
    if (rsq < target) break;
    ddag(g, mass, U, p.data(), v.data(), tmp.data());
    const double pv = dot(n, p.data(), v.data());
    if (!std::isfinite(pv) || std::abs(pv) < 1e-30) { if (iters_out) *iters_out = it; return E_BREAKDOWN; }
    const double alpha = rsq / pv;
    if (!std::isfinite(alpha)) { if (iters_out) *iters_out = it; return E_NONFINITE_ALPHA; }
#pragma omp parallel for
    for (long i = 0; i < (long)n; ++i) { X[i] += alpha * p[i]; r[i] -= alpha * v[i]; }
    const double rnew = dot(n, r.data(), r.data());
    const double beta = rnew / rsq;
    if (!std::isfinite(beta)) { if (iters_out) *iters_out = it + 1; return E_NONFINITE_BETA; }
    rsq = rnew;
#pragma omp parallel for
    for (long i = 0; i < (long)n; ++i) p[i] = r[i] + beta * p[i];
  }
  if (iters_out) *iters_out = it;
  if (rsq_out) *rsq_out = rsq;
  return rsq >= target ? E_NOCONV : OK;
}

// Plaquette staple sum A1 at link (s, mu): Lattice.cpp:249-313.
M3 staple(const Geo& g, const double* U, int s, int mu) {
  M3 sum = zero3();
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
    // lower: U_nu(n+mu-nu)^dag U_mu(n-nu)^dag U_nu(n-nu)
    const int a = g.step(s, nu, -1);
    const int b = g.step(a, mu, +1);
    M3 lo = mul(dag(load_link(U, b, nu)), mul(dag(load_link(U, a, mu)), load_link(U, a, nu)));
    // upper: U_nu(n+mu) U_mu(n+nu)^dag U_nu(n)^dag
    const int c = g.step(s, nu, +1);
    const int d = g.step(s, mu, +1);
    M3 hi = mul(load_link(U, d, nu), mul(dag(load_link(U, c, mu)), dag(load_link(U, s, nu))));
    sum = add(add(sum, lo), hi);
  }
  return sum;
}

// Rectangle staple sum A2 at link (s, mu), six 2x1 / 1x2 shapes per nu:
// Lattice.cpp:315-499.  Each term T satisfies ReTr(U_mu(n) T) = ReTr(loop).
M3 rect_staple(const Geo& g, const double* U, int s, int mu) {
  M3 sum = zero3();
  auto L = [&](int x, int d) { return load_link(U, x, d); };
  auto mv_ = [&](int x, int d, int k) { return g.step(x, d, k); };
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
    const int n = s;
    const int nP = mv_(n, mu, +1);            // n+mu
    const int n2P = mv_(nP, mu, +1);          // n+2mu
    const int nM = mv_(n, mu, -1);            // n-mu
    const int nU = mv_(n, nu, +1);            // n+nu
    const int n2U = mv_(nU, nu, +1);          // n+2nu
    const int nD = mv_(n, nu, -1);            // n-nu
    const int n2D = mv_(nD, nu, -1);          // n-2nu
    // 1: short side, below (2 nu-links down): loop n->n+mu->n+mu-nu->n+mu-2nu->n-2nu->n-nu->n
    {
      const int a = mv_(nP, nu, -1), b = mv_(a, nu, -1);
      M3 t = mul(dag(L(a, nu)), dag(L(b, nu)));
      t = mul(t, dag(L(n2D, mu)));
      t = mul(t, L(n2D, nu));
      t = mul(t, L(nD, nu));
      sum = add(sum, t);
    }
    // 2: short side, above: n->n+mu->n+mu+nu->n+mu+2nu->n+2nu->n+nu->n
    {
      const int a = mv_(nP, nu, +1);
      M3 t = mul(L(nP, nu), L(a, nu));
      t = mul(t, dag(L(n2U, mu)));
      t = mul(t, dag(L(nU, nu)));
      t = mul(t, dag(L(n, nu)));
      sum = add(sum, t);
    }
    // 3: long side, link first, below: n->n+mu->n+2mu->n+2mu-nu->n+mu-nu->n-nu->n
    {
      const int a = mv_(n2P, nu, -1), b = mv_(nP, nu, -1);
      M3 t = mul(L(nP, mu), dag(L(a, nu)));
      t = mul(t, dag(L(b, mu)));
      t = mul(t, dag(L(nD, mu)));
      t = mul(t, L(nD, nu));
      sum = add(sum, t);
    }
    // 4: long side, link second, above: n->n+mu->n+mu+nu->n+nu->n-mu+nu->n-mu->n
    {
      const int a = mv_(nM, nu, +1);
      M3 t = mul(L(nP, nu), dag(L(nU, mu)));
      t = mul(t, dag(L(a, mu)));
      t = mul(t, dag(L(nM, nu)));
      t = mul(t, L(nM, mu));
      sum = add(sum, t);
    }
    // 5: long side, link second, below: n->n+mu->n+mu-nu->n-nu->n-mu-nu->n-mu->n
    {
      const int a = mv_(nP, nu, -1), b = mv_(nM, nu, -1);
      M3 t = mul(dag(L(a, nu)), dag(L(nD, mu)));
      t = mul(t, dag(L(b, mu)));
      t = mul(t, L(b, nu));
      t = mul(t, L(nM, mu));
      sum = add(sum, t);
    }
    // 6: long side, link first, above: n->n+mu->n+2mu->n+2mu+nu->n+mu+nu->n+nu->n
    {
      const int a = mv_(nP, nu, +1);
      M3 t = mul(L(nP, mu), L(n2P, nu));
      t = mul(t, dag(L(a, mu)));
      t = mul(t, dag(L(nU, mu)));
      t = mul(t, dag(L(n, nu)));
      sum = add(sum, t);
    }
  }
  return sum;
}

inline M3 traceless(const M3& f) {
  M3 r = f; const cd t = trace(f) / 3.0;
  r.a[0] -= t; r.a[4] -= t; r.a[8] -= t;
  return r;
}

// Symanzik gauge force: Fermions.cpp:372-384 with Lattice.cpp:502-507.
M3 gauge_force_link(const Geo& g, const double* U, int s, int mu, double beta, double c11, double c12) {
  const M3 A = add(scale(c11, staple(g, U, s, mu)), scale(c12, rect_staple(g, U, s, mu)));
  const M3 Q = mul(load_link(U, s, mu), A);
  return traceless(scale(cd(0.0, -beta / 12.0), sub(Q, dag(Q))));
}

// Staggered fermion force: Fermions.cpp:443-515.
M3 fermion_force_link(const Geo& g, const double* U, const double* X, const double* Y, int s, int mu) {
  const int su = g.up[4 * s + mu];
  const double c = 0.25 * g.eta[4 * s + mu] * g.bup[4 * s + mu];
  const M3 u = load_link(U, s, mu);
  cd x0[3], x1[3], y0[3], y1[3], uy[3];
  load_vec(X, s, x0); load_vec(X, su, x1); load_vec(Y, s, y0); load_vec(Y, su, y1);
  mv(u, y1, uy);
  M3 af, ab0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      af.a[3 * i + j] = uy[i] * std::conj(x0[j]);
      ab0.a[3 * i + j] = y0[i] * std::conj(x1[j]);
    }
  const M3 ab = mul(ab0, dag(u));
  const M3 f = scale(cd(0.0, -c), add(sub(af, dag(af)), sub(ab, dag(ab))));
  return traceless(f);
}

// which: bit0 = gauge, bit1 = fermion (Fermions.cpp:640-659, :1114-1134, :1278-1294).
void force_field(const Geo& g, const double* U, const double* X, const double* Y, double beta, double c11,
                 double c12, int which, double* F) {
#pragma omp parallel for collapse(2)
  for (int s = 0; s < g.V; ++s)
    for (int mu = 0; mu < 4; ++mu) {
      M3 f = zero3();
      if (which & 1) f = gauge_force_link(g, U, s, mu, beta, c11, c12);
      if (which & 2) f = add(f, fermion_force_link(g, U, X, Y, s, mu));
      store_link(F, s, mu, f);
    }
}

// Gauge action, summed per link exactly as Lattice.cpp:759-781 (serial).
double action(const Geo& g, const double* U, double beta, double c11, double c12) {
  double total = 0.0;
  for (int s = 0; s < g.V; ++s)
    for (int mu = 0; mu < 4; ++mu) {
      const M3 u = load_link(U, s, mu);
      const double p = trace(mul(u, staple(g, U, s, mu))).real() / 4.0;
      const double r = trace(mul(u, rect_staple(g, U, s, mu))).real() / 6.0;
      total += -(beta / 3.0) * (c11 * p + c12 * r);
    }
  return total;
}

// Lattice.cpp:784-804 (note: normalised to 6 on a cold lattice).
double avg_plaq(const Geo& g, const double* U) {
  double total = 0.0;
  for (int s = 0; s < g.V; ++s)
    for (int mu = 0; mu < 4; ++mu)
      total += trace(mul(load_link(U, s, mu), staple(g, U, s, mu))).real() / 3.0;
  return total / (4.0 * g.V);
}

// Fermions.cpp:662-678.
double tr_p2(const Geo& g, const double* P) {
  double sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
  for (long l = 0; l < 4L * g.V; ++l) {
    const M3 p = load_link(P, int(l / 4), int(l % 4));
    sum += trace(mul(p, p)).real();
  }
  return sum;
}

// Fermions.cpp:834-841.
double sinc_stable(double x) {
  if (std::abs(x) <= 0.05) {
    const double x2 = x * x;
    return 1.0 - (x2 / 6.0) * (1.0 - (x2 / 20.0) * (1.0 - x2 / 42.0));
  }
  return std::sin(x) / x;
}

// exp(iQ) for hermitian traceless Q by Cayley-Hamilton (Morningstar & Peardon,
// hep-lat/0311018, eqs. 23-34), with the reference's guards: Fermions.cpp:843-902.
int exp_iq(const M3& Q, M3& out) {
  const M3 Q2 = mul(Q, Q);
  const M3 Q3 = mul(Q2, Q);
  if (std::abs(trace(Q3).imag()) > 1e-10) return E_EXP_QCUBED;
  double c0 = trace(Q3).real() / 3.0;
  const double c1 = trace(Q2).real() / 2.0;
  if (c1 < 0.0) return E_EXP_NEG_TRACE;
  if (c1 < 1e-12) {
    out = sub(add(eye3(), scale(cd(0, 1), Q)), scale(0.5, Q2));
    return OK;
  }
  const bool neg = c0 < 0.0;
  c0 = std::abs(c0);
  const double c0max = 2.0 * std::sqrt(c1 * c1 * c1 / 27.0);
  const double ratio = c0max > 0.0 ? std::clamp(c0 / c0max, -1.0, 1.0) : 1.0;
  const double theta = std::acos(ratio);
  const double w = std::sqrt(c1) * std::sin(theta / 3.0);
  const double u = std::sqrt(c1 / 3.0) * std::cos(theta / 3.0);
  const cd e2u = std::exp(cd(0, 2 * u)), emu = std::exp(cd(0, -u));
  const double xi = sinc_stable(w), cw = std::cos(w);
  const cd h0 = (u * u - w * w) * e2u + emu * (8.0 * u * u * cw + cd(0, 2 * u * (3 * u * u + w * w)) * xi);
  const cd h1 = 2.0 * u * e2u - emu * (2.0 * u * cw - cd(0, 3 * u * u - w * w) * xi);
  const cd h2 = e2u - emu * (cw + cd(0, 3 * u) * xi);
  const double den = 9.0 * u * u - w * w;
  cd f0 = h0 / den, f1 = h1 / den, f2 = h2 / den;
  if (neg) { f0 = std::conj(f0); f1 = -std::conj(f1); f2 = std::conj(f2); }
  out = add(add(scale(f0, eye3()), scale(f1, Q)), scale(f2, Q2));
  return OK;
}

// U <- exp(i eps P) U on every link with the hermiticity guard of
// Fermions.cpp:703-723 (check=1) or without it as in :982-993 (check=0).
int link_update(const Geo& g, double eps, const double* P, double* U, int check) {
  int status = OK;
#pragma omp parallel for collapse(2)
  for (int s = 0; s < g.V; ++s)
    for (int mu = 0; mu < 4; ++mu) {
      const M3 q = scale(eps, load_link(P, s, mu));
      if (check && (fnorm(sub(q, dag(q))) > 1e-10 || std::abs(trace(q)) > 1e-10)) {
#pragma omp critical
        status = E_NOT_HERMITIAN;
        continue;
      }
      M3 e;
      const int st = exp_iq(q, e);
      if (st != OK) {
#pragma omp critical
        status = st;
        continue;
      }
      store_link(U, s, mu, mul(e, load_link(U, s, mu)));
    }
  return status;
}

struct MD {
  double mass, tol, beta, c11, c12, eps;
  int max_iter, nsteps, which, check;
};

// One force evaluation + momentum kick P -= coef * F; returns X through xout.
int kick(const Geo& g, const MD& md, const double* phi, const double* U, double* P, double coef, double* xout) {
  const size_t n6 = size_t(g.V) * 6, n18 = size_t(g.V) * 72;
  std::vector<double> X(n6, 0.0), Y(n6, 0.0), F(n18);
  if (md.which & 2) {
    int it;
    double rr;
    const int st = cg(g, md.mass, md.tol, md.max_iter, U, phi, X.data(), &it, &rr);
    if (st != OK) return st;
    dirac(g, md.mass, -1, U, X.data(), Y.data());
  }
  force_field(g, U, X.data(), Y.data(), md.beta, md.c11, md.c12, md.which, F.data());
#pragma omp parallel for
  for (long i = 0; i < (long)n18; ++i) P[i] -= coef * F[i];
  if (xout) std::memcpy(xout, X.data(), n6 * sizeof(double));
  return OK;
}

// Leapfrog: Fermions.cpp:684-776 (which=3), :1114-1228 (which=2), :1278-1376 (which=1).
int leapfrog(const Geo& g, const MD& md, const double* phi, double* U, double* P, double* xfinal) {
  int st = kick(g, md, phi, U, P, md.eps * 0.5, nullptr);
  if (st) return st;
  for (int n = 1; n < md.nsteps; ++n) {
    if ((st = link_update(g, md.eps, P, U, md.check))) return st;
    if ((st = kick(g, md, phi, U, P, md.eps, nullptr))) return st;
  }
  if ((st = link_update(g, md.eps, P, U, md.check))) return st;
  return kick(g, md, phi, U, P, md.eps * 0.5, xfinal);
}

}  // namespace

extern "C" {

int oc_apply_d(const int* dims, double mass, int dagger, const double* U, const double* in, double* out) {
  Geo g(dims);
  dirac(g, mass, dagger ? -1 : +1, U, in, out);
  return OK;
}

int oc_apply_ddag(const int* dims, double mass, const double* U, const double* in, double* out) {
  Geo g(dims);
  std::vector<double> tmp(size_t(g.V) * 6);
  ddag(g, mass, U, in, out, tmp.data());
  return OK;
}

double oc_inner(long n, const double* a, const double* b) { return dot(size_t(n), a, b); }

int oc_cg(const int* dims, double mass, double tol, int max_iter, const double* U, const double* phi, double* X,
          int* iters, double* rsq) {
  Geo g(dims);
  return cg(g, mass, tol, max_iter, U, phi, X, iters, rsq);
}

// Times n repetitions of one CG iteration's work (D D^dagger + 2 dots +
// 2 axpy sweeps) and returns seconds per iteration: the CPU baseline kernel.
double oc_time_cg_iterations(const int* dims, double mass, const double* U, const double* phi, int n) {
  Geo g(dims);
  const size_t N = size_t(g.V) * 6;
  std::vector<double> x(N, 0.0), r(phi, phi + N), p(phi, phi + N), v(N), tmp(N);
  double rsq = dot(N, r.data(), r.data());
  const double t0 = omp_get_wtime();
  for (int it = 0; it < n; ++it) {
    ddag(g, mass, U, p.data(), v.data(), tmp.data());
    const double alpha = rsq / dot(N, p.data(), v.data());
#pragma omp parallel for
    for (long i = 0; i < (long)N; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * v[i]; }
    const double rn = dot(N, r.data(), r.data());
    const double beta = rn / rsq;
    rsq = rn;
#pragma omp parallel for
    for (long i = 0; i < (long)N; ++i) p[i] = r[i] + beta * p[i];
  }
  return (omp_get_wtime() - t0) / n;
}

int oc_staples(const int* dims, const double* U, int s, int mu, double* A1, double* A2) {
  Geo g(dims);
  const M3 a = staple(g, U, s, mu), b = rect_staple(g, U, s, mu);
  for (int i = 0; i < 9; ++i) {
    A1[2 * i] = a.a[i].real(); A1[2 * i + 1] = a.a[i].imag();
    A2[2 * i] = b.a[i].real(); A2[2 * i + 1] = b.a[i].imag();
  }
  return OK;
}

int oc_force(const int* dims, const double* U, const double* X, const double* Y, double beta, double c11,
             double c12, int which, double* F) {
  Geo g(dims);
  force_field(g, U, X, Y, beta, c11, c12, which, F);
  return OK;
}

double oc_action(const int* dims, const double* U, double beta, double c11, double c12) {
  Geo g(dims);
  return action(g, U, beta, c11, c12);
}

double oc_avg_plaq(const int* dims, const double* U) {
  Geo g(dims);
  return avg_plaq(g, U);
}

double oc_tr_p2(const int* dims, const double* P) {
  Geo g(dims);
  return tr_p2(g, P);
}

int oc_exp_iq(const double* Q18, double* out18) {
  M3 q, e;
  for (int i = 0; i < 9; ++i) q.a[i] = cd(Q18[2 * i], Q18[2 * i + 1]);
  const int st = exp_iq(q, e);
  if (st == OK)
    for (int i = 0; i < 9; ++i) { out18[2 * i] = e.a[i].real(); out18[2 * i + 1] = e.a[i].imag(); }
  return st;
}

int oc_link_update(const int* dims, double eps, const double* P, double* U, int check) {
  Geo g(dims);
  return link_update(g, eps, P, U, check);
}

// Leapfrog trajectory from (U, P) for a fixed phi.  which: 1 gauge, 2 fermion,
// 3 both.  On return U and P hold the end point and X the final solve.
int oc_leapfrog(const int* dims, double mass, double tol, int max_iter, double beta, double c11, double c12,
                double eps, int nsteps, int which, int check, const double* phi, double* U, double* P,
                double* X) {
  Geo g(dims);
  MD md{mass, tol, beta, c11, c12, eps, max_iter, nsteps, which, check};
  return leapfrog(g, md, phi, U, P, X);
}

// One HMC update as in Fermions.cpp:778-828, with chi, P and the Metropolis
// uniform r supplied by the caller (the reference draws them itself).
// H[0..5] = TrP2_i, Sg_i, Sf_i, TrP2_f, Sg_f, Sf_f; H[6] = H_i - H_f.
int oc_hmc_trajectory(const int* dims, double mass, double tol, int max_iter, double beta, double c11, double c12,
                      double eps, int nsteps, const double* chi, const double* P0, double r_uniform, double* U,
                      double* H, int* accepted) {
  Geo g(dims);
  const size_t n6 = size_t(g.V) * 6, n18 = size_t(g.V) * 72;
  MD md{mass, tol, beta, c11, c12, eps, max_iter, nsteps, 3, 1};
  std::vector<double> phi(n6), X(n6), P(P0, P0 + n18), Uc(U, U + n18), Xf(n6);
  const double sg_i = action(g, U, beta, c11, c12);
  dirac(g, mass, +1, U, chi, phi.data());
  const double tp_i = tr_p2(g, P.data());
  int it; double rr;
  int st = cg(g, mass, tol, max_iter, U, phi.data(), X.data(), &it, &rr);
  if (st) return st;
  const double sf_i = dot(n6, phi.data(), X.data());
  if ((st = leapfrog(g, md, phi.data(), Uc.data(), P.data(), Xf.data()))) return st;
  const double tp_f = tr_p2(g, P.data());
  const double sg_f = action(g, Uc.data(), beta, c11, c12);
  const double sf_f = dot(n6, phi.data(), Xf.data());
  const double comp = tp_i - tp_f + sg_i - sg_f + sf_i - sf_f;
  H[0] = tp_i; H[1] = sg_i; H[2] = sf_i; H[3] = tp_f; H[4] = sg_f; H[5] = sf_f; H[6] = comp;
  *accepted = r_uniform < std::exp(comp);
  if (*accepted) std::memcpy(U, Uc.data(), n18 * sizeof(double));
  return OK;
}

int oc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

}  // extern "C"

extern "C" {

// Polyakov loop traces, Lattice.cpp:849-884: Tr prod_{t=0}^{T-1} U_3(x,y,z,t)
// (right multiplication from t = 0), at [(x*L1 + y)*L2 + z] as (re, im).
int oc_polyakov_grid(const int* dims, const double* U, double* out) {
  Geo g(dims);
  const int ns = g.L[0] * g.L[1] * g.L[2];
#pragma omp parallel for
  for (int k = 0; k < ns; ++k) {
    const int x = k / (g.L[1] * g.L[2]), y = (k / g.L[2]) % g.L[1], z = k % g.L[2];
    M3 w = eye3();
    for (int t = 0; t < g.L[3]; ++t) {
      const int n[4] = {x, y, z, t};
      w = mul(w, load_link(U, g.index(n), 3));
    }
    const cd tr = trace(w);
    out[2 * k] = tr.real();
    out[2 * k + 1] = tr.imag();
  }
  return OK;
}

// correlator_over_fixed_distance_fast, Lattice.cpp:886-948: the
// displacements (dx, dy, dz) in ascending loop order with
// |d_k| <= min(floor(sqrt(r^2)), L_k / 2), d_k = L_k / 2 skipped for even L_k
// and d.d = r^2 (:892-913, one Spatial_Size generalised to L0, L1, L2); then
// the sum of P(m) conj(P(m + d)) over the grid in (x, y, z) order with the
// displacements innermost, periodic wrap, divided by the number of terms
// (:915-947).  Serial, in the reference's order.
int oc_polyakov_correlator(const int* spatial, const double* grid, int r_squared, double* out) {
  const int L0 = spatial[0], L1 = spatial[1], L2 = spatial[2];
  const int L[3] = {L0, L1, L2};
  int md[3];
  for (int k = 0; k < 3; ++k) md[k] = std::min(static_cast<int>(std::sqrt(r_squared)), L[k] / 2);
  std::vector<std::array<int, 3>> disp;
  for (int dx = -md[0]; dx <= md[0]; dx++) {
    if (dx == L0 / 2 && L0 % 2 == 0) continue;
    for (int dy = -md[1]; dy <= md[1]; dy++) {
      if (dy == L1 / 2 && L1 % 2 == 0) continue;
      for (int dz = -md[2]; dz <= md[2]; dz++) {
        if (dz == L2 / 2 && L2 % 2 == 0) continue;
        if (dx * dx + dy * dy + dz * dz == r_squared) disp.push_back({dx, dy, dz});
      }
    }
  }
  out[0] = out[1] = 0.0;
  if (disp.empty()) return OK;
  const std::complex<double>* P = reinterpret_cast<const std::complex<double>*>(grid);
  std::complex<double> acc = 0.0;
  long N = 0;
  for (int x = 0; x < L0; x++)
    for (int y = 0; y < L1; y++)
      for (int z = 0; z < L2; z++) {
        const std::complex<double> pm = P[(x * L1 + y) * L2 + z];
        for (const auto& d : disp) {
          int nx = x + d[0], ny = y + d[1], nz = z + d[2];
          if (nx < 0) nx += L0; else if (nx >= L0) nx -= L0;
          if (ny < 0) ny += L1; else if (ny >= L1) ny -= L1;
          if (nz < 0) nz += L2; else if (nz >= L2) nz -= L2;
          acc += pm * std::conj(P[(nx * L1 + ny) * L2 + nz]);
          N++;
        }
      }
  acc /= static_cast<double>(N);
  out[0] = acc.real();
  out[1] = acc.imag();
  return OK;
}

}  // extern "C"

// ------------------------------------------------------------------ heatbath
// Cabibbo-Marinari heatbath with Symanzik staples: Lattice.cpp:509-757 and
// SU3_Sampling.cpp:4-199.  The reference draws its uniforms from a per-thread
// std::mt19937 seeded from std::random_device (rng.cpp:8-27), which no run
// can reproduce; here every uniform comes from a counter-based Philox4x32-10
// stream (Salmon, Moraes, Dror, Shaw, SC'11) addressed by
// (seed; draw pair, link, sweep, tag), in the order the reference's sampler
// consumes them, so the device sweep can be checked draw for draw.
namespace {

struct Philox4x32 {
  uint32_t k0, k1;
  void block(uint32_t c[4]) const {
    uint32_t a = k0, b = k1;
    for (int r = 0; r < 10; ++r) {
      const uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
      const uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
      const uint32_t n0 = uint32_t(p1 >> 32) ^ c[1] ^ a;
      const uint32_t n2 = uint32_t(p0 >> 32) ^ c[3] ^ b;
      c[1] = uint32_t(p1);
      c[3] = uint32_t(p0);
      c[0] = n0;
      c[2] = n2;
      a += 0x9E3779B9u;
      b += 0xBB67AE85u;
    }
  }
};

constexpr uint32_t kHeatbathTag = 0x68623031u;  // separates these streams from the field generators'
// Tries before a rejection loop gives up (the reference's loops have no
// bound and would not end on non-finite links); the device uses the same.
constexpr int kMaxTries = 1 << 20;

// 53-bit uniform in [0, 1) from two 32-bit words.
inline double u53(uint32_t a, uint32_t b) {
  return (double(a >> 5) * 67108864.0 + double(b >> 6)) * (1.0 / 9007199254740992.0);
}

// The uniforms of one link update: draw k is half (k & 1) of Philox block
// {k >> 1, link, sweep, tag}.
struct LinkStream {
  Philox4x32 ph;
  uint32_t link, sweep, q = 0;
  int have = 0;
  double next_u = 0.0;
  bool fail = false;
  LinkStream(unsigned long long seed, uint32_t l, uint32_t s)
      : ph{uint32_t(seed), uint32_t(seed >> 32)}, link(l), sweep(s) {}
  double uniform() {
    if (have) { have = 0; return next_u; }
    uint32_t c[4] = {q++, link, sweep, kHeatbathTag};
    ph.block(c);
    next_u = u53(c[2], c[3]);
    have = 1;
    return u53(c[0], c[1]);
  }
};

constexpr double kPi = 3.14159265358979323846;  // SU3_Sampling.h:10

// SU3_Sampling.cpp:17-29 (Gattringer-Lang): x0 with density
// sqrt(1 - x0^2) exp(a beta x0).
double gen_x0(double a, double beta, LinkStream& rs) {
  for (int tries = 0; tries < kMaxTries; ++tries) {
    const double r1 = 1.0 - rs.uniform();
    const double r2 = 1.0 - rs.uniform();
    const double r3 = 1.0 - rs.uniform();
    const double c = std::cos(2 * kPi * r2);
    const double lam2 = -1 / (2 * a * beta) * (std::log(r1) + c * c * std::log(r3));
    const double x = 1 - lam2;
    const double r = rs.uniform();
    if (r * r <= x && x >= -1 && x <= 1) return 1 - 2 * lam2;
  }
  rs.fail = true;
  return 1.0;
}

// SU3_Sampling.cpp:53-68: unit vector by rejection from the cube [-1, 1)^3.
void unit3(LinkStream& rs, double v[3]) {
  for (int tries = 0; tries < kMaxTries; ++tries) {
    for (int k = 0; k < 3; ++k) v[k] = 2.0 * rs.uniform() - 1.0;
    const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len <= 1.0) {
      for (int k = 0; k < 3; ++k) v[k] /= len;
      return;
    }
  }
  rs.fail = true;
  v[0] = 1.0; v[1] = v[2] = 0.0;
}

// SU(2) as [[x0 + i x3, x2 + i x1], [-x2 + i x1, x0 - i x3]]
// (SU3_Sampling.cpp:82-95, :193-199).
struct M2 { cd a[4]; };
inline M2 su2_of(const double x[4]) {
  return M2{{cd(x[0], x[3]), cd(x[2], x[1]), cd(-x[2], x[1]), cd(x[0], -x[3])}};
}
inline M2 mul2(const M2& x, const M2& y) {
  return M2{{x.a[0] * y.a[0] + x.a[1] * y.a[2], x.a[0] * y.a[1] + x.a[1] * y.a[3],
             x.a[2] * y.a[0] + x.a[3] * y.a[2], x.a[2] * y.a[1] + x.a[3] * y.a[3]}};
}
inline M2 dag2(const M2& x) {
  return M2{{std::conj(x.a[0]), std::conj(x.a[2]), std::conj(x.a[1]), std::conj(x.a[3])}};
}

// SU3_Sampling.cpp:161-170 with :82-95.
M2 su2_heat(double a, double beta, LinkStream& rs, double* x0_out) {
  double x[4];
  x[0] = gen_x0(a, beta, rs);
  double v[3];
  unit3(rs, v);
  const double nrm = std::sqrt(1 - x[0] * x[0]);
  for (int k = 0; k < 3; ++k) x[k + 1] = v[k] * nrm;
  if (x0_out) *x0_out = x[0];
  return su2_of(x);
}

// SU3_Sampling.cpp:175-199.
M2 su2_random(LinkStream& rs) {
  for (int tries = 0; tries < kMaxTries; ++tries) {
    double r[4];
    for (auto& v : r) v = 2.0 * rs.uniform() - 1.0;
    const double len = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
    if (len <= 1.0) {
      for (auto& v : r) v /= len;
      return su2_of(r);
    }
  }
  rs.fail = true;
  const double one[4] = {1.0, 0.0, 0.0, 0.0};
  return su2_of(one);
}

// The subgroups in the reference's update order (Lattice.cpp:509-568,
// :696-723): R = rows/columns (0, 1), S = (0, 2), T = (1, 2).
constexpr int kSub[3][2] = {{0, 1}, {0, 2}, {1, 2}};

// One subgroup update U <- embed(R) U (Lattice.cpp:570-694): the 2 x 2 block
// of U A projected onto the quaternions (Lattice.cpp:570-612), heatbath
// sample X at a = sqrt(det), R = X (block / a)^dagger; below a = 1e-14 a
// Haar-random R.  Returns a and the drawn x0 (0 on the random branch).
M3 subgroup_heatbath(const M3& U, const M3& A, double beta, int type, LinkStream& rs, double* a_out,
                     double* x0_out) {
  if (rs.fail) return U;
  const int i = kSub[type][0], j = kSub[type][1];
  const M3 W = mul(U, A);
  const M2 B{{W.a[3 * i + i], W.a[3 * i + j], W.a[3 * j + i], W.a[3 * j + j]}};
  // b0 = Re tr(B) / 2 and b_l = i Im tr(B sigma_l) / 2
  const double b0 = 0.5 * (B.a[0].real() + B.a[3].real());
  const double k1 = 0.5 * (B.a[1].imag() + B.a[2].imag());
  const double k2 = 0.5 * (B.a[1].real() - B.a[2].real());
  const double k3 = 0.5 * (B.a[0].imag() - B.a[3].imag());
  const double q[4] = {b0, k1, k2, k3};
  const M2 P = su2_of(q);
  const double a = std::sqrt((P.a[0] * P.a[3] - P.a[1] * P.a[2]).real());
  M2 R;
  double x0 = 0.0;
  if (a <= 1e-14) {
    R = su2_random(rs);
  } else {
    M2 Vn = P;
    for (auto& v : Vn.a) v /= a;
    R = mul2(su2_heat(a, 2 * beta / 3, rs, &x0), dag2(Vn));
  }
  M3 out = U;
  for (int c = 0; c < 3; ++c) {
    out.a[3 * i + c] = R.a[0] * U.a[3 * i + c] + R.a[1] * U.a[3 * j + c];
    out.a[3 * j + c] = R.a[2] * U.a[3 * i + c] + R.a[3] * U.a[3 * j + c];
  }
  if (a_out) *a_out = a;
  if (x0_out) *x0_out = x0;
  return out;
}

inline uint32_t link_id(int s, int mu) { return uint32_t(s) * 4u + uint32_t(mu); }

// Lattice.cpp:696-723: staples from the current links, then R, S, T; the
// link's plaquette and action terms ReTr(U A1) / 3 and -(beta/3) ReTr(U A).
// A link whose sampler ran out of tries keeps its value, contributes 0 and
// returns false.
bool single_link_heatbath(const Geo& g, double* U, int s, int mu, double beta, double c11, double c12,
                          unsigned long long seed, uint32_t sweep, double* plaq, double* act) {
  const M3 A1 = staple(g, U, s, mu);
  const M3 A2 = rect_staple(g, U, s, mu);
  const M3 A = add(scale(c11, A1), scale(c12, A2));
  LinkStream rs(seed, link_id(s, mu), sweep);
  M3 u = load_link(U, s, mu);
  for (int t = 0; t < 3; ++t) u = subgroup_heatbath(u, A, beta, t, rs, nullptr, nullptr);
  if (rs.fail) {
    *plaq = *act = 0.0;
    return false;
  }
  store_link(U, s, mu, u);
  *plaq = trace(mul(u, A1)).real() / 3.0;
  *act = -(beta / 3.0) * trace(mul(u, A)).real();
  return true;
}

// Colour of link (s, mu) in the parallel sweep: (2 n_mu + sum of the other
// coordinates) & 3.  The mu-links in the staples of (n, mu) sit at offsets
// +-mu, +-nu, +-2 nu and +-mu +-nu (nu != mu), which change the colour by 2,
// 1, 2 and 1 or 3 (mod 4), so no staple of a link holds another mu-link of its
// colour when every extent is a multiple of 4.  Four colours is the fewest
// possible: links at n, n + mu, n + nu, n + mu + nu are pairwise coupled.
inline int link_colour(const Geo& g, int s, int mu) {
  const int* n = &g.coord[4 * s];
  return (n[0] + n[1] + n[2] + n[3] + n[mu]) & 3;
}

// One sweep, Lattice.cpp:725-757.  order 0: the reference's order (sites
// lexicographic, direction innermost), serial.  order 1: direction by
// direction, each in 4 colour classes, every class in parallel (the device
// order).  Returns the reference's (average plaquette, action density).
bool heatbath_sweep(const Geo& g, double* U, double beta, double c11, double c12, unsigned long long seed,
                    uint32_t sweep, int order, double* avg_plaq, double* act_density) {
  std::vector<double> pl(4 * size_t(g.V)), ac(4 * size_t(g.V));
  int failed = 0;
  if (order == 0) {
    for (int s = 0; s < g.V; ++s)
      for (int mu = 0; mu < 4; ++mu)
        failed += !single_link_heatbath(g, U, s, mu, beta, c11, c12, seed, sweep, &pl[4 * s + mu],
                                        &ac[4 * s + mu]);
  } else {
    for (int mu = 0; mu < 4; ++mu)
      for (int col = 0; col < 4; ++col) {
#pragma omp parallel for schedule(static) reduction(+ : failed)
        for (int s = 0; s < g.V; ++s)
          if (link_colour(g, s, mu) == col)
            failed += !single_link_heatbath(g, U, s, mu, beta, c11, c12, seed, sweep, &pl[4 * s + mu],
                                            &ac[4 * s + mu]);
      }
  }
  double tp = 0.0, ta = 0.0;
  for (size_t l = 0; l < pl.size(); ++l) {
    tp += pl[l];
    ta += ac[l];
  }
  *avg_plaq = (tp / 4.0) / (6.0 * g.V);
  *act_density = (ta / 4.0) / g.V;
  return failed == 0;
}

bool heatbath_dims_ok(const int* d) {
  for (int k = 0; k < 4; ++k)
    if (d[k] < 4 || d[k] % 4) return false;
  return true;
}

}  // namespace

extern "C" {

// Philox4x32-10 block: out = philox(key, ctr).
int oc_philox(const unsigned* key, const unsigned* ctr, unsigned* out) {
  const Philox4x32 ph{key[0], key[1]};
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  ph.block(c);
  for (int k = 0; k < 4; ++k) out[k] = c[k];
  return OK;
}

// n samples of x0 (SU3_Sampling.cpp:17-29) at (a, beta); sample k uses the
// stream of link k at the given sweep.
int oc_heatbath_x0(double a, double beta, unsigned long long seed, unsigned sweep, int n, double* out) {
  if (!(a > 0) || !(beta > 0) || n < 0) return E_INVALID;
#pragma omp parallel for
  for (int k = 0; k < n; ++k) {
    LinkStream rs(seed, uint32_t(k), sweep);
    out[k] = gen_x0(a, beta, rs);
  }
  return OK;
}

// One subgroup update (type 0, 1, 2 = R, S, T) of U18 at staple sum A18,
// with the stream of (link, sweep); returns a and x0 through ax[2].
int oc_heatbath_subgroup(const double* U18, const double* A18, double beta, int type, unsigned long long seed,
                         unsigned link, unsigned sweep, double* out18, double* ax) {
  if (type < 0 || type > 2 || !(beta > 0)) return E_INVALID;
  M3 u, a;
  for (int k = 0; k < 9; ++k) {
    u.a[k] = cd(U18[2 * k], U18[2 * k + 1]);
    a.a[k] = cd(A18[2 * k], A18[2 * k + 1]);
  }
  LinkStream rs(seed, link, sweep);
  const M3 r = subgroup_heatbath(u, a, beta, type, rs, &ax[0], &ax[1]);
  for (int k = 0; k < 9; ++k) {
    out18[2 * k] = r.a[k].real();
    out18[2 * k + 1] = r.a[k].imag();
  }
  return rs.fail ? E_SAMPLER : OK;
}

// nsweeps heatbath sweeps with sweep indices sweep0, sweep0 + 1, ...;
// plaq / act (may be null) receive each sweep's averages.
int oc_heatbath(const int* dims, double* U, double beta, double c11, double c12, unsigned long long seed,
                unsigned sweep0, int nsweeps, int order, double* plaq, double* act) {
  if (!heatbath_dims_ok(dims) || !(beta > 0) || nsweeps < 0 || order < 0 || order > 1) return E_INVALID;
  Geo g(dims);
  bool ok = true;
  for (int k = 0; k < nsweeps; ++k) {
    double p, a;
    ok &= heatbath_sweep(g, U, beta, c11, c12, seed, sweep0 + uint32_t(k), order, &p, &a);
    if (plaq) plaq[k] = p;
    if (act) act[k] = a;
  }
  return ok ? OK : E_SAMPLER;
}

}  // extern "C"

// ------------------------------------------------------------------ HEX smearing
// Hypercubic stout smearing, Smearing.cpp:76-331 (three nested stout levels,
// Smearing.h:7-9 alpha1 = 0.95, alpha2 = 0.76, alpha3 = 0.38).  Two places
// where the reference cannot run are restated as it intends:
//   * each level maps V = exp(proj(alpha/2 Gamma U^dagger)) U with proj the
//     traceless anti-hermitian part (Smearing.cpp:126-129, :143-145); the reference
//     passes that anti-hermitian matrix to its exp(iQ) routine for hermitian
//     Q (Fermions.cpp:843-902), whose guards reject it on any non-trivial
//     links; here exp(A) = exp(iQ) with Q = -iA;
//   * the level-3 staple sum starts at zero (Smearing.cpp:243 leaves it
//     uninitialised).
// The level-1 link V1[mu][nu][rho] depends on (nu, rho) only through the
// one direction sigma outside {mu, nu, rho} that its staples use
// (Smearing.cpp:89), so it is stored once per (mu, sigma): 12 per site.
// Likewise V2[mu][nu] (mu != nu) excludes nu: 12 per site.  Entries with
// repeated indices, which the reference also computes, feed nothing.
namespace {

// Slot of an ordered pair of distinct directions (a, b): 3 a + (b < a ? b : b - 1).
inline int pair_slot(int a, int b) { return 3 * a + (b < a ? b : b - 1); }

inline M3 antiherm_traceless(const M3& m) {
  M3 a = scale(0.5, sub(m, dag(m)));
  const cd t = trace(a) / 3.0;
  a.a[0] -= t; a.a[4] -= t; a.a[8] -= t;
  return a;
}

// exp(proj(alpha/2 Gamma U^dagger)) U (Smearing.cpp:126-129, :143-145, :223-225, :291-293).
int stout_step(const M3& Gamma, const M3& U, double alpha, M3& out) {
  const M3 A = scale(alpha * 0.5, antiherm_traceless(mul(Gamma, dag(U))));
  if (!std::isfinite(fnorm(A))) return E_INVALID;
  M3 e;
  const int st = exp_iq(scale(cd(0.0, -1.0), A), e);
  if (st != OK) return st;
  out = mul(e, U);
  return OK;
}

// Up plus down staple of link (s, mu) in the (mu, sig) plane, paths from s to
// s + mu: L(s, d) is the link of direction d at site s for the pair at hand.
template <class Link>
M3 staple_pair(const Geo& g, int s, int mu, int sig, Link L) {
  const int sp = g.step(s, sig, +1), smu = g.step(s, mu, +1);
  const int sm = g.step(s, sig, -1), smmu = g.step(sm, mu, +1);
  const M3 up = mul(mul(L(s, sig), L(sp, mu)), dag(L(smu, sig)));
  const M3 dn = mul(mul(dag(L(sm, sig)), L(sm, mu)), L(smmu, sig));
  return add(up, dn);
}

int hex_step(const Geo& g, double* U, const double alpha[3]) {
  const size_t V = g.V;
  std::vector<M3> V1(12 * V), V2(12 * V);
  int status = OK;
  // level 1 (Smearing.cpp:76-123, :132-155): V1 of direction mu, staples in sig
#pragma omp parallel for collapse(2)
  for (int s = 0; s < g.V; ++s)
    for (int k = 0; k < 12; ++k) {
      const int mu = k / 3, r = k % 3, sig = r < mu ? r : r + 1;
      auto thin = [&](int x, int d) { return load_link(U, x, d); };
      const M3 G = staple_pair(g, s, mu, sig, thin);
      const int st = stout_step(G, load_link(U, s, mu), alpha[2], V1[12 * s + k]);
      if (st) {
#pragma omp critical
        status = st;
      }
    }
  if (status) return status;
  // level 2 (Smearing.cpp:157-234): V2 of direction mu excluding nu, staples
  // in sig outside {mu, nu} built from level-1 links that exclude the fourth
  // direction rho
#pragma omp parallel for collapse(2)
  for (int s = 0; s < g.V; ++s)
    for (int k = 0; k < 12; ++k) {
      const int mu = k / 3, r = k % 3, nu = r < mu ? r : r + 1;
      M3 G = zero3();
      for (int sig = 0; sig < 4; ++sig) {
        if (sig == mu || sig == nu) continue;
        const int rho = 6 - mu - nu - sig;
        auto lvl1 = [&](int x, int d) { return V1[12 * x + pair_slot(d, rho)]; };
        G = add(G, staple_pair(g, s, mu, sig, lvl1));
      }
      const int st = stout_step(G, load_link(U, s, mu), alpha[1], V2[12 * s + k]);
      if (st) {
#pragma omp critical
        status = st;
      }
    }
  if (status) return status;
  // level 3 (Smearing.cpp:236-302): staples in every nu != mu from level-2
  // links that exclude the other direction of their plane
  std::vector<M3> out(4 * V);
#pragma omp parallel for collapse(2)
  for (int s = 0; s < g.V; ++s)
    for (int mu = 0; mu < 4; ++mu) {
      M3 G = zero3();
      for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        auto lvl2 = [&](int x, int d) { return V2[12 * x + pair_slot(d, d == mu ? nu : mu)]; };
        G = add(G, staple_pair(g, s, mu, nu, lvl2));
      }
      const int st = stout_step(G, load_link(U, s, mu), alpha[0], out[4 * s + mu]);
      if (st) {
#pragma omp critical
        status = st;
      }
    }
  if (status) return status;
  for (int s = 0; s < g.V; ++s)
    for (int mu = 0; mu < 4; ++mu) store_link(U, s, mu, out[4 * s + mu]);
  return OK;
}

}  // namespace

extern "C" {

// nsteps HEX smearing steps in place (Apply_HEX_smearing, Smearing.cpp:304-311;
// Apply_4HEX_smearing :313-318 is nsteps = 4); alpha = {alpha1, alpha2, alpha3}.
int oc_hex_smear(const int* dims, double* U, const double* alpha, int nsteps) {
  if (nsteps < 0) return E_INVALID;
  for (int k = 0; k < 4; ++k)
    if (dims[k] < 2) return E_INVALID;
  Geo g(dims);
  for (int n = 0; n < nsteps; ++n)
    if (int st = hex_step(g, U, alpha)) return st;
  return OK;
}

}  // extern "C"
