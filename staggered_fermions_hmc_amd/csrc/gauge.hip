// Gauge-side kernels: plaquette and rectangle staples, Symanzik gauge force,
// staggered fermion force, the fused momentum kick, the SU(3) exponential and
// link update, and the FP64 action / plaquette / Tr P^2 reductions.
// All FP64, one thread per link.  The staple kernels are FP64-VALU-bound
// (about 85 SU(3) products per link); the fermion force and link update are
// HBM-bound.
#include <cmath>
#include <vector>

#include "internal.h"
#include "reduce.h"
#include "staples.h"
#include "su3exp.h"

namespace sfh {


__device__ __forceinline__ M traceless(const M& f) {
  M r = f;
  const C t = mtrace(f);
  const C t3 = {t.re / 3.0, t.im / 3.0};
  r.e[0] = r.e[0] - t3;
  r.e[4] = r.e[4] - t3;
  r.e[8] = r.e[8] - t3;
  return r;
}

// -i s (A - A^dag) for a real s.
__device__ __forceinline__ M mi_antiherm(const M& A, double s) {
  M r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const C d = A.e[3 * i + j] - conj(A.e[3 * j + i]);
      r.e[3 * i + j] = {s * d.im, -s * d.re};
    }
  return r;
}

// Total force on link (i, mu) and optional P -= coef F (Fermions.cpp:372-384,
// :443-515, :640-659, :690-733).  R12: the links passed the FP64 SU(3)
// check, so the gauge part reads two rows per link and its staple paths take
// the SU(3) product forms; otherwise every link is read whole.
template <bool R12>
__global__ __launch_bounds__(128) void k_force(Geom g, const double2* __restrict__ U, const double2* __restrict__ X,
                                               const double2* __restrict__ Y, double beta, double c11, double c12,
                                               unsigned which, double2* __restrict__ F, double2* __restrict__ P,
                                               double coef, int nb) {
  // the four directions of one site range run as consecutive workgroups of
  // one XCD, so the links around those sites are read from HBM once and
  // from that XCD's L2 for the other directions; the ranges follow the
  // tiled column walk, so x0- and x1-neighbour ranges run close in time
  const int lb = xcd_block(nb);
  const int mu = lb & 3;
  const int i = range_site(g, lb >> 2, threadIdx.x);
  if (i < 0) return;
  const Site4 n = site_of(g, i);
  M f = mzero<double>();
  if (which & SFH_FORCE_GAUGE) {
    M Q = ldLs_after<R12>(U, g, i, mu, 0.0);
    rmul_nf(Q, staples_fused<R12>(U, g, n, mu, c11, c12));
    f = traceless(mi_antiherm(Q, beta / 12.0));
  }
  if (which & SFH_FORCE_FERMION) {
    const M u = ldL(U, g, i, mu);
    Site4 up = n;
    up.move(g, mu, +1);
    const int j = up.index(g);
    const int es = (mu > 0 ? n.x0 : 0) + (mu > 1 ? n.x1 : 0) + (mu > 2 ? n.x2 : 0);
    double c = (es & 1) ? -0.25 : 0.25;
    if (mu == 3 && n.x3 == g.L[3] - 1) c = -c;
    C x0[3], x1[3], y0[3], y1[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double2 v = X[(size_t)a * g.V + i]; x0[a] = {v.x, v.y};
      v = X[(size_t)a * g.V + j]; x1[a] = {v.x, v.y};
      v = Y[(size_t)a * g.V + i]; y0[a] = {v.x, v.y};
      v = Y[(size_t)a * g.V + j]; y1[a] = {v.x, v.y};
    }
    C uy[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      C s = u.e[3 * a] * y1[0];
      cmac(s, u.e[3 * a + 1], y1[1]);
      cmac(s, u.e[3 * a + 2], y1[2]);
      uy[a] = s;
    }
    M af, ab0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        af.e[3 * a + b] = uy[a] * conj(x0[b]);
        ab0.e[3 * a + b] = y0[a] * conj(x1[b]);
      }
    const M ab = mmul_nd(ab0, u);
    const M ff = traceless(mi_antiherm(madd(af, ab), c));
    f = madd(f, ff);
  }
  if (F) stL(F, g, i, mu, f);
  if (P) {
    M pm = ldL(P, g, i, mu);
    pm = msub(pm, mscale(coef, f));
    stL(P, g, i, mu, pm);
  }
}

bool staple_dims_ok(const Geom& g) {
  for (int d = 0; d < 4; ++d)  // the staple walks pack each coordinate into 8 bits
    if (g.L[d] > 255) return false;
  return true;
}

int force_field(sfh_ctx* c, const Geom& g, const double2* U, const double2* X, const double2* Y, double beta,
                double c11, double c12, unsigned which, double2* F, double2* P, double coef, bool r12) {
  if (!staple_dims_ok(g)) return set_err(c, SFH_E_INVALID, "force: lattice extents above 255");
  KTimer kt(c, "force");
  const int nb = 4 * ((g.Vh + kRangeHalf - 1) / kRangeHalf);
  if (r12)
    hipLaunchKernelGGL(k_force<true>, dim3(nb), dim3(2 * kRangeHalf), 0, c->stream, g, U, X, Y, beta, c11, c12,
                       which, F, P, coef, nb);
  else
    hipLaunchKernelGGL(k_force<false>, dim3(nb), dim3(2 * kRangeHalf), 0, c->stream, g, U, X, Y, beta, c11, c12,
                       which, F, P, coef, nb);
  return hip_check(c, hipGetLastError(), "force kernel launch");
}

// ------------------------------------------------------------------ reductions over links
// kind 0: action, 1: plaquette sum ReTr(U A1)/3, 2: ReTr(P^2).  The action
// -(beta/3) [c11 ReTr(U A1)/4 + c12 ReTr(U A2)/6] (Lattice.cpp:759-781) is
// linear in the staple sums, so it is ReTr(U A) of one shared-factor staple
// sum with coefficients c11/4 and c12/6.
template <bool R12>
__global__ __launch_bounds__(128) void k_link_sum(Geom g, const double2* __restrict__ U, double beta, double c11,
                                                  double c12, int kind, double* partials, unsigned* cnt,
                                                  double* out, int nb) {
  double v = 0.0;
  // blocks of 128 links of one direction; the four directions of a site
  // range are consecutive blocks of one XCD (shared L2)
  const int lb = xcd_block(nb);
  const int nvb = 4 * ((g.V + blockDim.x - 1) / blockDim.x);
  for (int vb = lb; vb < nvb; vb += nb) {
    const int mu = vb & 3;
    const int i = (vb >> 2) * blockDim.x + threadIdx.x;
    if (i >= g.V) continue;
    const M u = ldL(U, g, i, mu);
    if (kind == 2) {
      v += retr_mul(u, u);
    } else {
      const Site4 n = site_of(g, i);
      if (kind == 1) v += retr_mul(u, staples_fused<R12>(U, g, n, mu, 1.0, 0.0)) / 3.0;
      else v += -(beta / 3.0) * retr_mul(u, staples_fused<R12>(U, g, n, mu, c11 / 4.0, c12 / 6.0));
    }
  }
  grid_reduce(v, partials, cnt, nb, lb, FinStore{out});
}

static double link_sum(sfh_ctx* c, const Geom& g, const double2* U, double beta, double c11, double c12, int kind,
                       bool r12) {
  if (kind != 2 && !staple_dims_ok(g)) {
    set_err(c, SFH_E_INVALID, "action and plaquette: lattice extents above 255");
    return NAN;
  }
  int nb = (4 * g.V + 127) / 128;
  if (nb > kPartialSlot - kRedGroups) nb = kPartialSlot - kRedGroups;
  KTimer kt(c, "force");
  if (r12)
    hipLaunchKernelGGL(k_link_sum<true>, dim3(nb), dim3(128), 0, c->stream, g, U, beta, c11, c12, kind,
                       c->d_partials + 2 * kPartialSlot, c->aux_ctr(), &c->d_state[1].pv, nb);
  else
    hipLaunchKernelGGL(k_link_sum<false>, dim3(nb), dim3(128), 0, c->stream, g, U, beta, c11, c12, kind,
                       c->d_partials + 2 * kPartialSlot, c->aux_ctr(), &c->d_state[1].pv, nb);
  hipMemcpyAsync(c->h_scalar, &c->d_state[1].pv, sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (hip_check(c, hipStreamSynchronize(c->stream), "link reduction")) return NAN;
  return *c->h_scalar;
}

double gauge_action(sfh_ctx* c, const Geom& g, const double2* U, double beta, double c11, double c12, bool r12) {
  return link_sum(c, g, U, beta, c11, c12, 0, r12);
}
double avg_plaq(sfh_ctx* c, const Geom& g, const double2* U, bool r12) {
  return link_sum(c, g, U, 0, 0, 0, 1, r12) / (4.0 * g.V);
}
double tr_p2(sfh_ctx* c, const Geom& g, const double2* P) { return link_sum(c, g, P, 0, 0, 0, 2, false); }

// ------------------------------------------------------------------ Polyakov loops
// Tr prod_t U_3(x, y, z, t), right-multiplied in t order from t = 0
// (Lattice.cpp:849-866), one thread per spatial site; out is complex
// [(x*L1 + y)*L2 + z] as in precompute_polyakov_grid (Lattice.cpp:870-884).
__global__ void k_polyakov(Geom g, const double2* __restrict__ U, double2* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int ns = g.L[0] * g.L[1] * g.L[2];
  if (s >= ns) return;
  Site4 x{s / (g.L[2] * g.L[1]), (s / g.L[2]) % g.L[1], s % g.L[2], 0};
  M w = ldL(U, g, x.index(g), 3);
  for (int t = 1; t < g.L[3]; ++t) {
    x.x3 = t;
    rmul_n(w, ldL(U, g, x.index(g), 3));
  }
  const C tr = mtrace(w);
  out[s] = make_double2(tr.re, tr.im);
}

static double2* polyakov_dev(sfh_ctx* c, const Geom& g, const double2* U) {
  const int ns = g.L[0] * g.L[1] * g.L[2];
  double2* d = static_cast<double2*>(scratch(c, "polyakov", ns * sizeof(double2)));
  if (d) hipLaunchKernelGGL(k_polyakov, dim3((ns + 127) / 128), dim3(128), 0, c->stream, g, U, d);
  return d;
}

int polyakov_grid(sfh_ctx* c, const Geom& g, const double2* U, double* out) {
  const int ns = g.L[0] * g.L[1] * g.L[2];
  double2* d = polyakov_dev(c, g, U);
  if (!d) return SFH_E_NOMEM;
  hipMemcpyAsync(out, d, ns * sizeof(double2), hipMemcpyDeviceToHost, c->stream);
  return hip_check(c, hipStreamSynchronize(c->stream), "polyakov");
}

// ------------------------------------------------------------------ Polyakov-loop correlators
// C(r^2) = (1/N) sum_m sum_{d in D(r^2)} P(m) conj(P(m + d)), N = |grid| |D(r^2)|,
// with D(r^2) the integer displacements of correlator_over_fixed_distance_fast
// (Lattice.cpp:886-948): |d_k| <= min(floor(sqrt(r^2)), L_k / 2), d_k = +L_k / 2
// left out (it is the same neighbour as -L_k / 2), d . d = r^2; C = 0 when D
// is empty.  One lane per grid point m and one block row per r^2; each lane
// sums its displacements in table order, blocks reduce in a fixed order, and a
// second kernel sums the block partials of each r^2 in block order, so the
// result does not depend on timing.
constexpr int kCorrB = 256;

__global__ __launch_bounds__(kCorrB) void k_poly_corr(int L0, int L1, int L2, const double2* __restrict__ P,
                                                     const int4* __restrict__ disp, const int* __restrict__ off,
                                                     double* __restrict__ part) {
  const int r = blockIdx.y;
  const int ns = L0 * L1 * L2;
  const int m = blockIdx.x * kCorrB + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (m < ns) {
    const int x = m / (L1 * L2), y = (m / L2) % L1, z = m % L2;
    const double2 pm = P[m];
    for (int k = off[r]; k < off[r + 1]; ++k) {
      const int4 d = disp[k];
      int nx = x + d.x, ny = y + d.y, nz = z + d.z;
      nx += nx < 0 ? L0 : (nx >= L0 ? -L0 : 0);
      ny += ny < 0 ? L1 : (ny >= L1 ? -L1 : 0);
      nz += nz < 0 ? L2 : (nz >= L2 ? -L2 : 0);
      const double2 pn = P[(nx * L1 + ny) * L2 + nz];
      v[0] += pm.x * pn.x + pm.y * pn.y;  // Re P(m) conj(P(n))
      v[1] += pm.y * pn.x - pm.x * pn.y;  // Im
    }
  }
  block_sum_n<2>(v);
  if (threadIdx.x == 0) {
    part[2 * ((size_t)r * gridDim.x + blockIdx.x)] = v[0];
    part[2 * ((size_t)r * gridDim.x + blockIdx.x) + 1] = v[1];
  }
}

__global__ __launch_bounds__(kCorrB) void k_poly_corr_fin(const double* __restrict__ part, int nblk,
                                                         const int* __restrict__ off, int ns,
                                                         double* __restrict__ out) {
  const int r = blockIdx.x;
  double v[2] = {0.0, 0.0};
  for (int b = threadIdx.x; b < nblk; b += kCorrB) {
    v[0] += part[2 * ((size_t)r * nblk + b)];
    v[1] += part[2 * ((size_t)r * nblk + b) + 1];
  }
  block_sum_n<2>(v);
  if (threadIdx.x == 0) {
    const double n = (double)ns * (double)(off[r + 1] - off[r]);
    out[2 * r] = n > 0 ? v[0] / n : 0.0;
    out[2 * r + 1] = n > 0 ? v[1] / n : 0.0;
  }
}

// Displacement tables of r^2 = r2_first .. r2_first + count - 1 in the
// reference's loop order (dx, then dy, then dz ascending).
void corr_tables(const int L[3], int r2_first, int count, std::vector<int4>& disp, std::vector<int>& off) {
  disp.clear();
  off.assign(1, 0);
  for (int r2 = r2_first; r2 < r2_first + count; ++r2) {
    int md[3];
    for (int k = 0; k < 3; ++k) {
      int s = (int)std::sqrt((double)r2);
      while ((s + 1) * (s + 1) <= r2) ++s;
      while (s * s > r2) --s;
      md[k] = s < L[k] / 2 ? s : L[k] / 2;
    }
    for (int dx = -md[0]; dx <= md[0]; ++dx) {
      if (L[0] % 2 == 0 && dx == L[0] / 2) continue;
      for (int dy = -md[1]; dy <= md[1]; ++dy) {
        if (L[1] % 2 == 0 && dy == L[1] / 2) continue;
        for (int dz = -md[2]; dz <= md[2]; ++dz) {
          if (L[2] % 2 == 0 && dz == L[2] / 2) continue;
          if (dx * dx + dy * dy + dz * dz == r2) disp.push_back(make_int4(dx, dy, dz, 0));
        }
      }
    }
    off.push_back((int)disp.size());
  }
}

void polyakov_grid_into(sfh_ctx* c, const Geom& g, const double2* U, double2* P) {
  const int ns = g.L[0] * g.L[1] * g.L[2];
  hipLaunchKernelGGL(k_polyakov, dim3((ns + 127) / 128), dim3(128), 0, c->stream, g, U, P);
}

int poly_corr_blocks(const int L[3]) { return (L[0] * L[1] * L[2] + kCorrB - 1) / kCorrB; }

// Correlators of the device grid P (L[0] x L[1] x L[2] complex) with the
// tables disp / off into out[2 k], out[2 k + 1] = C(r2_first + k), k < count
// (host memory), through the device buffers d_disp (>= disp.size() entries),
// d_off (count + 1), d_part (2 count poly_corr_blocks(L) doubles) and d_out
// (2 count).  Synchronises the stream.
int poly_corr_run(sfh_ctx* c, const int L[3], const double2* P, const std::vector<int4>& disp,
                  const std::vector<int>& off, int4* d_disp, int* d_off, double* d_part, double* d_out, double* out) {
  const int count = (int)off.size() - 1;
  const int ns = L[0] * L[1] * L[2];
  const int nblk = poly_corr_blocks(L);
  if (!disp.empty())
    hipMemcpyAsync(d_disp, disp.data(), disp.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(d_off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
  hipLaunchKernelGGL(k_poly_corr, dim3(nblk, count), dim3(kCorrB), 0, c->stream, L[0], L[1], L[2], P, d_disp, d_off,
                     d_part);
  hipLaunchKernelGGL(k_poly_corr_fin, dim3(count), dim3(kCorrB), 0, c->stream, d_part, nblk, d_off, ns, d_out);
  hipMemcpyAsync(out, d_out, 2 * (size_t)count * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  // the host tables stay alive until the stream has drained
  return hip_check(c, hipStreamSynchronize(c->stream), "polyakov correlators");
}

static int poly_corr_dev(sfh_ctx* c, const int L[3], const double2* P, int r2_first, int count, double* out) {
  std::vector<int4> disp;
  std::vector<int> off;
  corr_tables(L, r2_first, count, disp, off);
  const int nblk = poly_corr_blocks(L);
  int4* d_disp = static_cast<int4*>(scratch(c, "pc_disp", (disp.size() + 1) * sizeof(int4)));
  int* d_off = static_cast<int*>(scratch(c, "pc_off", off.size() * sizeof(int)));
  double* d_part = static_cast<double*>(scratch(c, "pc_part", 2 * (size_t)count * nblk * sizeof(double)));
  double* d_out = static_cast<double*>(scratch(c, "pc_out", 2 * (size_t)count * sizeof(double)));
  if (!d_disp || !d_off || !d_part || !d_out) return SFH_E_NOMEM;
  return poly_corr_run(c, L, P, disp, off, d_disp, d_off, d_part, d_out, out);
}

int polyakov_correlators(sfh_ctx* c, const Geom& g, const double2* U, int r2_first, int count, double* out) {
  double2* P = polyakov_dev(c, g, U);
  if (!P) return SFH_E_NOMEM;
  return poly_corr_dev(c, g.L, P, r2_first, count, out);
}

int polyakov_correlators_grid(sfh_ctx* c, const int L[3], const double* grid, int r2_first, int count, double* out) {
  const int ns = L[0] * L[1] * L[2];
  double2* P = static_cast<double2*>(scratch(c, "polyakov", ns * sizeof(double2)));
  if (!P) return SFH_E_NOMEM;
  hipMemcpyAsync(P, grid, ns * sizeof(double2), hipMemcpyHostToDevice, c->stream);
  return poly_corr_dev(c, L, P, r2_first, count, out);
}

// ------------------------------------------------------------------ exponential and link update
__global__ __launch_bounds__(128) void k_link_update(Geom g, double2* __restrict__ U, const double2* __restrict__ P,
                                                     double eps, int check, int* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (i >= g.V) return;
  const M q = mscale(eps, ldL(P, g, i, mu));
  if (check) {
    double hn = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const C d = q.e[3 * a + b] - conj(q.e[3 * b + a]);
        hn += d.re * d.re + d.im * d.im;
      }
    const C t = mtrace(q);
    if (sqrt(hn) > 1e-10 || sqrt(t.re * t.re + t.im * t.im) > 1e-10) {
      atomicMax(status, SFH_E_NOT_HERMITIAN);
      return;
    }
  }
  M e;
  const int st = exp_iq(q, e);
  if (st != SFH_OK) {
    atomicMax(status, st);
    return;
  }
  stL(U, g, i, mu, mmul(e, ldL(U, g, i, mu)));
}

int link_update(sfh_ctx* c, const Geom& g, double2* U, const double2* P, double eps, int check) {
  int* d_status = &c->d_state[1].status;
  hipMemsetAsync(d_status, 0, sizeof(int), c->stream);
  {
    KTimer kt(c, "link_update");
    dim3 grid((g.V + 127) / 128, 4);
    hipLaunchKernelGGL(k_link_update, grid, dim3(128), 0, c->stream, g, U, P, eps, check, d_status);
  }
  int* h = reinterpret_cast<int*>(c->h_scalar);
  hipMemcpyAsync(h, d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (int e = hip_check(c, hipStreamSynchronize(c->stream), "link update")) return e;
  if (*h == SFH_E_NOT_HERMITIAN) return set_err(c, *h, "P_pre is not Hermitian or not traceless");
  if (*h == SFH_E_EXP_QCUBED) return set_err(c, *h, "Error: Q cubed matrix has non-zero imaginary trace.");
  if (*h == SFH_E_EXP_NEG_TRACE) return set_err(c, *h, "Error: Q squared matrix has negative trace.");
  return SFH_OK;
}

__global__ void k_exp_one(const double* Q18, double* out18, int* status) {
  M q, e;
  for (int k = 0; k < 9; ++k) q.e[k] = {Q18[2 * k], Q18[2 * k + 1]};
  *status = exp_iq(q, e);
  for (int k = 0; k < 9; ++k) {
    out18[2 * k] = e.e[k].re;
    out18[2 * k + 1] = e.e[k].im;
  }
}

int exp_iq_one(sfh_ctx* c, const double* Q18, double* out18) {
  double* d = static_cast<double*>(scratch(c, "exp_one", 40 * sizeof(double)));
  if (!d) return SFH_E_NOMEM;
  int* dst = reinterpret_cast<int*>(d + 36);
  hipMemcpyAsync(d, Q18, 18 * sizeof(double), hipMemcpyHostToDevice, c->stream);
  hipLaunchKernelGGL(k_exp_one, dim3(1), dim3(1), 0, c->stream, d, d + 18, dst);
  hipMemcpyAsync(out18, d + 18, 18 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  int* h = reinterpret_cast<int*>(c->h_scalar);
  hipMemcpyAsync(h, dst, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (int e = hip_check(c, hipStreamSynchronize(c->stream), "exp")) return e;
  if (*h == SFH_E_EXP_QCUBED) return set_err(c, *h, "Error: Q cubed matrix has non-zero imaginary trace.");
  if (*h == SFH_E_EXP_NEG_TRACE) return set_err(c, *h, "Error: Q squared matrix has negative trace.");
  return SFH_OK;
}

}  // namespace sfh
