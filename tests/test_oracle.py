"""Pins the FP64 CPU oracle (oracle/oracle.cpp) before it is trusted as the
checker of the HIP path.  Each check is chosen so that a plausible slip in
the restatement (dropped term, wrong sign or index, transposed operand)
fails: known answers of the reference's own checks (ai_check.cpp:172-333),
closed forms, invariants, brute force and finite differences.  CPU only."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as oc
import fields as F

DIMS4 = (4, 4, 4, 4)
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def eta(n, mu):
    return (-1) ** int(sum(n[:mu]))


# ---------------------------------------------------------------- known answers
def test_cold_staples_known_answer():
    """ai_check.cpp:172-223: identity links give A1 = 6 I and A2 = 18 I."""
    ka = json.load(open(os.path.join(GOLDEN, "known_answers.json")))
    U = F.cold_links(DIMS4)
    for site in (0, 37, 255):
        for mu in range(4):
            A1, A2 = oc.staples(DIMS4, U, site, mu)
            A1 = A1.view(np.complex128).reshape(3, 3)
            A2 = A2.view(np.complex128).reshape(3, 3)
            assert np.linalg.norm(A1 - ka["cold_A1_diag"] * np.eye(3)) < 1e-12
            assert np.linalg.norm(A2 - ka["cold_A2_diag"] * np.eye(3)) < 1e-12


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 4, 4, 8), (6, 4, 4, 2)])
@pytest.mark.parametrize("c11,c12", [(5.0 / 3.0, -1.0 / 12.0), (1.0, 0.0)])
def test_cold_action_known_answer(dims, c11, c12):
    """ai_check.cpp:227-333: S_cold = -beta (6V c11 + 12V c12)."""
    beta = 4.1
    V = F.volume(dims)
    S = oc.action(dims, F.cold_links(dims), beta, c11, c12)
    expect = -beta * (6 * V * c11 + 12 * V * c12)
    # ai_check.cpp:323 uses 1e-10 absolute at 4^4; a serial sum of 4V equal
    # per-link terms carries ~1e-14 relative rounding, so scale with |S|.
    assert abs(S - expect) < 1e-13 * abs(expect)


def test_cold_avg_plaq():
    """Lattice.cpp:784-804 normalises to 6 on a cold lattice."""
    assert abs(oc.avg_plaq(DIMS4, F.cold_links(DIMS4)) - 6.0) < 1e-13


# ---------------------------------------------------------------- gauge action
@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 6, 4, 2)])
def test_action_matches_brute_force_loops(dims):
    """Every plaquette and rectangle enumerated once, by walking closed paths."""
    U = F.random_links(dims, 0.4, seed=11)
    for c11, c12 in [(1.0, 0.0), (0.0, 1.0), (5.0 / 3.0, -1.0 / 12.0)]:
        S_bf, sp, sr = F.brute_force_action(dims, U, 5.6, c11, c12)
        S = oc.action(dims, U, 5.6, c11, c12)
        assert abs(S - S_bf) < 1e-10 * max(1.0, abs(S_bf))
    # avg plaquette: (1/(3V)) sum_plaq ReTr in the reference's normalisation
    assert abs(oc.avg_plaq(dims, U) - sp / (3.0 * F.volume(dims))) < 1e-12


def test_action_gauge_invariant():
    U = F.random_links(DIMS4, 0.5, seed=3)
    Ug, _G = F.gauge_transform(DIMS4, U, None, seed=4)
    for c in [(1.0, 0.0), (5.0 / 3.0, -1.0 / 12.0)]:
        assert abs(oc.action(DIMS4, U, 4.0, *c) - oc.action(DIMS4, Ug, 4.0, *c)) < 1e-10


# ---------------------------------------------------------------- Dirac operator
def test_dirac_delta_source_stencil():
    """Hand-derived scatter form of the stencil from a delta source, at the
    APBC corner (0,0,0,T-1) and an eta-sensitive site (1,1,1,0)."""
    dims = DIMS4
    T = dims[3]
    m = 0.37
    U = F.random_links(dims, 0.6, seed=5)
    Um = F.flat_to_mat(U)
    e = np.eye(4, dtype=int)
    for n0 in [(0, 0, 0, T - 1), (1, 1, 1, 0), (2, 3, 1, 2)]:
        n0 = np.array(n0)
        for c0 in range(3):
            z = 0.3 - 0.8j
            psi = np.zeros((F.volume(dims), 3), complex)
            s0 = F.site_index(dims, n0)
            psi[s0, c0] = z
            for dagger in (False, True):
                sg = -1.0 if dagger else 1.0
                exp = np.zeros_like(psi)
                exp[s0, c0] += m * z
                for mu in range(4):
                    # site n = n0 - mu sees the source as its forward neighbour
                    n = n0 - e[mu]
                    b_up = -1.0 if (mu == 3 and n[3] == -1) else 1.0
                    n = n % np.array(dims)
                    s = F.site_index(dims, n)
                    exp[s] += sg * 0.5 * eta(n, mu) * b_up * Um[s, mu][:, c0] * z
                    # site n = n0 + mu sees the source as its backward neighbour
                    n = n0 + e[mu]
                    b_dn = -1.0 if (mu == 3 and n[3] == T) else 1.0
                    n = n % np.array(dims)
                    s = F.site_index(dims, n)
                    exp[s] -= sg * 0.5 * eta(n, mu) * b_dn * Um[s0, mu].conj().T[:, c0] * z
                got = F.flat_to_vec(oc.apply_d(dims, m, U, F.vec_to_flat(psi), dagger=dagger))
                assert np.abs(got - exp).max() < 1e-15


def test_hop_antihermitian_and_eps_anticommutes():
    dims = (4, 4, 4, 6)
    U = F.random_links(dims, 0.5, seed=7)
    a = F.gaussian_spinor(dims, 1)
    b = F.gaussian_spinor(dims, 2)
    ca, cb = F.flat_to_vec(a), F.flat_to_vec(b)
    Ka = F.flat_to_vec(oc.apply_d(dims, 0.0, U, a))
    Kb = F.flat_to_vec(oc.apply_d(dims, 0.0, U, b))
    assert abs(np.vdot(ca, Kb) + np.vdot(Ka, cb)) < 1e-12
    # D^dagger = eps D eps with eps(n) = (-1)^{sum n}
    es = F.eps_sign(dims)[:, None]
    m = 0.2
    Dd = F.flat_to_vec(oc.apply_d(dims, m, U, a, dagger=True))
    eDe = es * F.flat_to_vec(oc.apply_d(dims, m, U, F.vec_to_flat(es * ca)))
    assert np.abs(Dd - eDe).max() < 1e-14
    # D^dagger is the adjoint of D
    Db = F.flat_to_vec(oc.apply_d(dims, m, U, b))
    Dda = F.flat_to_vec(oc.apply_d(dims, m, U, a, dagger=True))
    assert abs(np.vdot(ca, Db) - np.vdot(Dda, cb)) < 1e-12


def test_dirac_gauge_covariant():
    U = F.random_links(DIMS4, 0.5, seed=8)
    psi = F.gaussian_spinor(DIMS4, 9)
    Ug, psig, G = F.gauge_transform(DIMS4, U, psi, seed=10)
    D = F.flat_to_vec(oc.apply_d(DIMS4, 0.1, U, psi))
    Dg = F.flat_to_vec(oc.apply_d(DIMS4, 0.1, Ug, psig))
    assert np.abs(Dg - np.einsum("vij,vj->vi", G, D)).max() < 1e-13


@pytest.mark.parametrize("k", [(0, 0, 0, 0), (1, 2, 0, 1), (3, 1, 2, 3)])
def test_free_field_plane_wave(k):
    """U = 1: D D^dag psi = (m^2 + sum_mu sin^2 p_mu) psi for p_i = 2 pi k_i/L,
    p_3 = (2 k_3 + 1) pi / T (antiperiodic)."""
    dims = (4, 4, 4, 8)
    m = 0.3
    p = np.array([2 * np.pi * k[i] / dims[i] for i in range(3)] + [(2 * k[3] + 1) * np.pi / dims[3]])
    c = F.coords(dims)
    color = np.array([0.3 + 0.1j, -0.5j, 0.8])
    psi = np.exp(1j * (c @ p))[:, None] * color[None, :]
    out = oc.apply_ddag(dims, m, F.cold_links(dims), F.vec_to_flat(psi))
    lam = m * m + np.sum(np.sin(p) ** 2)
    assert np.abs(F.flat_to_vec(out) - lam * psi).max() < 1e-13
    X, iters, _, _ = oc.cg(dims, m, F.cold_links(dims), F.vec_to_flat(psi), tol=1e-12)
    assert iters == 1
    assert np.abs(F.flat_to_vec(X) - psi / lam).max() < 1e-12


# ---------------------------------------------------------------- CG
def test_cg_against_dense_solve():
    dims = (2, 2, 2, 4)
    m = 0.1
    U = F.random_links(dims, 0.5, seed=12)
    V = F.volume(dims)
    n = 6 * V
    A = np.empty((n, n))
    for i in range(n):
        e = np.zeros(n); e[i] = 1.0
        A[:, i] = oc.apply_ddag(dims, m, U, e)
    assert np.abs(A - A.T).max() < 1e-13  # D D^dag symmetric as a real operator
    assert np.linalg.eigvalsh(A).min() > m * m * (1 - 1e-12)
    b = F.gaussian_spinor(dims, 13)
    X, iters, rsq, st = oc.cg(dims, m, U, b, tol=1e-12)
    assert st == 0
    xd = np.linalg.solve(A, b)
    assert np.abs(X - xd).max() < 1e-9 * np.abs(xd).max()


def test_cg_pseudofermion_identities():
    """Y = D^dag (D D^dag)^-1 D chi = chi and S_f = Re<D chi, X> = |chi|^2."""
    dims = DIMS4
    m = 0.1
    U = F.random_links(dims, 0.3, seed=14)
    chi = F.gaussian_spinor(dims, 15)
    phi = oc.apply_d(dims, m, U, chi)
    X, iters, rsq, st = oc.cg(dims, m, U, phi, tol=1e-12)
    Y = oc.apply_d(dims, m, U, X, dagger=True)
    assert np.abs(Y - chi).max() < 1e-9
    assert abs(oc.inner(phi, X) - oc.inner(chi, chi)) < 1e-9 * oc.inner(chi, chi)
    r = phi - oc.apply_ddag(dims, m, U, X)
    assert np.sqrt(oc.inner(r, r) / oc.inner(phi, phi)) < 1e-11


def test_cg_error_paths():
    dims = DIMS4
    U = F.random_links(dims, 0.3, seed=16)
    phi = F.gaussian_spinor(dims, 17)
    # zero source: silent early return (Fermions.cpp:562-564)
    X, it, _, st = oc.cg(dims, 0.1, U, np.zeros_like(phi))
    assert st == 0 and it == 0 and not X.any()
    # non-convergence (Fermions.cpp:609-614)
    _, it, _, st = oc.cg(dims, 0.1, U, phi, max_iter=3, raise_on_error=False)
    assert st == 8 and it == 3
    # NaN source: p.v is non-finite -> breakdown (Fermions.cpp:573-578)
    bad = phi.copy(); bad[5] = np.nan
    _, _, _, st = oc.cg(dims, 0.1, U, bad, raise_on_error=False)
    assert st in (5, 8)


# ---------------------------------------------------------------- exponential
def _herm_traceless(rng):
    return F.random_algebra(rng, ())


@pytest.mark.parametrize("scale", [1e-7, 1e-3, 0.05, 0.7, 3.0])
def test_exp_matches_eigendecomposition(scale):
    rng = np.random.default_rng(int(scale * 1e7) % 1000 + 1)
    for _ in range(20):
        Q = scale * _herm_traceless(rng)
        got = oc.exp_iq(F.mat_to_flat(Q)).view(np.complex128).reshape(3, 3)
        ref = F.expi_hermitian(Q)
        assert np.abs(got - ref).max() < 5e-14
        assert np.abs(got @ got.conj().T - np.eye(3)).max() < 1e-13
        assert abs(np.linalg.det(got) - 1) < 1e-13


def test_exp_degenerate_and_branches():
    for Q in [np.diag([0.4, 0.4, -0.8]), np.diag([-0.4, -0.4, 0.8]), np.zeros((3, 3)),
              np.diag([1e-8, -1e-8, 0.0])]:
        Q = Q.astype(complex)
        got = oc.exp_iq(F.mat_to_flat(Q)).view(np.complex128).reshape(3, 3)
        assert np.abs(got - F.expi_hermitian(Q)).max() < 1e-14
    # non-hermitian input trips the reference's guard (Fermions.cpp:848-850)
    bad = np.diag([0.5j, 0.1, -0.1])
    with pytest.raises(oc.OracleError):
        oc.exp_iq(F.mat_to_flat(bad))


def test_link_update_guard_and_value():
    dims = (2, 2, 2, 2)
    U = F.random_links(dims, 0.3, seed=20)
    P = F.random_momenta(dims, 21)
    eps = 0.1
    U2 = oc.link_update(dims, eps, P, U)
    ref = np.einsum("lij,ljk->lik", F.expi_hermitian(eps * F.flat_to_mat(P).reshape(-1, 3, 3)),
                    F.flat_to_mat(U).reshape(-1, 3, 3))
    assert np.abs(F.flat_to_mat(U2).reshape(-1, 3, 3) - ref).max() < 1e-13
    Pbad = P.copy(); Pbad[0] += 1.0  # Re P_00 += 1 makes P non-traceless
    with pytest.raises(oc.OracleError):
        oc.link_update(dims, eps, Pbad, U)


# ---------------------------------------------------------------- forces
def _fd_check(dims, action_fn, F_field, U, links, eps=1e-4, rtol=2e-6):
    rng = np.random.default_rng(99)
    Fm = F.flat_to_mat(F_field)
    for (s, mu) in links:
        H = _herm_traceless(rng)
        Um = F.flat_to_mat(U).copy()
        base = Um[s, mu].copy()
        Um[s, mu] = F.expi_hermitian(eps * H) @ base
        Sp = action_fn(F.mat_to_flat(Um))
        Um[s, mu] = F.expi_hermitian(-eps * H) @ base
        Sm = action_fn(F.mat_to_flat(Um))
        num = (Sp - Sm) / (2 * eps)
        ana = 2.0 * np.trace(Fm[s, mu] @ H).real
        assert abs(num - ana) < rtol * max(1.0, abs(ana)), (s, mu, num, ana)
        # hermitian and traceless
        assert np.abs(Fm[s, mu] - Fm[s, mu].conj().T).max() < 1e-13
        assert abs(np.trace(Fm[s, mu])) < 1e-13


@pytest.mark.parametrize("c11,c12", [(1.0, 0.0), (0.0, 1.0), (5.0 / 3.0, -1.0 / 12.0)])
def test_gauge_force_is_action_gradient(c11, c12):
    """ai_check.cpp:509-660: dS/de = 2 ReTr(F H) under U -> exp(i e H) U."""
    dims = (4, 4, 4, 4)
    beta = 4.0
    U = F.random_links(dims, 0.5, seed=22)
    Fg = oc.force(dims, U, None, None, beta, c11, c12, which=1)
    links = [(0, 0), (77, 1), (130, 2), (255, 3), (200, 0)]
    _fd_check(dims, lambda u: oc.action(dims, u, beta, c11, c12), Fg, U, links)


def test_fermion_force_is_action_gradient():
    """S_f = Re<phi, (D D^dag)^-1 phi>; dS_f/de = 2 ReTr(F_f H) (App. A.7)."""
    dims = (4, 4, 4, 4)
    m = 0.2
    U = F.random_links(dims, 0.4, seed=23)
    phi = F.gaussian_spinor(dims, 24)
    X, *_ = oc.cg(dims, m, U, phi, tol=1e-14, max_iter=10000)
    Y = oc.apply_d(dims, m, U, X, dagger=True)
    Ff = oc.force(dims, U, X, Y, 0.0, 0.0, 0.0, which=2)

    def Sf(u):
        x, *_ = oc.cg(dims, m, u, phi, tol=1e-14, max_iter=10000)
        return oc.inner(phi, x)

    links = [(0, 3), (63, 3), (100, 1), (255, 0), (3, 3)]
    _fd_check(dims, Sf, Ff, U, links, eps=1e-4, rtol=1e-6)


def test_forces_gauge_covariant():
    dims = DIMS4
    m = 0.2
    U = F.random_links(dims, 0.4, seed=25)
    phi = F.gaussian_spinor(dims, 26)
    X, *_ = oc.cg(dims, m, U, phi)
    Y = oc.apply_d(dims, m, U, X, dagger=True)
    Ft = oc.force(dims, U, X, Y, 4.0, 5 / 3, -1 / 12, which=3)
    Ug, Xg, G = F.gauge_transform(dims, U, X, seed=27)
    _, Yg, _ = F.gauge_transform(dims, U, Y, seed=27)
    Ftg = oc.force(dims, Ug, Xg, Yg, 4.0, 5 / 3, -1 / 12, which=3)
    Fm, Fgm = F.flat_to_mat(Ft), F.flat_to_mat(Ftg)
    expect = np.einsum("vij,vmjk,vlk->vmil", G, Fm, G.conj())
    assert np.abs(Fgm - expect).max() < 1e-11


def test_polyakov_grid_against_numpy_walk():
    """Lattice.cpp:849-884: product of the T time links at each spatial site,
    right-multiplied from t = 0; cold links give Tr 1 = 3."""
    dims = (4, 6, 4, 8)
    U = F.random_links(dims, 0.5, seed=60)
    Um = F.flat_to_mat(U)
    ref = np.empty(dims[0] * dims[1] * dims[2], complex)
    for k, (x, y, z) in enumerate(np.ndindex(*dims[:3])):
        W = np.eye(3, dtype=complex)
        for t in range(dims[3]):
            W = W @ Um[F.site_index(dims, np.array([x, y, z, t])), 3]
        ref[k] = np.trace(W)
    assert np.abs(oc.polyakov_grid(dims, U) - ref).max() < 1e-13
    assert np.abs(oc.polyakov_grid(dims, F.cold_links(dims)) - 3.0).max() == 0.0
    # gauge invariant
    Ug, _ = F.gauge_transform(dims, U, None, seed=61)
    assert np.abs(oc.polyakov_grid(dims, Ug) - ref).max() < 1e-12


def test_tr_p2():
    P = F.random_momenta(DIMS4, 30)
    Pm = F.flat_to_mat(P)
    assert abs(oc.tr_p2(DIMS4, P) - np.einsum("lmij,lmji->", Pm, Pm).real) < 1e-10


# ---------------------------------------------------------------- integrator
def test_leapfrog_reversible():
    """Fermions.cpp:908-955: forward, flip P, forward returns to the start."""
    dims = DIMS4
    m = 0.1
    U0 = F.random_links(dims, 0.3, seed=31)
    P0 = F.random_momenta(dims, 32)
    phi = oc.apply_d(dims, m, U0, F.gaussian_spinor(dims, 33))
    kw = dict(beta=5.6, c11=1.0, c12=0.0, eps=0.1, nsteps=5, tol=1e-13)
    U1, P1, _ = oc.leapfrog(dims, m, U0, P0, phi, **kw)
    U2, P2, _ = oc.leapfrog(dims, m, U1, -P1, phi, **kw)
    assert np.abs(U2 - U0).max() < 1e-10
    assert np.abs(-P2 - P0).max() < 1e-10


@pytest.mark.parametrize("which", [1, 2])
def test_dH_scales_as_eps_squared(which):
    """Fermions.cpp:1230-1273 and :1378-1418: |dH(e)|/|dH(e/2)| in (3.5, 4.5)."""
    dims = DIMS4
    m = 0.2
    U0 = F.random_links(dims, 0.3, seed=40)
    P0 = F.random_momenta(dims, 41)
    phi = oc.apply_d(dims, m, U0, F.gaussian_spinor(dims, 42))
    beta, c11, c12 = 4.0, 5 / 3, -1 / 12

    def H(U, P):
        h = oc.tr_p2(dims, P)
        if which & 1:
            h += oc.action(dims, U, beta, c11, c12)
        if which & 2:
            X, *_ = oc.cg(dims, m, U, phi, tol=1e-13)
            h += oc.inner(phi, X)
        return h

    H0 = H(U0, P0)
    dH = []
    for eps, n in [(0.05, 10), (0.025, 20)]:
        U1, P1, _ = oc.leapfrog(dims, m, U0, P0, phi, beta=beta, c11=c11, c12=c12, eps=eps, nsteps=n,
                                which=which, tol=1e-13, check=False)
        dH.append(abs(H(U1, P1) - H0))
    assert 3.5 < dH[0] / dH[1] < 4.5


def test_hmc_trajectory_energy_bookkeeping():
    dims = DIMS4
    m = 0.1
    U0 = F.random_links(dims, 0.3, seed=50)
    P0 = F.random_momenta(dims, 51)
    chi = F.gaussian_spinor(dims, 52)
    U, H, acc = oc.hmc_trajectory(dims, m, U0, chi, P0, 0.0, beta=5.6, c11=1.0, c12=0.0, eps=0.1, nsteps=10,
                                  tol=1e-12)
    # S_f(initial) = |chi|^2 exactly when solved (App. A.5)
    assert abs(H[2] - oc.inner(chi, chi)) < 1e-8 * H[2]
    assert abs(H[0] - oc.tr_p2(dims, P0)) < 1e-12 * H[0]
    assert abs(H[1] - oc.action(dims, U0, 5.6, 1.0, 0.0)) < 1e-9
    assert abs(H[6] - (H[0] - H[3] + H[1] - H[4] + H[2] - H[5])) < 1e-9
    assert acc == (0.0 < np.exp(H[6]))


def test_leapfrog_stage_order_momentum_first():
    """Fermions.cpp:684-776: P half kick first, then (link update, kick)
    steps, a last link update and a closing half kick:
      P_1/2 = P_0 - (e/2) F(U_0), U_1 = exp(i e P_1/2) U_0,
      P_3/2 = P_1/2 - e F(U_1),   U_2 = exp(i e P_3/2) U_1,
      P_2   = P_3/2 - (e/2) F(U_2),
    rebuilt from the oracle's force with numpy's eigendecomposition for the
    exponential.  A position-first leapfrog (half link update, full kick,
    half link update) is also reversible with an e^2 error, and fails this."""
    dims = DIMS4
    m = 0.2
    U0 = F.random_links(dims, 0.3, seed=70)
    P0 = F.random_momenta(dims, 71)
    phi = oc.apply_d(dims, m, U0, F.gaussian_spinor(dims, 72))
    beta, c11, c12, eps = 4.0, 5 / 3, -1 / 12, 0.07

    def force(U, which):
        if which & 2:
            X, *_ = oc.cg(dims, m, U, phi, tol=1e-14, max_iter=10000)
            Y = oc.apply_d(dims, m, U, X, dagger=True)
            return oc.force(dims, U, X, Y, beta, c11, c12, which=which)
        return oc.force(dims, U, None, None, beta, c11, c12, which=which)

    def step_links(U, P):
        E = F.expi_hermitian(eps * F.flat_to_mat(P).reshape(-1, 3, 3))
        return F.mat_to_flat(np.einsum("lij,ljk->lik", E, F.flat_to_mat(U).reshape(-1, 3, 3)).reshape(
            F.flat_to_mat(U).shape))

    for which in (1, 3):
        P = P0 - 0.5 * eps * force(U0, which)
        U1 = step_links(U0, P)
        P = P - eps * force(U1, which)
        U2 = step_links(U1, P)
        P2 = P - 0.5 * eps * force(U2, which)
        Ur, Pr, _ = oc.leapfrog(dims, m, U0, P0, phi, beta=beta, c11=c11, c12=c12, eps=eps, nsteps=2,
                                which=which, tol=1e-14, check=False)
        assert np.abs(Ur - U2).max() < 1e-12, which
        assert np.abs(Pr - P2).max() < 1e-9, which
        # the position-first order lands elsewhere
        Uh = step_links(U0, 0.5 * P0)  # exp(i (e/2) P0) U0
        Pp = P0 - eps * force(Uh, which)
        assert np.abs(step_links(Uh, 0.5 * Pp) - Ur).max() > 1e-4


# ---------------------------------------------------------------- Polyakov correlators
def _corr_pairs(spatial, P, r2):
    """Brute force over every pair of grid points (m, n): the displacement
    n - m wrapped into [-L/2, L/2) per axis (so -L/2 and not +L/2 stands for
    the half-way neighbour), the pairs with |d|^2 = r^2 averaged."""
    L = np.asarray(spatial)
    pts = np.array(list(np.ndindex(*spatial)))
    d = (pts[None, :, :] - pts[:, None, :]) % L
    d = np.where(d >= L // 2, d - L, d)
    sel = (d ** 2).sum(-1) == r2
    if not sel.any():
        return 0j
    Pm = np.broadcast_to(P[:, None], sel.shape)[sel]
    Pn = np.broadcast_to(P[None, :], sel.shape)[sel]
    return (Pm * Pn.conj()).sum() / sel.sum()


@pytest.mark.parametrize("spatial", [(4, 4, 4), (6, 6, 6), (8, 8, 8), (6, 4, 8)])
def test_polyakov_correlator_against_pair_sum(spatial):
    """correlator_over_fixed_distance_fast (Lattice.cpp:888-948) against an
    independent sum over all pairs of grid points: pins the displacement set
    (the +L/2 exclusion, the floor(sqrt r^2) bound), the conjugation order and
    the normalisation.  At L = 6 and 8 some r^2 mix displacements with and
    without an L/2 component, so counting +L/2 as well changes the average."""
    rng = np.random.default_rng(sum(spatial))
    n = int(np.prod(spatial))
    P = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    L = np.asarray(spatial)
    for r2 in range(1, int(((L // 2) ** 2).sum()) + 2):
        ref = _corr_pairs(spatial, P, r2)
        got = oc.polyakov_correlator(spatial, P, r2)
        assert abs(got - ref) < 1e-12 * max(1.0, abs(ref)), (spatial, r2, got, ref)


def test_polyakov_correlator_cold_and_empty():
    """Cold links: P = 3 everywhere, so C = 9 wherever a displacement in
    [-L/2, L/2)^3 has |d|^2 = r^2, else 0 (r^2 = 7, 15, ... are not sums of
    three squares; some larger r^2 need a component beyond L/2)."""
    dims = (8, 8, 8, 4)
    P = oc.polyakov_grid(dims, F.cold_links(dims))
    reach = {dx * dx + dy * dy + dz * dz for dx in range(-4, 4) for dy in range(-4, 4) for dz in range(-4, 4)}
    for r2 in range(1, 52):
        c = oc.polyakov_correlator(dims[:3], P, r2)
        assert c == (9.0 if r2 in reach else 0.0), (r2, c)
    assert 7 not in reach and 37 not in reach and 48 in reach
