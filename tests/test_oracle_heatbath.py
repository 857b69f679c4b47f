"""Pins of the heatbath oracle (Lattice.cpp:509-757, SU3_Sampling.cpp:17-199).

The reference ships no test of its heatbath, and its uniforms come from
std::random_device seeds (rng.cpp:8-27), so no run of it can be replayed.
The oracle is pinned instead by what the mathematics fixes:
  * Philox4x32-10 against the published Random123 known-answer vectors;
  * the x0 sampler against its exact law sqrt(1 - x^2) exp(k x): mean
    I_2(k) / I_1(k) and the Kolmogorov-Smirnov distance to the exact CDF;
  * one subgroup update against the identity ReTr((U' A)_block) = 2 a x0,
    with a from an independent quaternion projection (B + s2 conj(B) s2) / 2,
    unitarity, and the untouched row;
  * the beta' = 2 beta / 3 of the subgroup sampler through the mean of x0 over
    many updates of one link;
  * the colour classes of the parallel sweep by brute force: re-drawing every
    other link of a class leaves a link's staples bit-identical;
  * the whole chain at strong coupling: the plaquette (Wilson) and rectangle
    (pure rectangle action) averages against the exact single-plaquette SU(3)
    integral u(beta) by Weyl's formula (lattice corrections are O(u^5)).
"""
import numpy as np
import pytest
from scipy import integrate
from scipy.special import iv

from oracle import oracle as O

from fields import random_links, u_exact

SIGMA_Y = np.array([[0, -1j], [1j, 0]])
SUB = [(0, 1), (0, 2), (1, 2)]


def _m(v18):
    return (v18[0::2] + 1j * v18[1::2]).reshape(3, 3)


def _v(m):
    return np.stack([m.real, m.imag], -1).reshape(-1).copy()


def _rand_su3(rng):
    z = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    q, r = np.linalg.qr(z)
    q = q * (np.diag(r) / abs(np.diag(r)))
    return q / np.linalg.det(q) ** (1 / 3)


def _x0_cdf(k, n=40001):
    x = np.linspace(-1, 1, n)
    dens = np.sqrt(np.clip(1 - x * x, 0, None)) * np.exp(k * (x - 1))
    cdf = integrate.cumulative_trapezoid(dens, x, initial=0)
    return x, cdf / cdf[-1]


def _ks(samples, k):
    x, cdf = _x0_cdf(k)
    s = np.sort(samples)
    F = np.interp(s, x, cdf)
    n = len(s)
    return max(np.max(np.arange(1, n + 1) / n - F), np.max(F - np.arange(n) / n)) * np.sqrt(n)


def test_u_exact_matches_the_strong_coupling_series():
    # u = beta/18 + beta^2/216 + O(beta^3) for SU(3)
    for b in (0.05, 0.1):
        assert abs(u_exact(b) - (b / 18 + b * b / 216)) < b ** 3 / 100


def test_philox_known_answer_vectors():
    # Random123 kat_vectors, philox4x32_10
    kat = [((0, 0), (0, 0, 0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff, 0xffffffff), (0xffffffff,) * 4, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0xa4093822, 0x299f31d0), (0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for key, ctr, want in kat:
        assert [int(x) for x in O.philox(key, ctr)] == list(want)


@pytest.mark.parametrize("k", [0.5, 3.0, 20.0])
def test_x0_follows_its_density(k):
    n = 100000
    x = O.heatbath_x0(1.0, k, n, seed=11)
    assert np.all(np.abs(x) <= 1)
    mean = iv(2, k) / iv(1, k)
    assert abs(x.mean() - mean) < 5 * x.std() / np.sqrt(n)
    assert _ks(x, k) < 1.95  # 0.1 % level
    # a and beta enter only through their product
    assert np.array_equal(O.heatbath_x0(2.0, k / 2, 1000, seed=11), x[:1000])


def test_x0_rejects_nonpositive_coupling():
    with pytest.raises(O.OracleError):
        O.heatbath_x0(1.0, 0.0, 10)


def _projection_norm(B):
    P = 0.5 * (B + SIGMA_Y @ B.conj() @ SIGMA_Y)
    return np.sqrt(np.linalg.det(P).real), P


def test_subgroup_update_identities():
    rng = np.random.default_rng(3)
    for trial in range(20):
        U = _rand_su3(rng)
        A = sum(_rand_su3(rng) for _ in range(6)) * rng.uniform(0.2, 1.5)
        for sub, (i, j) in enumerate(SUB):
            out, a, x0 = O.heatbath_subgroup(_v(U), _v(A), 4.0, sub, seed=5, link=trial, sweep=3)
            Un = _m(out)
            B = (U @ A)[np.ix_([i, j], [i, j])]
            a_ref, _ = _projection_norm(B)
            assert abs(a - a_ref) < 1e-12 * max(1.0, a_ref)
            assert -1 <= x0 <= 1
            assert np.abs(Un.conj().T @ Un - np.eye(3)).max() < 1e-14
            assert abs(np.linalg.det(Un) - 1) < 1e-14
            k = 3 - i - j
            assert np.array_equal(Un[k], U[k])
            Bn = (Un @ A)[np.ix_([i, j], [i, j])]
            assert abs(np.trace(Bn).real - 2 * a * x0) < 1e-12 * max(1.0, a)
            # only rows i, j change: Un = embed(R) U with R in SU(2)
            R = (Un @ U.conj().T)[np.ix_([i, j], [i, j])]
            assert np.abs(R.conj().T @ R - np.eye(2)).max() < 1e-13
            assert abs(np.linalg.det(R) - 1) < 1e-13


def test_subgroup_update_degenerate_staple_draws_haar_element():
    rng = np.random.default_rng(4)
    U = _rand_su3(rng)
    out, a, x0 = O.heatbath_subgroup(_v(U), np.zeros(18), 4.0, 0, seed=5, link=1, sweep=0)
    Un = _m(out)
    assert a == 0.0 and x0 == 0.0
    assert np.abs(Un.conj().T @ Un - np.eye(3)).max() < 1e-14
    assert np.abs(Un - U).max() > 1e-3


def test_sampler_gives_up_on_non_finite_links():
    """The rejection loops end after 2^20 tries (the reference's would not);
    the link keeps its value and the sweep reports the failure."""
    U = np.eye(3, dtype=complex)
    A = np.full((3, 3), np.nan, dtype=complex)
    with pytest.raises(O.OracleError) as e:
        O.heatbath_subgroup(_v(U), _v(A), 4.0, 0, seed=1)
    assert e.value.status == 12
    dims = (4, 4, 4, 4)
    W = random_links(dims, 0.5, 3)
    W[5 * 72 + 2 * 18] = np.nan
    with pytest.raises(O.OracleError) as e:
        O.heatbath(dims, W, 4.1, 5 / 3, -1 / 12, seed=1)
    assert e.value.status == 12


def test_subgroup_sampler_uses_two_thirds_beta():
    rng = np.random.default_rng(8)
    U = _rand_su3(rng)
    A = sum(_rand_su3(rng) for _ in range(4)) * 1.5
    beta = 2.5
    xs = []
    a0 = None
    for link in range(6000):
        _, a, x0 = O.heatbath_subgroup(_v(U), _v(A), beta, 1, seed=21, link=link, sweep=0)
        a0 = a
        xs.append(x0)
    xs = np.array(xs)
    k = a0 * 2 * beta / 3
    assert abs(xs.mean() - iv(2, k) / iv(1, k)) < 5 * xs.std() / np.sqrt(len(xs))
    assert _ks(xs, k) < 1.95
    # the law at beta (not 2 beta / 3) is far away
    assert abs(xs.mean() - iv(2, a0 * beta) / iv(1, a0 * beta)) > 10 * xs.std() / np.sqrt(len(xs))


def _colour(n, mu):
    return (sum(n) + n[mu]) & 3


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 8, 4, 8)])
def test_colour_classes_do_not_share_staples(dims):
    """Re-drawing every other link of a link's colour class leaves its staples
    unchanged; re-drawing one coupled link of another class changes them."""
    V = int(np.prod(dims))
    U = random_links(dims, 0.7, 5)
    rng = np.random.default_rng(2)
    coords = np.array(np.unravel_index(np.arange(V), dims)).T
    for mu in range(4):
        for s in rng.choice(V, 6, replace=False):
            col = _colour(coords[s], mu)
            A1, A2 = O.staples(dims, U, int(s), mu)
            W = U.copy()
            for t in range(V):
                if t != s and _colour(coords[t], mu) == col:
                    W[(t * 4 + mu) * 18:(t * 4 + mu + 1) * 18] = _v(_rand_su3(rng))
            B1, B2 = O.staples(dims, W, int(s), mu)
            assert np.array_equal(A1, B1) and np.array_equal(A2, B2)
            # control: the mu-link at n + 2 nu is in a rectangle staple
            nu = (mu + 1) % 4
            n2 = coords[s].copy()
            n2[nu] = (n2[nu] + 2) % dims[nu]
            t = int(np.ravel_multi_index(n2, dims))
            assert _colour(n2, mu) != col
            W = U.copy()
            W[(t * 4 + mu) * 18:(t * 4 + mu + 1) * 18] = _v(_rand_su3(rng))
            C1, C2 = O.staples(dims, W, int(s), mu)
            assert np.abs(C2 - A2).max() > 1e-6


def test_colour_order_is_thread_independent_and_shares_the_first_link():
    dims = (4, 4, 4, 8)
    U = random_links(dims, 0.5, 9)
    try:
        O.set_threads(1)
        U1, p1, a1 = O.heatbath(dims, U, 4.1, 5 / 3, -1 / 12, seed=3, sweep0=2, nsweeps=1, order=1)
        O.set_threads(8)
        U8, p8, a8 = O.heatbath(dims, U, 4.1, 5 / 3, -1 / 12, seed=3, sweep0=2, nsweeps=1, order=1)
    finally:
        O.set_threads(8)
    assert np.array_equal(U1, U8) and np.array_equal(p1, p8) and np.array_equal(a1, a8)
    # link (0, mu=0) is the first update of both orders, from the same links
    # and stream; the orders then differ
    U0, _, _ = O.heatbath(dims, U, 4.1, 5 / 3, -1 / 12, seed=3, sweep0=2, nsweeps=1, order=0)
    assert np.array_equal(U0[:18], U1[:18])
    assert np.abs(U0 - U1).max() > 1e-3
    # every link moved, and stays in SU(3)
    M = (U1[0::2] + 1j * U1[1::2]).reshape(-1, 3, 3)
    assert np.abs(np.einsum("lji,ljk->lik", M.conj(), M) - np.eye(3)).max() < 1e-13
    assert np.min(np.abs(U1 - U).reshape(-1, 18).max(axis=1)) > 0


def test_sweeps_compose():
    dims = (4, 4, 4, 4)
    U = random_links(dims, 0.5, 4)
    Ua, pa, aa = O.heatbath(dims, U, 2.0, 1.0, 0.0, seed=8, sweep0=10, nsweeps=2)
    Ub, pb, ab = O.heatbath(dims, U, 2.0, 1.0, 0.0, seed=8, sweep0=10, nsweeps=1)
    Ub, pc, ac = O.heatbath(dims, Ub, 2.0, 1.0, 0.0, seed=8, sweep0=11, nsweeps=1)
    assert np.array_equal(Ua, Ub)
    assert np.array_equal(pa, np.concatenate([pb, pc])) and np.array_equal(aa, np.concatenate([ab, ac]))


def test_heatbath_rejects_bad_input():
    U = random_links((4, 4, 4, 6), 0.5, 1)
    with pytest.raises(O.OracleError):
        O.heatbath((4, 4, 4, 6), U, 2.0, 1.0, 0.0, seed=1)
    U = random_links((4, 4, 4, 4), 0.5, 1)
    with pytest.raises(O.OracleError):
        O.heatbath((4, 4, 4, 4), U, 0.0, 1.0, 0.0, seed=1)


def test_chain_sensitivity_to_rounding_depends_on_coupling():
    """How far the heatbath chain carries a rounding difference: the bound
    behind the device-parity tolerances.  A 1e-16 relative perturbation of
    the links stays below 1e-13 over two sweeps at beta = 4.1 but grows about
    1e4-fold per sweep at beta = 1 (small a, large 1/a in the subgroup
    rotations)."""
    dims = (4, 4, 4, 4)
    U0 = random_links(dims, 0.6, 31)
    U1 = U0 * (1 + 1e-16 * np.random.default_rng(0).standard_normal(U0.size))
    growth = {}
    for act in [(4.1, 5 / 3, -1 / 12), (4.1, 1.0, 0.0), (1.0, 1.0, 0.0)]:
        Ua, Ub, d = U0, U1, []
        for k in range(2):
            Ua, _, _ = O.heatbath(dims, Ua, *act, seed=0x1234abcd5678, sweep0=5 + k, nsweeps=1)
            Ub, _, _ = O.heatbath(dims, Ub, *act, seed=0x1234abcd5678, sweep0=5 + k, nsweeps=1)
            d.append(np.abs(Ua - Ub).max())
        growth[act] = d
    assert max(growth[(4.1, 5 / 3, -1 / 12)]) < 1e-13 and max(growth[(4.1, 1.0, 0.0)]) < 1e-13
    d = growth[(1.0, 1.0, 0.0)]
    assert d[0] < 1e-10 and d[1] > 1e2 * d[0]


def _chain(dims, beta, c11, c12, order, ntherm, nmeas, seed):
    V = int(np.prod(dims))
    U = np.zeros(V * 72)
    U[0::18] = U[8::18] = U[16::18] = 1.0
    U, _, _ = O.heatbath(dims, U, beta, c11, c12, seed=seed, sweep0=0, nsweeps=ntherm, order=order)
    P, R = [], []
    for k in range(nmeas):
        U, _, _ = O.heatbath(dims, U, beta, c11, c12, seed=seed, sweep0=ntherm + k, nsweeps=1, order=order)
        P.append(O.avg_plaq(dims, U) / 6.0)
        R.append(-O.action(dims, U, 1.0, 0.0, 1.0) / (12 * V))
    return np.array(P), np.array(R)


@pytest.mark.parametrize("beta,c11,c12,order,which,coupling", [
    (1.0, 1.0, 0.0, 1, "plaq", 1.0),        # Wilson: <P> = u(beta)
    (1.0, 0.0, 1.0, 0, "rect", 1.0),        # pure rectangle action: <R> = u(beta)
    (0.6, 5 / 3, -1 / 12, 1, "plaq", 1.0),  # Symanzik: <P> = u(c11 beta) + O(u^2 c12)
])
def test_strong_coupling_averages(beta, c11, c12, order, which, coupling):
    P, R = _chain((4, 4, 4, 4), beta, c11, c12, order, 20, 150, seed=17)
    x = P if which == "plaq" else R
    err = x.std() / np.sqrt(len(x))
    assert abs(x.mean() - u_exact(coupling)) < 4 * err + 2e-4, (x.mean(), u_exact(coupling), err)
