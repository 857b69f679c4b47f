"""Pins of the HEX smearing oracle (Smearing.cpp:76-331).

The reference ships no smearing test, and its smearing cannot run on a
non-trivial configuration (its stout step hands an anti-hermitian matrix to
the exp(iQ) routine, whose guards reject it), so the oracle is pinned by what
the construction fixes:
  * cold links are a fixed point;
  * gauge covariance: smearing gauge-transformed links gives the transformed
    smeared links (catches a wrong path orientation or a misplaced dagger);
  * hypercubic locality: the smeared link (x, mu) depends on exactly the thin
    links inside the hypercubes that hold it; re-drawing one outside leaves it
    bit-identical, re-drawing one inside changes it (catches a wrong excluded
    direction at levels 1 and 2);
  * alpha2 = alpha3 = 0 leaves levels 1 and 2 at the thin links, and the step
    is then one stout step exp(proj(alpha1/2 C U^dagger)) U with C the six
    plaquette staples, computed here from the oracle's own staple sums and
    scipy's expm (catches the exp convention, the projection and alpha/2);
  * outputs stay in SU(3); at small alpha every step raises the plaquette
    (smoothing).  At the reference's full alpha with its alpha/2 per level the
    outer step overshoots on smooth links (6 staples x 0.475), so smoothing
    is checked at alpha / 10 and, with the HEX paper's per-level
    normalisation (alpha1/6, alpha2/4, alpha3/2, i.e. alpha = (alpha1/3,
    alpha2/2, alpha3) here), at full strength.
"""
import itertools

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import oracle as O

from fields import random_links

DIMS = (4, 4, 4, 4)


def _mats(U):
    return (U[0::2] + 1j * U[1::2]).reshape(-1, 4, 3, 3)


def _flat(M):
    return np.stack([M.real, M.imag], -1).reshape(-1).copy()


def _rand_su3(rng, n):
    z = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / abs(d))[:, None, :]
    return q / np.linalg.det(q)[:, None, None] ** (1 / 3)


def _site(n, dims=DIMS):
    return int(np.ravel_multi_index(tuple(np.mod(n, dims)), dims))


def test_cold_links_are_a_fixed_point():
    V = int(np.prod(DIMS))
    U = np.zeros(V * 72)
    U[0::18] = U[8::18] = U[16::18] = 1.0
    assert np.array_equal(O.hex_smear(DIMS, U, nsteps=2), U)


def test_gauge_covariance():
    rng = np.random.default_rng(1)
    V = int(np.prod(DIMS))
    U = random_links(DIMS, 0.8, 2)
    G = _rand_su3(rng, V)
    M = _mats(U)
    Mg = np.empty_like(M)
    for s in range(V):
        n = np.array(np.unravel_index(s, DIMS))
        for mu in range(4):
            e = np.zeros(4, int)
            e[mu] = 1
            Mg[s, mu] = G[s] @ M[s, mu] @ G[_site(n + e)].conj().T
    S = _mats(O.hex_smear(DIMS, U))
    Sg = _mats(O.hex_smear(DIMS, _flat(Mg)))
    for s in range(V):
        n = np.array(np.unravel_index(s, DIMS))
        for mu in range(4):
            e = np.zeros(4, int)
            e[mu] = 1
            assert np.abs(Sg[s, mu] - G[s] @ S[s, mu] @ G[_site(n + e)].conj().T).max() < 1e-12


def test_hypercubic_locality():
    dims = (6, 6, 6, 6)
    U = random_links(dims, 0.8, 3)
    S = _mats(O.hex_smear(dims, U))
    rng = np.random.default_rng(4)
    x = np.array([2, 3, 2, 3])
    for mu in range(4):
        target = (_site(x, dims), mu)
        # the sites of the hypercubes holding link (x, mu)
        cube = set()
        for off in itertools.product(*[(0, 1) if d == mu else (-1, 0, 1) for d in range(4)]):
            cube.add(_site(x + np.array(off), dims))
        inside, outside = [], []
        for s in range(int(np.prod(dims))):
            n = np.array(np.unravel_index(s, dims))
            for lam in range(4):
                e = np.zeros(4, int)
                e[lam] = 1
                (inside if s in cube and _site(n + e, dims) in cube else outside).append((s, lam))
        for s, lam in [outside[k] for k in rng.choice(len(outside), 12, replace=False)]:
            W = U.copy()
            W[(s * 4 + lam) * 18:(s * 4 + lam + 1) * 18] = _flat(_rand_su3(rng, 1)[0])
            assert np.array_equal(_mats(O.hex_smear(dims, W))[target], S[target])
        for s, lam in [inside[k] for k in rng.choice(len(inside), 12, replace=False)]:
            W = U.copy()
            W[(s * 4 + lam) * 18:(s * 4 + lam + 1) * 18] = _flat(_rand_su3(rng, 1)[0])
            assert np.abs(_mats(O.hex_smear(dims, W))[target] - S[target]).max() > 1e-8


def test_outer_level_alone_is_one_stout_step():
    U = random_links(DIMS, 0.8, 5)
    a1 = 0.95
    S = _mats(O.hex_smear(DIMS, U, alpha=(a1, 0.0, 0.0)))
    M = _mats(U)
    for s in range(0, int(np.prod(DIMS)), 7):
        for mu in range(4):
            A1, _ = O.staples(DIMS, U, s, mu)
            C = ((A1[0::2] + 1j * A1[1::2]).reshape(3, 3)).conj().T  # paths x -> x + mu
            W = a1 / 2 * C @ M[s, mu].conj().T
            P = 0.5 * (W - W.conj().T)
            P -= np.trace(P) / 3 * np.eye(3)
            assert np.abs(S[s, mu] - expm(P) @ M[s, mu]).max() < 1e-13


def test_outputs_stay_in_su3_and_smooth():
    U = random_links(DIMS, 1.0, 6)
    S = O.hex_smear(DIMS, U, nsteps=4)
    M = _mats(S).reshape(-1, 3, 3)
    assert np.abs(np.matmul(np.conj(np.swapaxes(M, 1, 2)), M) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-13
    a1, a2, a3 = O.HEX_ALPHA
    for eps in (0.3, 1.0):
        W = random_links(DIMS, eps, 6)
        p = [O.avg_plaq(DIMS, O.hex_smear(DIMS, W, alpha=(a1 / 10, a2 / 10, a3 / 10), nsteps=k)) for k in range(4)]
        assert p[0] < p[1] < p[2] < p[3] <= 6.0, (eps, p)
        p = [O.avg_plaq(DIMS, O.hex_smear(DIMS, W, alpha=(a1 / 3, a2 / 2, a3), nsteps=k)) for k in range(3)]
        assert p[0] + 0.5 < p[1] <= 6.0 and p[0] + 0.5 < p[2] <= 6.0, (eps, p)


def test_rejects_bad_input():
    with pytest.raises(O.OracleError):
        O.hex_smear(DIMS, random_links(DIMS, 0.5, 1), nsteps=-1)
