"""Seeded numpy field generators and independent brute-force checks.

Layouts are the reference's AoS layouts: links [V,4,3,3] complex flattened to
float64 V*4*18 (Lattice.cpp:135-151), spinors [V,3] complex flattened to V*6
(Fermions.cpp:99-116), site s = ((n0*L1+n1)*L2+n2)*L3+n3 (Fermions.cpp:42).

Nothing here calls the oracle: these are the independent references the
oracle is pinned against (closed loops enumerated directly, matrix
exponentials by eigendecomposition).
"""
import itertools

import numpy as np

GELL_MANN = np.zeros((8, 3, 3), dtype=np.complex128)
GELL_MANN[0][0, 1] = GELL_MANN[0][1, 0] = 1
GELL_MANN[1][0, 1] = -1j; GELL_MANN[1][1, 0] = 1j
GELL_MANN[2][0, 0] = 1; GELL_MANN[2][1, 1] = -1
GELL_MANN[3][0, 2] = GELL_MANN[3][2, 0] = 1
GELL_MANN[4][0, 2] = -1j; GELL_MANN[4][2, 0] = 1j
GELL_MANN[5][1, 2] = GELL_MANN[5][2, 1] = 1
GELL_MANN[6][1, 2] = -1j; GELL_MANN[6][2, 1] = 1j
GELL_MANN[7] = np.diag([1, 1, -2]) / np.sqrt(3)


def volume(dims):
    return int(np.prod(dims))


def coords(dims):
    """[V,4] integer coordinates in reference site order (last index fastest)."""
    return np.array(list(itertools.product(*[range(L) for L in dims])), dtype=np.int64)


def site_index(dims, n):
    n = np.asarray(n) % np.asarray(dims)
    return ((n[..., 0] * dims[1] + n[..., 1]) * dims[2] + n[..., 2]) * dims[3] + n[..., 3]


def random_algebra(rng, shape):
    """H = sum_a r_a lambda_a / 2 with r_a ~ N(0,1) (Fermions.cpp:342-360)."""
    r = rng.standard_normal(shape + (8,))
    return 0.5 * np.einsum("...a,aij->...ij", r, GELL_MANN)


def expi_hermitian(H):
    """exp(iH) for hermitian H by eigendecomposition."""
    w, v = np.linalg.eigh(H)
    return np.einsum("...ij,...j,...kj->...ik", v, np.exp(1j * w), v.conj())


def random_links(dims, eps, seed):
    """U = exp(i eps H) per link, H as random_algebra: float64 V*4*18."""
    rng = np.random.default_rng(seed)
    V = volume(dims)
    U = expi_hermitian(eps * random_algebra(rng, (V, 4)))
    return mat_to_flat(U)


def cold_links(dims):
    V = volume(dims)
    U = np.broadcast_to(np.eye(3, dtype=np.complex128), (V, 4, 3, 3)).copy()
    return mat_to_flat(U)


def random_momenta(dims, seed):
    rng = np.random.default_rng(seed)
    return mat_to_flat(random_algebra(rng, (volume(dims), 4)))


def gaussian_spinor(dims, seed, sigma=np.sqrt(0.5)):
    """Complex Gaussian, sigma per real part (Fermions.cpp:207-212)."""
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, sigma, size=volume(dims) * 6)


def mat_to_flat(U):
    U = np.ascontiguousarray(U)
    return U.view(np.float64).reshape(-1).copy()


def flat_to_mat(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.complex128).reshape(-1, 4, 3, 3)


def flat_to_vec(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.complex128).reshape(-1, 3)


def vec_to_flat(v):
    return np.ascontiguousarray(v, dtype=np.complex128).view(np.float64).reshape(-1).copy()


def eps_sign(dims):
    c = coords(dims)
    return np.where(c.sum(axis=1) % 2 == 0, 1.0, -1.0)


def loop_traces(dims, Umat, mu, nu, a, b):
    """ReTr of every a x b Wilson loop in the (mu,nu) plane (a steps along mu,
    b along nu), enumerated by walking the closed path link by link."""
    dims = np.asarray(dims)
    c = coords(dims)
    V = len(c)
    e = np.eye(4, dtype=np.int64)
    W = np.broadcast_to(np.eye(3, dtype=np.complex128), (V, 3, 3)).copy()
    pos = c.copy()
    path = [(mu, +1)] * a + [(nu, +1)] * b + [(mu, -1)] * a + [(nu, -1)] * b
    for d, s in path:
        if s > 0:
            W = W @ Umat[site_index(dims, pos), d]
            pos = pos + e[d]
        else:
            pos = pos - e[d]
            W = W @ Umat[site_index(dims, pos), d].conj().transpose(0, 2, 1)
    return np.einsum("vii->v", W).real


def brute_force_action(dims, U, beta, c11, c12):
    """S = -(beta/3)[c11 sum_plaq ReTr + c12 sum_rect ReTr], every plaquette and
    every 2x1 and 1x2 rectangle counted once."""
    Umat = flat_to_mat(U)
    sp = sr = 0.0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            sp += loop_traces(dims, Umat, mu, nu, 1, 1).sum()
            sr += loop_traces(dims, Umat, mu, nu, 2, 1).sum()
            sr += loop_traces(dims, Umat, mu, nu, 1, 2).sum()
    return -(beta / 3.0) * (c11 * sp + c12 * sr), sp, sr


def gauge_transform(dims, U, psi, seed):
    """U'_mu(n) = G(n) U_mu(n) G(n+mu)^dag, psi'(n) = G(n) psi(n)."""
    rng = np.random.default_rng(seed)
    dims = np.asarray(dims)
    V = volume(dims)
    G = expi_hermitian(random_algebra(rng, (V,)) * 2.0)
    c = coords(dims)
    Um = flat_to_mat(U)
    Up = np.empty_like(Um)
    for mu in range(4):
        nb = site_index(dims, c + np.eye(4, dtype=np.int64)[mu])
        Up[:, mu] = G @ Um[:, mu] @ G[nb].conj().transpose(0, 2, 1)
    out = [mat_to_flat(Up)]
    if psi is not None:
        out.append(vec_to_flat(np.einsum("vij,vj->vi", G, flat_to_vec(psi))))
    out.append(G)
    return out


def u_exact(beta, n=256):
    """<ReTr U / 3> under exp((beta/3) ReTr U) dU on SU(3), by Weyl's formula
    over the eigenphases (the integrand is smooth and periodic, so the
    rectangle rule converges exponentially)."""
    th = np.linspace(0, 2 * np.pi, n, endpoint=False)
    t1, t2 = np.meshgrid(th, th, indexing="ij")
    t3 = -t1 - t2
    e = [np.exp(1j * t) for t in (t1, t2, t3)]
    vdm = abs(e[0] - e[1]) ** 2 * abs(e[0] - e[2]) ** 2 * abs(e[1] - e[2]) ** 2
    p = (np.cos(t1) + np.cos(t2) + np.cos(t3)) / 3
    w = vdm * np.exp(beta * p)
    return float((w * p).sum() / w.sum())
