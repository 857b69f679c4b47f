"""Parity of the device HEX smearing (sfh_hex_smear, through the C ABI) with
the oracle, plus size-independent properties at 16^4 and 32^4.

Tolerance: each level is a smooth map of its inputs, so device and oracle
agree to rounding: links within 1e-12 absolute after up to four steps."""
import numpy as np
import pytest

from oracle import oracle as oc
from staggered_fermions_hmc_amd import device as dev

pytestmark = pytest.mark.gpu

ALPHA = (0.95, 0.76, 0.38)          # Smearing.h:7-9
PAPER = (0.95 / 3, 0.76 / 2, 0.38)  # the same with the HEX paper's per-level staple normalisation


@pytest.fixture(scope="module")
def ctx():
    c = dev.Context(0)
    yield c
    c.close()


def _mats(U, dims):
    return (U[0::2] + 1j * U[1::2]).reshape(tuple(dims) + (4, 3, 3))


def _flat(M):
    return np.stack([M.real, M.imag], -1).reshape(-1).copy()


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (6, 4, 4, 8), (8, 8, 8, 8)])
@pytest.mark.parametrize("alpha", [ALPHA, PAPER], ids=["reference", "paper"])
@pytest.mark.parametrize("nsteps", [1, 4])
def test_hex_matches_oracle(ctx, dims, alpha, nsteps):
    U = dev.random_links(dims, 0.6, seed=41)
    g = dev.Gauge(ctx, dims, U)
    g.hex_smear(alpha, nsteps)
    ref = oc.hex_smear(dims, U, alpha=alpha, nsteps=nsteps)
    assert np.abs(g.download() - ref).max() < 1e-12


def test_hex_non_su3_links_match_oracle(ctx):
    """SU(3) links make every level SU(3), and the kernels then store and read
    two rows per link; links scaled by 0.9 are not SU(3), take the nine-element
    path, and still match the oracle (which makes no such assumption)."""
    dims = (6, 4, 4, 8)
    U = 0.9 * dev.random_links(dims, 0.6, seed=43)
    g = dev.Gauge(ctx, dims, U)
    g.hex_smear(ALPHA, 2)
    ref = oc.hex_smear(dims, U, alpha=ALPHA, nsteps=2)
    assert np.abs(g.download() - ref).max() < 1e-12


def test_hex_cold_links_fixed_and_zero_steps_noop(ctx):
    dims = (16, 16, 16, 16)
    U = np.zeros(int(np.prod(dims)) * 72)
    U[0::18] = U[8::18] = U[16::18] = 1.0
    g = dev.Gauge(ctx, dims, U)
    g.hex_smear(ALPHA, 2)
    assert np.array_equal(g.download(), U)
    W = dev.random_links(dims, 0.5, seed=2)
    g.upload(W)
    g.hex_smear(ALPHA, 0)
    assert np.array_equal(g.download(), W)


def test_hex_gauge_covariance_8x8x8x8(ctx):
    dims = (8, 8, 8, 8)
    U = dev.random_links(dims, 0.7, seed=6)
    G = _mats(dev.random_links(dims, 2.0, seed=7), dims)[..., 0, :, :]  # one SU(3) matrix per site
    M = _mats(U, dims)
    Mg = np.empty_like(M)
    for mu in range(4):
        Gs = np.roll(G, -1, axis=mu)  # G(x + mu)
        Mg[..., mu, :, :] = G @ M[..., mu, :, :] @ np.conj(np.swapaxes(Gs, -1, -2))
    g = dev.Gauge(ctx, dims, U)
    g.hex_smear(ALPHA, 1)
    S = _mats(g.download(), dims)
    g.upload(_flat(Mg))
    g.hex_smear(ALPHA, 1)
    Sg = _mats(g.download(), dims)
    for mu in range(4):
        Gs = np.roll(G, -1, axis=mu)
        want = G @ S[..., mu, :, :] @ np.conj(np.swapaxes(Gs, -1, -2))
        assert np.abs(Sg[..., mu, :, :] - want).max() < 1e-12


def test_hex_full_size_stays_in_su3_and_smooths(ctx):
    dims = (32, 32, 32, 32)
    U = dev.random_links(dims, 0.6, seed=8)
    g = dev.Gauge(ctx, dims, U)
    p0 = g.avg_plaq()
    g.hex_smear(PAPER, 1)
    p1 = g.avg_plaq()
    M = _mats(g.download(), dims).reshape(-1, 3, 3)
    assert np.abs(np.matmul(np.conj(np.swapaxes(M, 1, 2)), M) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-13
    assert p1 > p0 + 0.5


def test_hex_rejects_bad_input(ctx):
    dims = (4, 4, 4, 4)
    U = dev.random_links(dims, 0.5, seed=1)
    g = dev.Gauge(ctx, dims, U)
    with pytest.raises(dev.SfhError):
        g.hex_smear(ALPHA, -1)
    with pytest.raises(dev.SfhError):
        g.hex_smear((np.nan, 0.76, 0.38), 1)
    U[7] = np.inf
    g.upload(U)
    with pytest.raises(dev.SfhError):
        g.hex_smear(ALPHA, 1)
