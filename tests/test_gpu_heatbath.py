"""Parity of the device heatbath (sfh_heatbath, through the C ABI) with the
oracle's colour-ordered sweep on the same links, seed and sweep indices, plus
size-independent properties at 16^4 and 32^4.

Tolerance: the device and the oracle draw the same Philox uniforms and make
the same accept/reject decisions, so links agree to rounding as carried by
the chain.  How far the chain carries a rounding difference depends on the
coupling (tests/test_oracle_heatbath.py pins it): at beta = 4.1 a 1e-16
relative perturbation stays below 1e-13 over two sweeps, so those cases take
links within 1e-12 absolute and per-sweep averages within 1e-12 relative; at
beta = 1 it grows about 1e4-fold per sweep, so the strong-coupling case runs
one sweep and takes links within 1e-9."""
import numpy as np
import pytest

from oracle import oracle as oc
from staggered_fermions_hmc_amd import device as dev
import fields as F

pytestmark = pytest.mark.gpu

SYMANZIK = (4.1, 5 / 3, -1 / 12)   # beta_min and c_11, c_12 of Parameters.h:29, :51-52
WILSON = (4.1, 1.0, 0.0)
STRONG = (1.0, 1.0, 0.0)


@pytest.fixture(scope="module")
def ctx():
    c = dev.Context(0)
    yield c
    c.close()


def _cold(dims):
    U = np.zeros(int(np.prod(dims)) * 72)
    U[0::18] = U[8::18] = U[16::18] = 1.0
    return U


def _unitarity(U):
    M = (U[0::2] + 1j * U[1::2]).reshape(-1, 3, 3)
    dev_u = np.abs(np.matmul(np.conj(np.swapaxes(M, 1, 2)), M) - np.eye(3)).max()
    dev_d = np.abs(np.linalg.det(M) - 1).max()
    return dev_u, dev_d


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 8, 4, 8), (8, 8, 8, 8)])
@pytest.mark.parametrize("action", [SYMANZIK, WILSON], ids=["symanzik", "wilson"])
def test_heatbath_matches_oracle(ctx, dims, action):
    beta, c11, c12 = action
    U = dev.random_links(dims, 0.6, seed=31)
    g = dev.Gauge(ctx, dims, U)
    pl, ac = g.heatbath(beta, c11, c12, seed=0x1234abcd5678, sweep0=5, nsweeps=2)
    Uref, plr, acr = oc.heatbath(dims, U, beta, c11, c12, seed=0x1234abcd5678, sweep0=5, nsweeps=2, order=1)
    out = g.download()
    assert np.abs(out - Uref).max() < 1e-12
    assert np.abs(out - U).max() > 0.1
    assert np.allclose(pl, plr, rtol=1e-12, atol=0) and np.allclose(ac, acr, rtol=1e-12, atol=0)


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (8, 8, 8, 8)])
def test_heatbath_strong_coupling_sweep_matches_oracle(ctx, dims):
    U = dev.random_links(dims, 0.6, seed=31)
    g = dev.Gauge(ctx, dims, U)
    pl, ac = g.heatbath(*STRONG, seed=0x1234abcd5678, sweep0=5, nsweeps=1)
    Uref, plr, acr = oc.heatbath(dims, U, *STRONG, seed=0x1234abcd5678, sweep0=5, nsweeps=1, order=1)
    assert np.abs(g.download() - Uref).max() < 1e-9
    assert np.allclose(pl, plr, rtol=1e-10, atol=0) and np.allclose(ac, acr, rtol=1e-10, atol=0)


def test_heatbath_non_su3_links_match_oracle(ctx):
    """SU(3) links are kept as two rows in the heatbath's copy; links scaled
    by 0.9 are not SU(3) and keep all three rows, and still match the
    oracle (which makes no assumption about the links)."""
    dims = (4, 4, 4, 8)
    U = 0.9 * dev.random_links(dims, 0.6, seed=33)
    g = dev.Gauge(ctx, dims, U)
    pl, ac = g.heatbath(*SYMANZIK, seed=5, sweep0=1, nsweeps=1)
    Uref, plr, acr = oc.heatbath(dims, U, *SYMANZIK, seed=5, sweep0=1, nsweeps=1, order=1)
    assert np.abs(g.download() - Uref).max() < 1e-12
    assert np.allclose(pl, plr, rtol=1e-12, atol=0) and np.allclose(ac, acr, rtol=1e-12, atol=0)


def test_heatbath_cold_start_matches_oracle(ctx):
    dims = (4, 4, 4, 8)
    U = _cold(dims)
    g = dev.Gauge(ctx, dims, U)
    pl, ac = g.heatbath(*SYMANZIK, seed=7, sweep0=0, nsweeps=3)
    Uref, plr, acr = oc.heatbath(dims, U, *SYMANZIK, seed=7, sweep0=0, nsweeps=3, order=1)
    assert np.abs(g.download() - Uref).max() < 1e-12
    assert np.allclose(pl, plr, rtol=1e-12, atol=0) and np.allclose(ac, acr, rtol=1e-12, atol=0)


def test_heatbath_is_deterministic_and_sweeps_compose(ctx):
    dims = (8, 8, 8, 8)
    U = dev.random_links(dims, 0.5, seed=2)
    g1 = dev.Gauge(ctx, dims, U)
    p1, a1 = g1.heatbath(*SYMANZIK, seed=99, sweep0=3, nsweeps=2)
    g2 = dev.Gauge(ctx, dims, U)
    p2a, a2a = g2.heatbath(*SYMANZIK, seed=99, sweep0=3, nsweeps=1)
    p2b, a2b = g2.heatbath(*SYMANZIK, seed=99, sweep0=4, nsweeps=1)
    assert np.array_equal(g1.download(), g2.download())
    assert np.array_equal(p1, np.concatenate([p2a, p2b])) and np.array_equal(a1, np.concatenate([a2a, a2b]))
    g3 = dev.Gauge(ctx, dims, U)
    g3.heatbath(*SYMANZIK, seed=100, sweep0=3, nsweeps=1)
    assert np.abs(g3.download() - g2.download()).max() > 0.1


def test_heatbath_then_solver_sees_new_links(ctx):
    """The reduced-precision link copies are rebuilt after a heatbath."""
    dims = (8, 8, 8, 8)
    U = dev.random_links(dims, 0.5, seed=12)
    g = dev.Gauge(ctx, dims, U)
    psi = dev.random_spinor(dims, seed=3)
    g.cg(psi, 0.1, tol=1e-10, precision="f32")
    g.heatbath(*SYMANZIK, seed=5, sweep0=0, nsweeps=1)
    Unew = g.download()
    x1, _ = g.cg(psi, 0.1, tol=1e-10, precision="f32")
    xr, _, _, _ = oc.cg(dims, 0.1, Unew, psi, tol=1e-12)
    assert np.abs(x1 - xr).max() / np.abs(xr).max() < 1e-8


def test_heatbath_rejects_bad_input(ctx):
    g = dev.Gauge(ctx, (4, 4, 4, 6), dev.random_links((4, 4, 4, 6), 0.5, seed=1))
    with pytest.raises(dev.SfhError):
        g.heatbath(*SYMANZIK, seed=1)
    g = dev.Gauge(ctx, (4, 4, 4, 4), dev.random_links((4, 4, 4, 4), 0.5, seed=1))
    with pytest.raises(dev.SfhError):
        g.heatbath(0.0, 1.0, 0.0, seed=1)
    with pytest.raises(dev.SfhError):
        g.heatbath(float("nan"), 1.0, 0.0, seed=1)


def test_heatbath_reports_sampler_failure_on_non_finite_links(ctx):
    """A NaN link makes its own and its neighbours' rejection loops exhaust
    their 2^20 tries; the call fails with SFH_E_SAMPLER and those links keep
    their values (the reference's loops would never end)."""
    dims = (4, 4, 4, 4)
    U = dev.random_links(dims, 0.5, seed=3)
    U[5 * 72 + 2 * 18] = np.nan
    g = dev.Gauge(ctx, dims, U)
    with pytest.raises(dev.SfhError) as e:
        g.heatbath(*SYMANZIK, seed=1)
    assert e.value.status == 12
    out = g.download()
    assert np.isnan(out[5 * 72 + 2 * 18])
    assert np.isnan(out).sum() == 1


def test_heatbath_strong_coupling_plaquette(ctx):
    """Wilson action at beta = 1 on 8^4: the measured plaquette and the
    plaquette the sweep reports both average to the exact single-plaquette
    SU(3) value u(1) (lattice corrections O(u^5) ~ 1e-6)."""
    dims = (8, 8, 8, 8)
    g = dev.Gauge(ctx, dims, _cold(dims))
    g.heatbath(*STRONG, seed=17, sweep0=0, nsweeps=20)
    meas, rep = [], []
    for k in range(100):
        pl, _ = g.heatbath(*STRONG, seed=17, sweep0=20 + k, nsweeps=1)
        rep.append(pl[0])
        meas.append(g.avg_plaq() / 6.0)
    u = F.u_exact(1.0)
    for x in (np.array(meas), np.array(rep)):
        assert abs(x.mean() - u) < 4 * x.std() / np.sqrt(len(x)) + 1e-4, (x.mean(), u)
    # 120 sweeps of links stored as two rows stay unitary to rounding
    assert max(_unitarity(g.download())) < 1e-12


@pytest.mark.parametrize("dims", [(16, 16, 16, 16), (32, 32, 32, 32)])
def test_heatbath_full_size_stays_in_su3(ctx, dims):
    U = dev.random_links(dims, 0.6, seed=8)
    g = dev.Gauge(ctx, dims, U)
    pl, ac = g.heatbath(*SYMANZIK, seed=3, sweep0=0, nsweeps=1)
    out = g.download()
    du, dd = _unitarity(out)
    assert du < 1e-13 and dd < 1e-13
    # every link was updated once
    assert np.abs(out - U).reshape(-1, 18).max(axis=1).min() > 0
    assert 0.0 < pl[0] < 1.0 and ac[0] < 0.0
