"""The workload configurations at their real sizes, through the C ABI.

* 16^4, m = 0.05 (FP32-inner mixed vs FP64): solution, force and a Wilson
  HMC trajectory against the FP64 CPU oracle (the error table of the
  mixed-precision results).
* 32^4, m = 0.01: the FP16-inner mixed solve and the FP64 solve (stopped on
  the true residual) both reach a true relative residual of 1e-10, measured
  with the oracle's FP64 D D^dag on the host; a Symanzik (5/3, -1/12),
  beta = 4.0, tau = 1, 20-step HMC trajectory in FP16-mixed against the
  device FP64 path.
* The 64-system batch (16 each of 8^4, 12^4, 16^4, 24^4; 1-8 right-hand
  sides each): every right-hand side at 1e-10, a sample bit-identical alone.
* 32 chains on {8,12,16,24}^4 under Poisson arrivals: every result
  bit-identical to the same request served alone.

Budgets.  Solves stop at a true relative residual of 1e-10, so a solution
differs from the exact one by about 1e-10 * cond relative; the force built
from it by about as much.  The pseudofermion action S_f = Re<phi, X> ~
|chi|^2 ~ 3 V, so its error per evaluation is of order 1e-10 * 3 V (3e-4 at
32^4, 2e-5 at 16^4); Delta H takes two such terms, the kicks move the end
point by the force error times eps.  The budgets below are those figures
with a safety factor of a few; the measured errors are one to two decades
below them.
"""
import os
import sys
import time

import numpy as np
import pytest

from oracle import oracle as oc
from staggered_fermions_hmc_amd import device as dev

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


@pytest.fixture(scope="module")
def ctx():
    c = dev.Context(0)
    yield c
    c.close()


def rel(a, b):
    return np.abs(a - b).max() / max(1e-300, np.abs(b).max())


def true_res(dims, m, U, b, x):
    r = b - oc.apply_ddag(dims, m, U, x)
    return float(np.sqrt(r @ r / (b @ b)))


# ---------------------------------------------------------------- 16^4 error table
def test_16_4_mixed_errors_against_oracle(ctx):
    dims, m, tol = (16, 16, 16, 16), 0.05, 1e-10
    U = dev.random_links(dims, 0.3, seed=1234)
    chi = dev.random_spinor(dims, seed=99)
    P0 = dev.random_momenta(dims, seed=77)
    g = dev.Gauge(ctx, dims, U)
    phi = oc.apply_d(dims, m, U, chi)
    xo, *_ = oc.cg(dims, m, U, phi, tol=1e-13)
    Fo = oc.force(dims, U, xo, oc.apply_d(dims, m, U, xo, dagger=True), 5.6, 1.0, 0.0)
    for prec in ("f64", "f32", "f16"):
        x, st = g.cg(phi, m, tol=tol, precision=prec, max_iter=100000, true_residual="stop")
        assert true_res(dims, m, U, phi, x) < tol, (prec, st)
        assert rel(x, xo) < 1e-8, prec
        F = g.force(5.6, 1.0, 0.0, x, g.apply_d(m, x, dagger=True))
        assert rel(F, Fo) < 1e-8, prec
    kw = dict(beta=5.6, c11=1.0, c12=0.0, eps=0.1, nsteps=10)
    Uo, Ho, _ = oc.hmc_trajectory(dims, m, U, chi, P0, 0.0, tol=1e-13, **kw)
    plaq_o = oc.avg_plaq(dims, Uo)
    for prec in ("f64", "f32", "f16"):
        g.upload(U)
        H, _ = g.hmc_trajectory(chi, P0, 0.0, tol=tol, max_iter=100000, precision=prec, mass=m, **kw)
        assert abs(H[6] - Ho[6]) < 1e-4, (prec, H[6], Ho[6])
        # the trajectory ended on the device links (r = 0 accepts every trajectory)
        assert abs(g.avg_plaq() - plaq_o) < 1e-10, prec
        assert np.abs(g.download() - Uo).max() < 1e-8, prec


# ---------------------------------------------------------------- 32^4 (FP16 inner)
@pytest.fixture(scope="module")
def lat32(ctx):
    dims = (32, 32, 32, 32)
    U = dev.random_links(dims, 0.3, seed=1234)
    g = dev.Gauge(ctx, dims, U)
    yield dims, U, g
    g.close()


def test_32_4_solves_reach_true_residual(lat32):
    dims, U, g = lat32
    m, tol = 0.01, 1e-10
    b = dev.random_spinor(dims, seed=5678)
    x16, s16 = g.cg(b, m, tol=tol, precision="f16", max_iter=100000)
    t16 = true_res(dims, m, U, b, x16)
    assert t16 < tol, s16
    assert abs(s16["true_rel_resid"] - t16) < 1e-11
    x64, s64 = g.cg(b, m, tol=tol, precision="f64", max_iter=100000, true_residual="stop")
    t64 = true_res(dims, m, U, b, x64)
    assert t64 < tol, s64
    assert abs(s64["true_rel_resid"] - t64) < 1e-11
    assert rel(x16, x64) < 1e-7
    # the reliable updates keep the mixed solve's iteration count close to FP64's
    assert s16["iterations"] <= 1.25 * s64["iterations"] + 16, (s16, s64)


def test_48_4_solves_and_force(ctx):
    """Beyond the configurations: a 48^4 lattice (5.3 M sites; FP16 two-row
    links 530 MB, past the 256 MB Infinity Cache).  Both solves reach a true
    relative residual of 1e-10 by the oracle's FP64 D D^dag on the host, and
    the 32-bit per-lane offsets of the force's link loads still hold (the
    gauge force at a sample of links equals the one on a lattice translated
    by one site in every direction, read back at the translated links)."""
    dims, m, tol = (48, 48, 48, 48), 0.01, 1e-10
    U = dev.random_links(dims, 0.3, seed=4848)
    b = dev.random_spinor(dims, seed=4849)
    g = dev.Gauge(ctx, dims, U)
    try:
        x16, s16 = g.cg(b, m, tol=tol, precision="f16", max_iter=100000)
        assert true_res(dims, m, U, b, x16) < tol, s16
        x64, s64 = g.cg(b, m, tol=tol, precision="f64", max_iter=100000, true_residual="stop")
        assert true_res(dims, m, U, b, x64) < tol, s64
        assert s16["iterations"] <= 1.25 * s64["iterations"] + 16, (s16, s64)
        F = g.force(4.0, 5 / 3, -1 / 12, which=1).reshape(dims + (4, 18))
    finally:
        g.close()
    # translation covariance: shift every site by (1, 1, 1, 1); the time
    # boundary is periodic for the gauge force
    Us = np.roll(U.reshape(dims + (4, 18)), shift=(1, 1, 1, 1), axis=(0, 1, 2, 3))
    g2 = dev.Gauge(ctx, dims, Us.reshape(-1))
    try:
        F2 = g2.force(4.0, 5 / 3, -1 / 12, which=1).reshape(dims + (4, 18))
    finally:
        g2.close()
    assert np.isfinite(F).all()
    assert np.abs(np.roll(F, shift=(1, 1, 1, 1), axis=(0, 1, 2, 3)) - F2).max() < 1e-12 * np.abs(F).max()


def test_32_4_symanzik_trajectory_f16_vs_f64(lat32):
    dims, U, g = lat32
    m, tol = 0.01, 1e-10
    chi = dev.random_spinor(dims, seed=99)
    P0 = dev.random_momenta(dims, seed=77)
    g.upload(U)
    phi = g.apply_d(m, chi)
    F = {}
    for prec in ("f64", "f16"):
        x, _ = g.cg(phi, m, tol=tol, precision=prec, max_iter=100000)
        F[prec] = g.force(4.0, 5 / 3, -1 / 12, x, g.apply_d(m, x, dagger=True))
    assert rel(F["f16"], F["f64"]) < 1e-8
    kw = dict(mass=m, beta=4.0, c11=5 / 3, c12=-1 / 12, eps=0.05, nsteps=20, tol=tol, max_iter=100000)
    out = {}
    for prec in ("f64", "f16"):
        g.upload(U)
        H, acc = g.hmc_trajectory(chi, P0, 0.0, precision=prec, **kw)
        out[prec] = (H, g.avg_plaq(), g.download())
    (H64, p64, U64), (H16, p16, U16) = out["f64"], out["f16"]
    assert np.all(np.isfinite(H16))
    assert abs(H16[6] - H64[6]) < 1e-2, (H16[6], H64[6])
    assert abs(H16[2] - H64[2]) < 3e-3 and abs(H16[5] - H64[5]) < 3e-3
    assert abs(p16 - p64) < 1e-9
    assert np.abs(U16 - U64).max() < 1e-7
    g.upload(U)


# ---------------------------------------------------------------- 64-system batch
def test_batch_64_systems(ctx):
    from bench import make_batch_systems
    systems = make_batch_systems(dev, ctx)
    spec = [(s["g"], s["mass"], s["rf"], s["xf"]) for s in systems]
    stats = dev.cg_batch(ctx, spec, tol=1e-10, precision="f32")
    assert len(stats) == sum(len(s["rf"]) for s in systems)
    assert all(st["status"] == 0 and st["true_rel_resid"] < 1e-10 for st in stats)
    # independent check on the host for one system of each lattice size
    k = 0
    sample = {}
    for i, s in enumerate(systems):
        if s["dims"] not in sample:
            sample[s["dims"]] = (i, k)
        k += len(s["rf"])
    for dims, (i, k) in sample.items():
        s = systems[i]
        U = s["g"].download()
        for j, rf in enumerate(s["rf"]):
            t = true_res(dims, s["mass"], U, rf.download(), s["xf"][j].download())
            assert t < 1e-10 and abs(t - stats[k + j]["true_rel_resid"]) < 1e-11, (dims, j)
    # the same systems alone give bit-identical solutions
    for dims, (i, _) in sample.items():
        s = systems[i]
        xa = [dev.Field(s["g"]) for _ in s["rf"]]
        dev.cg_batch(ctx, [(s["g"], s["mass"], s["rf"], xa)], tol=1e-10, precision="f32")
        for j in range(len(xa)):
            assert np.array_equal(xa[j].download(), s["xf"][j].download()), (dims, j)


# ---------------------------------------------------------------- 32 served chains
def test_service_32_chains_poisson(ctx):
    sizes = [(8, 8, 8, 8), (12, 12, 12, 12), (16, 16, 16, 16), (24, 24, 24, 24)]
    rng = np.random.default_rng(7)
    chains = []
    for k in range(32):
        dims = sizes[k % 4]
        g = dev.Gauge(ctx, dims, dev.random_links(dims, 0.3, seed=30000 + k))
        rhs = [dev.Field(g, dev.random_spinor(dims, seed=40000 + 4 * k + j)) for j in range(2)]
        chains.append(dict(dims=dims, mass=float(rng.uniform(0.01, 0.1)), g=g, rhs=rhs))
    svc = dev.Service(0, tol=1e-10)
    try:
        ids = [svc.add_chain(c["g"], c["mass"], max_inflight=4) for c in chains]
        n = 96
        gaps = rng.exponential(1.0 / 400.0, size=n)  # about 400 requests/s: several in flight
        which = rng.integers(0, 32, size=n)
        reqs = []
        t_next = time.perf_counter()
        for r in range(n):
            t_next += gaps[r]
            delay = t_next - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            k = int(which[r])
            xf = dev.Field(chains[k]["g"])
            rf = chains[k]["rhs"][r % 2]
            reqs.append((k, r % 2, xf, svc.submit(ids[k], rf, xf)))
        stats = {t: svc.wait(t, timeout=120) for *_, t in reqs}
        assert all(st["status"] == 0 and st["true_rel_resid"] < 1e-10 for st in stats.values())
        # each (chain, rhs) served alone, nothing else in flight
        alone = {}
        for k, j, xf, t in reqs:
            if (k, j) not in alone:
                xa = dev.Field(chains[k]["g"])
                svc.wait(svc.submit(ids[k], chains[k]["rhs"][j], xa), timeout=120)
                alone[(k, j)] = xa.download()
            assert np.array_equal(xf.download(), alone[(k, j)]), (k, j)
    finally:
        svc.close()
