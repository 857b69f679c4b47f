"""Solve service (sfh_service_*): continuous batching across chains, FIFO
admission by free slot memory, and each result identical to the solve alone."""
import numpy as np
import pytest

from oracle import oracle as oc
from staggered_fermions_hmc_amd import device as dev
from staggered_fermions_hmc_amd._abi import SfhError

pytestmark = pytest.mark.gpu

CHAINS = [((4, 4, 4, 4), 0.1), ((8, 8, 8, 8), 0.05), ((6, 4, 4, 2), 0.08), ((8, 6, 4, 4), 0.03)]


@pytest.fixture(scope="module")
def ctx():
    c = dev.Context(0)
    yield c
    c.close()


def make_chains(ctx, seed=700):
    out = []
    for k, (dims, m) in enumerate(CHAINS):
        U = dev.random_links(dims, 0.3, seed=seed + k)
        out.append(dict(dims=dims, mass=m, U=U, g=dev.Gauge(ctx, dims, U)))
    return out


def test_service_results_match_solo_solves(ctx):
    chains = make_chains(ctx)
    svc = dev.Service(0, tol=1e-10)
    try:
        ids = [svc.add_chain(c["g"], c["mass"], max_inflight=2) for c in chains]
        rng = np.random.default_rng(1)
        reqs = []
        for r in range(14):
            k = int(rng.integers(len(chains)))
            c = chains[k]
            b = dev.random_spinor(c["dims"], seed=900 + r)
            rf, xf = dev.Field(c["g"], b), dev.Field(c["g"])
            reqs.append((k, b, rf, xf, svc.submit(ids[k], rf, xf)))
        for k, b, rf, xf, t in reqs:
            st = svc.wait(t, timeout=120)
            c = chains[k]
            x = xf.download()
            res = b - oc.apply_ddag(c["dims"], c["mass"], c["U"], x)
            assert np.sqrt(res @ res / (b @ b)) < 1e-10
            assert st["seconds"] > 0 and st["status"] == 0
            # the same request solved alone (batched engine, one system, one RHS)
            xs = dev.Field(c["g"])
            dev.cg_batch(ctx, [(c["g"], c["mass"], [rf], [xs])], tol=1e-10, precision="f32")
            assert np.array_equal(xs.download(), x)
    finally:
        svc.close()


def test_service_queues_beyond_pool(ctx):
    chains = make_chains(ctx, seed=800)
    svc = dev.Service(0, tol=1e-10)
    try:
        cid = svc.add_chain(chains[0]["g"], chains[0]["mass"], max_inflight=1)
        fields = []
        for r in range(5):
            b = dev.random_spinor(chains[0]["dims"], seed=1000 + r)
            rf, xf = dev.Field(chains[0]["g"], b), dev.Field(chains[0]["g"])
            fields.append((b, rf, xf, svc.submit(cid, rf, xf)))
        got = {}
        while len(got) < 5:
            for t, st in svc.poll(timeout=30.0):
                got[t] = st
        for b, rf, xf, t in fields:
            assert got[t]["status"] == 0 and got[t]["true_rel_resid"] < 1e-10
    finally:
        svc.close()


def test_service_memory_budget_refuses_chain(ctx):
    chains = make_chains(ctx, seed=850)
    svc = dev.Service(0, mem_budget=1 << 20, tol=1e-10)  # 1 MiB: an 8^4 pool does not fit
    try:
        with pytest.raises(SfhError):
            svc.add_chain(chains[1]["g"], chains[1]["mass"], max_inflight=4)
    finally:
        svc.close()


def test_service_update_chain_uses_new_links(ctx):
    """A chain whose links change between requests (an MD step) is refreshed
    with update_chain; the next result equals a solo solve on the new links
    bit for bit, and differs from the one on the old links.  Updating a
    chain with requests queued or in flight is refused."""
    chains = make_chains(ctx, seed=870)
    c = chains[1]
    svc = dev.Service(0, tol=1e-10)
    try:
        cid = svc.add_chain(c["g"], c["mass"], max_inflight=2)
        b = dev.random_spinor(c["dims"], seed=871)
        rf, x0, x1 = dev.Field(c["g"], b), dev.Field(c["g"]), dev.Field(c["g"])
        # eight requests on two slots: some are still queued when update_chain runs
        busy = [dev.Field(c["g"]) for _ in range(8)]
        tickets = [svc.submit(cid, rf, xb) for xb in busy]
        with pytest.raises(SfhError):
            svc.update_chain(cid, c["g"], c["mass"])
        for t in tickets:
            svc.wait(t, timeout=120)
        svc.wait(svc.submit(cid, rf, x0), timeout=120)
        # an MD step on the chain's own device links, then the refresh
        P = dev.random_momenta(c["dims"], seed=872)
        c["g"].link_update(0.05, P)
        U1 = c["g"].download()
        svc.update_chain(cid, c["g"], c["mass"])
        svc.wait(svc.submit(cid, rf, x1), timeout=120)
        x = x1.download()
        res = b - oc.apply_ddag(c["dims"], c["mass"], U1, x)
        assert np.sqrt(res @ res / (b @ b)) < 1e-10
        xs = dev.Field(c["g"])
        dev.cg_batch(ctx, [(c["g"], c["mass"], [rf], [xs])], tol=1e-10, precision="f32")
        assert np.array_equal(xs.download(), x)
        assert not np.array_equal(x0.download(), x)
    finally:
        svc.close()


def test_service_remove_chain_reuses_slots(ctx):
    """A removed chain's solve slots serve the next chain of the same lattice
    without new device memory; the removed chain refuses requests."""
    chains = make_chains(ctx, seed=880)
    svc = dev.Service(0, tol=1e-10)
    try:
        a = chains[1]
        ida = svc.add_chain(a["g"], a["mass"], max_inflight=2)
        fields = [(dev.Field(a["g"], dev.random_spinor(a["dims"], seed=881 + j)), dev.Field(a["g"])) for j in range(2)]
        for t in [svc.submit(ida, r, x) for r, x in fields]:
            assert svc.wait(t, timeout=120)["status"] == 0
        held, _ = svc.memory()
        assert held > 0
        svc.remove_chain(ida)
        with pytest.raises(SfhError):
            svc.submit(ida, *fields[0])
        g2 = dev.Gauge(ctx, a["dims"], dev.random_links(a["dims"], 0.3, seed=889))
        idb = svc.add_chain(g2, 0.04, max_inflight=2)
        b = dev.random_spinor(a["dims"], seed=890)
        rf, xf = dev.Field(g2, b), dev.Field(g2)
        st = svc.wait(svc.submit(idb, rf, xf), timeout=120)
        assert st["status"] == 0 and st["true_rel_resid"] < 1e-10
        assert svc.memory()[0] == held  # no new slot memory
    finally:
        svc.close()


def test_service_admission_by_memory_budget(ctx):
    """A budget of two slots for a chain allowed eight in flight: requests
    beyond two wait for a slot (each refusal is counted) and all complete,
    each equal to its solo solve."""
    chains = make_chains(ctx, seed=900)
    c = chains[1]
    slot = None
    probe = dev.Service(0, tol=1e-10)
    try:
        pid = probe.add_chain(c["g"], c["mass"], max_inflight=1)
        rf, xf = dev.Field(c["g"], dev.random_spinor(c["dims"], seed=901)), dev.Field(c["g"])
        probe.wait(probe.submit(pid, rf, xf), timeout=120)
        slot, _ = probe.memory()
    finally:
        probe.close()
    svc = dev.Service(0, mem_budget=2 * slot + slot // 2, tol=1e-10)
    try:
        cid = svc.add_chain(c["g"], c["mass"], max_inflight=8)
        reqs = []
        for j in range(6):
            b = dev.random_spinor(c["dims"], seed=910 + j)
            rf, xf = dev.Field(c["g"], b), dev.Field(c["g"])
            reqs.append((rf, xf, svc.submit(cid, rf, xf)))
        for rf, xf, t in reqs:
            assert svc.wait(t, timeout=120)["status"] == 0
        held, refused = svc.memory()
        assert held == 2 * slot and refused >= 1
        for rf, xf, t in reqs:
            xs = dev.Field(c["g"])
            dev.cg_batch(ctx, [(c["g"], c["mass"], [rf], [xs])], tol=1e-10, precision="f32")
            assert np.array_equal(xs.download(), xf.download())
    finally:
        svc.close()
