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
