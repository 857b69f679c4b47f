"""The allocation-free correlator entry point (sfh_polyakov_correlators_noalloc)
against the allocating one (sfh_polyakov_correlators), through the C ABI:
bit-identical results (same kernels, same summation order) on two calls in a
row, the second on a smaller lattice, and the size limit as an error."""
import numpy as np
import pytest

from staggered_fermions_hmc_amd import _abi
from staggered_fermions_hmc_amd import device as dev

pytestmark = pytest.mark.gpu


def noalloc(ctx, g, r2_first, count):
    out = np.empty(2 * count)
    ctx.check(_abi.lib().sfh_polyakov_correlators_noalloc(ctx.h, g.h, int(r2_first), int(count), out.ctypes.data))
    return out.view(np.complex128)


def test_noalloc_correlators_match_allocating_entry_point():
    ctx = dev.Context(0)
    try:
        for dims, r2_max in (((32, 32, 32, 32), 15), ((16, 16, 16, 8), 7)):
            print("no-alloc correlators", dims, flush=True)
            g = dev.Gauge(ctx, dims, dev.random_links(dims, 0.6, seed=sum(dims)))
            ref = g.polyakov_correlators(r2_max)
            got = noalloc(ctx, g, 1, r2_max)
            assert np.array_equal(got, ref), (dims, np.abs(got - ref).max())
            got2 = noalloc(ctx, g, 3, 2)
            assert np.array_equal(got2, ref[2:4])
        # beyond the stated maximum (256 values of r^2): a clear error
        with pytest.raises(dev.SfhError, match="at most 256"):
            noalloc(ctx, g, 1, 300)
    finally:
        ctx.close()
