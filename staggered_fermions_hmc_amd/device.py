"""Object wrapper over the C ABI: a context (one HIP stream), device-resident
gauge fields and spinor fields.  All host arrays are float64 in the
reference's AoS layouts (links V*4*18, spinors V*6)."""
import ctypes

import numpy as np

from . import _abi
from ._abi import SfhError, f64, ptr

# Reference exception classes by status (Fermions.cpp: std::runtime_error for
# CG and the hermiticity guard, std::invalid_argument for the exponential).
_EXC = {5: RuntimeError, 6: RuntimeError, 7: RuntimeError, 8: RuntimeError, 9: RuntimeError,
        10: ValueError, 11: ValueError, 12: RuntimeError, 1: ValueError}


class SolverError(SfhError):
    pass


def _raise(status, msg):
    base = _EXC.get(status)
    if base is None:
        raise SfhError(status, msg)
    cls = type(f"Sfh{base.__name__}", (SfhError, base), {})
    raise cls(status, msg)


def _true_mode(true_residual):
    """sfh_cg_params.compute_true_residual: False / 0 none, True / 1 report
    the FP64 true residual (FP64 CG), "stop" / 2 also stop on it."""
    if true_residual == "stop":
        return 2
    if isinstance(true_residual, bool) or true_residual is None:
        return int(bool(true_residual))
    return int(true_residual)


class Context:
    def __init__(self, device=0):
        self._lib = _abi.lib()
        h = ctypes.c_void_p()
        st = self._lib.sfh_create(int(device), ctypes.byref(h))
        if st:
            raise SfhError(st, "sfh_create failed (is a GPU visible?)")
        self.h = h

    def check(self, st):
        if st:
            msg = self._lib.sfh_last_error(self.h)
            _raise(st, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self._lib.sfh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def synchronize(self):
        self.check(self._lib.sfh_synchronize(self.h))

    def set_profiling(self, on):
        self.check(self._lib.sfh_set_profiling(self.h, int(on)))

    def set_graph_replay(self, on):
        """True (default): CG chunks replay as HIP graphs; False: direct launches."""
        self.check(self._lib.sfh_set_graph_replay(self.h, int(bool(on))))

    def kernel_time(self, name):
        ms = ctypes.c_double()
        n = ctypes.c_long()
        self.check(self._lib.sfh_kernel_time(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def inner(self, a, b):
        a = f64(a); b = f64(b)
        assert a.size == b.size
        out = ctypes.c_double()
        self.check(self._lib.sfh_inner(self.h, a.size, ptr(a), ptr(b), ctypes.byref(out)))
        return out.value

    def exp_iq(self, Q18):
        Q = f64(Q18)
        out = np.empty(18)
        self.check(self._lib.sfh_exp_iq(self.h, ptr(Q), ptr(out)))
        return out


class Gauge:
    """Device-resident links for one lattice (dims = L0, L1, L2, T)."""

    def __init__(self, ctx, dims, links=None):
        self.ctx = ctx
        self.dims = tuple(int(d) for d in dims)
        self.V = int(np.prod(self.dims))
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.sfh_gauge_create(ctx.h, ptr(_abi.dims_arr(self.dims)), ctypes.byref(h)))
        self.h = h
        if links is not None:
            self.upload(links)

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx._lib.sfh_gauge_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, links):
        a = f64(links)
        assert a.size == self.V * 72
        self.ctx.check(self.ctx._lib.sfh_gauge_upload(self.ctx.h, self.h, ptr(a)))

    def download(self):
        out = np.empty(self.V * 72)
        self.ctx.check(self.ctx._lib.sfh_gauge_download(self.ctx.h, self.h, ptr(out)))
        return out

    def _spinor(self, a):
        a = f64(a)
        assert a.size == self.V * 6, (a.size, self.V * 6)
        return a

    def apply_d(self, mass, psi, dagger=False, precision="f64"):
        """D psi (or D^dagger psi); precision "f32" / "f16" applies the mixed
        CG's inner-iteration operator (storage in that precision)."""
        a = self._spinor(psi)
        out = np.empty_like(a)
        if precision == "f64":
            self.ctx.check(self.ctx._lib.sfh_apply_d(self.ctx.h, self.h, mass, int(bool(dagger)), ptr(a), ptr(out)))
        else:
            self.ctx.check(self.ctx._lib.sfh_apply_d_prec(self.ctx.h, self.h, mass, int(bool(dagger)),
                                                          _abi.PREC[precision], ptr(a), ptr(out)))
        return out

    def apply_ddag(self, mass, psi):
        a = self._spinor(psi)
        out = np.empty_like(a)
        self.ctx.check(self.ctx._lib.sfh_apply_ddag(self.ctx.h, self.h, mass, ptr(a), ptr(out)))
        return out

    def cg(self, phi, mass, tol=1e-12, max_iter=50000, precision="f64", reliable_delta=0.0,
           true_residual=False):
        a = self._spinor(phi)
        x = np.empty_like(a)
        prm = _abi.CGParams(mass, tol, max_iter, _abi.PREC[precision], reliable_delta, _true_mode(true_residual))
        stats = _abi.CGStats()
        self.ctx.check(self.ctx._lib.sfh_cg(self.ctx.h, self.h, ctypes.byref(prm), ptr(a), ptr(x),
                                            ctypes.byref(stats)))
        return x, stats.as_dict()

    def cg_field(self, phi_field, x_field, mass, tol=1e-12, max_iter=50000, precision="f64", reliable_delta=0.0,
                 true_residual=False):
        prm = _abi.CGParams(mass, tol, max_iter, _abi.PREC[precision], reliable_delta, _true_mode(true_residual))
        stats = _abi.CGStats()
        self.ctx.check(self.ctx._lib.sfh_cg_field(self.ctx.h, self.h, ctypes.byref(prm), phi_field.h,
                                                  x_field.h, ctypes.byref(stats)))
        return stats.as_dict()

    def force(self, beta, c11, c12, X=None, Y=None, which=3):
        out = np.empty(self.V * 72)
        Xa = self._spinor(X) if X is not None else None
        Ya = self._spinor(Y) if Y is not None else None
        self.ctx.check(self.ctx._lib.sfh_force(self.ctx.h, self.h, beta, c11, c12,
                                               ptr(Xa) if Xa is not None else None,
                                               ptr(Ya) if Ya is not None else None, int(which), ptr(out)))
        return out

    def tune_hop(self, precision="f64", n=20):
        """Time the hop forms on this lattice, keep the faster; returns
        (variant, average microseconds)."""
        v = ctypes.c_int()
        us = ctypes.c_double()
        self.ctx.check(self.ctx._lib.sfh_tune_hop(self.ctx.h, self.h, _abi.PREC[precision], int(n), ctypes.byref(v),
                                                  ctypes.byref(us)))
        return v.value, us.value

    def time_hop(self, precision="f64", with_dot=False, n=50):
        """Average microseconds of the solver's hop kernel on this lattice,
        n back-to-back launches between HIP events on the context stream."""
        us = ctypes.c_double()
        self.ctx.check(self.ctx._lib.sfh_time_hop(self.ctx.h, self.h, _abi.PREC[precision], int(bool(with_dot)),
                                                  int(n), ctypes.byref(us)))
        return us.value

    def tune_cg_form(self, precision="f32", n=20):
        """Time both mixed-iteration forms on this lattice, keep the faster;
        returns (form, microseconds per iteration)."""
        f = ctypes.c_int()
        us = ctypes.c_double()
        self.ctx.check(self.ctx._lib.sfh_tune_cg_form(self.ctx.h, self.h, _abi.PREC[precision], int(n),
                                                      ctypes.byref(f), ctypes.byref(us)))
        return f.value, us.value

    def time_cg_kernels(self, precision="f32", n=50):
        """Average microseconds of the two steps of one mixed-precision CG
        iteration on this lattice (update + first hop, second hop + dots)."""
        ua, ub = ctypes.c_double(), ctypes.c_double()
        self.ctx.check(self.ctx._lib.sfh_time_cg_kernels(self.ctx.h, self.h, _abi.PREC[precision], int(n),
                                                         ctypes.byref(ua), ctypes.byref(ub)))
        return ua.value, ub.value

    def time_force(self, beta=4.0, c11=5.0 / 3.0, c12=-1.0 / 12.0, which=1, n=10):
        """Average microseconds of the force kernel as the leapfrog runs it
        (momentum kick fused, no force field stored)."""
        us = ctypes.c_double()
        self.ctx.check(self.ctx._lib.sfh_time_force(self.ctx.h, self.h, float(beta), float(c11), float(c12),
                                                    int(which), int(n), ctypes.byref(us)))
        return us.value

    def action(self, beta, c11, c12):
        out = ctypes.c_double()
        self.ctx.check(self.ctx._lib.sfh_action(self.ctx.h, self.h, beta, c11, c12, ctypes.byref(out)))
        return out.value

    def avg_plaq(self):
        out = ctypes.c_double()
        self.ctx.check(self.ctx._lib.sfh_avg_plaq(self.ctx.h, self.h, ctypes.byref(out)))
        return out.value

    def polyakov_grid(self):
        """Tr prod_t U_3 at each spatial site, complex [(x*L1 + y)*L2 + z]."""
        out = np.empty(2 * self.dims[0] * self.dims[1] * self.dims[2])
        self.ctx.check(self.ctx._lib.sfh_polyakov_grid(self.ctx.h, self.h, ptr(out)))
        return out.view(np.complex128)

    def polyakov_correlators(self, r2_max, r2_first=1):
        """C(r^2) for r^2 = r2_first .. r2_max from the device links (grid and
        sums on the device): complex array of r2_max - r2_first + 1 values."""
        n = int(r2_max) - int(r2_first) + 1
        out = np.empty(2 * n)
        self.ctx.check(self.ctx._lib.sfh_polyakov_correlators(self.ctx.h, self.h, int(r2_first), n, ptr(out)))
        return out.view(np.complex128)

    def heatbath(self, beta, c11, c12, seed, sweep0=0, nsweeps=1):
        """nsweeps Cabibbo-Marinari heatbath sweeps of the Symanzik action on
        the device-resident links (heatbath_update, Lattice.cpp:725-757),
        sweep indices sweep0, sweep0 + 1, ...  Returns the per-sweep
        (average plaquette, action density) arrays the reference returns."""
        pl = np.empty(max(int(nsweeps), 1))
        ac = np.empty(max(int(nsweeps), 1))
        self.ctx.check(self.ctx._lib.sfh_heatbath(self.ctx.h, self.h, float(beta), float(c11), float(c12), int(seed),
                                                  int(sweep0), int(nsweeps), ptr(pl), ptr(ac)))
        return pl[:nsweeps], ac[:nsweeps]

    def hex_smear(self, alpha=(0.95, 0.76, 0.38), nsteps=1):
        """nsteps HEX smearing steps of the device-resident links in place
        (Apply_HEX_smearing, Smearing.cpp:304-311); alpha = (alpha1, alpha2,
        alpha3) as in Smearing.h:7-9."""
        a = f64(np.asarray(alpha, dtype=np.float64))
        self.ctx.check(self.ctx._lib.sfh_hex_smear(self.ctx.h, self.h, ptr(a), int(nsteps)))

    def tr_p2(self, P):
        a = f64(P)
        assert a.size == self.V * 72
        out = ctypes.c_double()
        self.ctx.check(self.ctx._lib.sfh_tr_p2(self.ctx.h, self.h, ptr(a), ctypes.byref(out)))
        return out.value

    def link_update(self, eps, P, check=True):
        a = f64(P)
        assert a.size == self.V * 72
        self.ctx.check(self.ctx._lib.sfh_link_update(self.ctx.h, self.h, eps, ptr(a), int(bool(check))))

    def md_params(self, *, mass=0.0, tol=1e-12, max_iter=50000, precision="f64", beta=0.0, c11=1.0, c12=0.0,
                  eps=0.01, nsteps=100, which=3, check=True):
        return _abi.MDParams(mass, tol, max_iter, _abi.PREC[precision], beta, c11, c12, eps, nsteps, which,
                             int(bool(check)), 0)

    def leapfrog(self, P, phi=None, **kw):
        prm = self.md_params(**kw)
        Pa = f64(P).copy()
        X = np.zeros(self.V * 6)
        ph = self._spinor(phi) if phi is not None else None
        self.ctx.check(self.ctx._lib.sfh_leapfrog(self.ctx.h, self.h, ctypes.byref(prm),
                                                  ptr(ph) if ph is not None else None, ptr(Pa), ptr(X)))
        return Pa, X

    def hmc_trajectory(self, chi, P0, r_uniform, **kw):
        prm = self.md_params(**kw)
        c = self._spinor(chi)
        P = f64(P0)
        H = np.zeros(7)
        acc = ctypes.c_int()
        self.ctx.check(self.ctx._lib.sfh_hmc_trajectory(self.ctx.h, self.h, ctypes.byref(prm), ptr(c), ptr(P),
                                                        float(r_uniform), ptr(H), ctypes.byref(acc)))
        return H, bool(acc.value)


class Field:
    """Device-resident spinor for the lattice of a Gauge."""

    def __init__(self, gauge, host=None):
        self.ctx = gauge.ctx
        self.V = gauge.V
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx._lib.sfh_field_create(self.ctx.h, gauge.h, ctypes.byref(h)))
        self.h = h
        if host is not None:
            self.upload(host)

    def upload(self, host):
        a = f64(host)
        assert a.size == self.V * 6
        self.ctx.check(self.ctx._lib.sfh_field_upload(self.ctx.h, self.h, ptr(a)))

    def download(self):
        out = np.empty(self.V * 6)
        self.ctx.check(self.ctx._lib.sfh_field_download(self.ctx.h, self.h, ptr(out)))
        return out

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx._lib.sfh_field_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def polyakov_correlators_grid(ctx, spatial, grid, r2_max, r2_first=1):
    """C(r^2), r^2 = r2_first .. r2_max, of a host grid (complex, L0*L1*L2)."""
    g = np.ascontiguousarray(np.asarray(grid, dtype=np.complex128)).view(np.float64)
    sp = np.ascontiguousarray(np.asarray(spatial, dtype=np.int32))
    assert sp.shape == (3,) and g.size == 2 * int(np.prod(sp))
    n = int(r2_max) - int(r2_first) + 1
    out = np.empty(2 * n)
    ctx.check(ctx._lib.sfh_polyakov_correlators_grid(ctx.h, ptr(sp), ptr(g), int(r2_first), n, ptr(out)))
    return out.view(np.complex128)


def save_links(ctx, gauges, path):
    """Checkpoint in the reference's byte format (Lattice.cpp:810-828)."""
    arr = (ctypes.c_void_p * len(gauges))(*[g.h.value for g in gauges])
    ctx.check(ctx._lib.sfh_gauge_save(ctx.h, arr, len(gauges), str(path).encode()))


def load_links(ctx, gauges, path):
    """Restore a checkpoint written by save_links or the reference (Lattice.cpp:830-847)."""
    arr = (ctypes.c_void_p * len(gauges))(*[g.h.value for g in gauges])
    ctx.check(ctx._lib.sfh_gauge_load(ctx.h, arr, len(gauges), str(path).encode()))


def _batch_array(systems):
    n = len(systems)
    arr = (_abi.BatchSystem * n)()
    keep = []
    total = 0
    for k, (g, mass, rhs, xs) in enumerate(systems):
        assert len(rhs) == len(xs) and 1 <= len(rhs) <= 8
        r = (ctypes.c_void_p * len(rhs))(*[f.h.value for f in rhs])
        x = (ctypes.c_void_p * len(xs))(*[f.h.value for f in xs])
        keep += [r, x]
        arr[k].gauge = g.h
        arr[k].mass = mass
        arr[k].nrhs = len(rhs)
        arr[k].rhs = ctypes.cast(r, ctypes.POINTER(ctypes.c_void_p))
        arr[k].x = ctypes.cast(x, ctypes.POINTER(ctypes.c_void_p))
        total += len(rhs)
    return arr, keep, total


def cg_batch(ctx, systems, tol=1e-10, max_iter=100000, precision="f32", reliable_delta=0.0):
    """Solve many systems together.  systems: list of (gauge, mass, [rhs
    Field, ...], [x Field, ...]).  Returns one stats dict per right-hand side,
    systems in order."""
    arr, keep, total = _batch_array(systems)
    prm = _abi.CGParams(0.0, tol, max_iter, _abi.PREC[precision], reliable_delta, 0)
    stats = (_abi.CGStats * total)()
    ctx.check(ctx._lib.sfh_cg_batch(ctx.h, len(systems), arr, ctypes.byref(prm), stats))
    return [s.as_dict() for s in stats]


def tune_batch(ctx, systems, tol=1e-10, max_iter=100000, precision="f32", reliable_delta=0.0, shape=None):
    """Autotune the batched hop's launch shape for this batch (kept in the
    context for batches of the same shape); the x fields are work space.
    shape=(sites per block, RHS per block) records that shape instead.
    Returns (sites per block, right-hand sides per block, microseconds)."""
    arr, keep, _ = _batch_array(systems)
    prm = _abi.CGParams(0.0, tol, max_iter, _abi.PREC[precision], reliable_delta, 0)
    th, tile, us = ctypes.c_int(shape[0] if shape else 0), ctypes.c_int(shape[1] if shape else 0), ctypes.c_double()
    ctx.check(ctx._lib.sfh_tune_batch(ctx.h, len(systems), arr, ctypes.byref(prm), ctypes.byref(th),
                                      ctypes.byref(tile), ctypes.byref(us)))
    return th.value, tile.value, us.value


class Service:
    """Continuous-batching solve service (sfh_service_*): chains register
    device-resident links; requests are (rhs Field, x Field) pairs."""

    def __init__(self, device=0, mem_budget=0, tol=1e-10, max_iter=100000):
        self._lib = _abi.lib()
        h = ctypes.c_void_p()
        st = self._lib.sfh_service_create(int(device), int(mem_budget), float(tol), int(max_iter), ctypes.byref(h))
        if st:
            raise SfhError(st, "sfh_service_create")
        self.h = h

    def _check(self, st):
        if st:
            msg = self._lib.sfh_service_last_error(self.h)
            _raise(st, msg.decode() if msg else "")

    def add_chain(self, gauge, mass, max_inflight=4):
        cid = ctypes.c_int()
        self._check(self._lib.sfh_service_add_chain(self.h, gauge.h, float(mass), int(max_inflight),
                                                    ctypes.byref(cid)))
        return cid.value

    def update_chain(self, chain, gauge, mass):
        """New links (gauge, same lattice) or mass for a chain with nothing
        queued or in flight."""
        self._check(self._lib.sfh_service_update_chain(self.h, int(chain), gauge.h, float(mass)))

    def remove_chain(self, chain):
        self._check(self._lib.sfh_service_remove_chain(self.h, int(chain)))

    def memory(self):
        """(bytes of solve-slot memory, slots refused for memory so far)."""
        b, n = ctypes.c_size_t(), ctypes.c_long()
        self._check(self._lib.sfh_service_memory(self.h, ctypes.byref(b), ctypes.byref(n)))
        return b.value, n.value

    def submit(self, chain, rhs, x):
        t = ctypes.c_longlong()
        self._check(self._lib.sfh_service_submit(self.h, int(chain), rhs.h, x.h, ctypes.byref(t)))
        return t.value

    def wait(self, ticket, timeout=60.0):
        st = _abi.CGStats()
        self._check(self._lib.sfh_service_wait(self.h, int(ticket), float(timeout), ctypes.byref(st)))
        return st.as_dict()

    def poll(self, timeout=1.0, max_n=256):
        tk = (ctypes.c_longlong * max_n)()
        sts = (_abi.CGStats * max_n)()
        n = ctypes.c_int()
        self._check(self._lib.sfh_service_poll(self.h, float(timeout), int(max_n), tk, sts, ctypes.byref(n)))
        return [(tk[i], sts[i].as_dict()) for i in range(n.value)]

    def close(self):
        if getattr(self, "h", None):
            self._lib.sfh_service_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_hop_variant(precision, variant):
    """0 one lane per site (default), 1 four lanes per site, 2 / 3 one lane
    per site stepping one / two directions at a time, -1 the default."""
    st = _abi.lib().sfh_set_hop_variant(_abi.PREC[precision], int(variant))
    if st:
        raise SfhError(st, "sfh_set_hop_variant")


def set_cg_form(precision, form):
    """Mixed-precision iteration form: 0 fused (two kernels), 1 split (update
    kernel + plain hop + second hop), -1 by lattice size (default)."""
    st = _abi.lib().sfh_set_cg_form(_abi.PREC[precision], int(form))
    if st:
        raise SfhError(st, "sfh_set_cg_form")


def random_links(dims, eps, seed):
    out = np.empty(int(np.prod(dims)) * 72)
    st = _abi.lib().sfh_random_links(ptr(_abi.dims_arr(dims)), float(eps), int(seed), ptr(out))
    if st:
        raise SfhError(st, "sfh_random_links")
    return out


def random_spinor(dims, seed, sigma=np.sqrt(0.5)):
    out = np.empty(int(np.prod(dims)) * 6)
    st = _abi.lib().sfh_random_spinor(ptr(_abi.dims_arr(dims)), float(sigma), int(seed), ptr(out))
    if st:
        raise SfhError(st, "sfh_random_spinor")
    return out


def random_momenta(dims, seed):
    out = np.empty(int(np.prod(dims)) * 72)
    st = _abi.lib().sfh_random_momenta(ptr(_abi.dims_arr(dims)), int(seed), ptr(out))
    if st:
        raise SfhError(st, "sfh_random_momenta")
    return out
