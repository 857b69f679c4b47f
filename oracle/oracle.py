"""ctypes binding of the FP64 CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The product path never
imports it.

Arrays use the reference's array-of-structures layouts (Lattice.cpp:135-151,
Fermions.cpp:99-116): links/momenta/forces are float64 of length V*4*18,
spinors float64 of length V*6, site order s = ((n0*L1+n1)*L2+n2)*L3+n3.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "oc_apply_d": (_I, [_P, _D, _I, _P, _P, _P]),
            "oc_apply_ddag": (_I, [_P, _D, _P, _P, _P]),
            "oc_inner": (_D, [ctypes.c_long, _P, _P]),
            "oc_cg": (_I, [_P, _D, _D, _I, _P, _P, _P, _P, _P]),
            "oc_time_cg_iterations": (_D, [_P, _D, _P, _P, _I]),
            "oc_staples": (_I, [_P, _P, _I, _I, _P, _P]),
            "oc_force": (_I, [_P, _P, _P, _P, _D, _D, _D, _I, _P]),
            "oc_action": (_D, [_P, _P, _D, _D, _D]),
            "oc_avg_plaq": (_D, [_P, _P]),
            "oc_tr_p2": (_D, [_P, _P]),
            "oc_exp_iq": (_I, [_P, _P]),
            "oc_link_update": (_I, [_P, _D, _P, _P, _I]),
            "oc_leapfrog": (_I, [_P, _D, _D, _I, _D, _D, _D, _D, _I, _I, _I, _P, _P, _P, _P]),
            "oc_hmc_trajectory": (_I, [_P, _D, _D, _I, _D, _D, _D, _D, _I, _P, _P, _D, _P, _P, _P]),
            "oc_set_threads": (_I, [_I]),
            "oc_polyakov_grid": (_I, [_P, _P, _P]),
            "oc_polyakov_correlator": (_I, [_P, _P, _I, _P]),
            "oc_philox": (_I, [_P, _P, _P]),
            "oc_heatbath_x0": (_I, [_D, _D, ctypes.c_ulonglong, ctypes.c_uint, _I, _P]),
            "oc_heatbath_subgroup": (_I, [_P, _P, _D, _I, ctypes.c_ulonglong, ctypes.c_uint, ctypes.c_uint, _P,
                                          _P]),
            "oc_heatbath": (_I, [_P, _P, _D, _D, _D, ctypes.c_ulonglong, ctypes.c_uint, _I, _I, _P, _P]),
            "oc_hex_smear": (_I, [_P, _P, _P, _I]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dims(dims):
    d = np.ascontiguousarray(np.asarray(dims, dtype=np.int32))
    assert d.shape == (4,) and np.all(d % 2 == 0) and np.all(d >= 2)
    return d


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def volume(dims):
    return int(np.prod(dims))


class OracleError(RuntimeError):
    def __init__(self, status, what):
        super().__init__(f"{what}: oracle status {status}")
        self.status = status


def set_threads(n):
    return lib().oc_set_threads(int(n))


def apply_d(dims, mass, U, psi, dagger=False):
    d = _dims(dims); U = _f64(U); psi = _f64(psi)
    out = np.empty_like(psi)
    lib().oc_apply_d(_p(d), mass, int(bool(dagger)), _p(U), _p(psi), _p(out))
    return out


def apply_ddag(dims, mass, U, psi):
    d = _dims(dims); U = _f64(U); psi = _f64(psi)
    out = np.empty_like(psi)
    lib().oc_apply_ddag(_p(d), mass, _p(U), _p(psi), _p(out))
    return out


def inner(a, b):
    a = _f64(a); b = _f64(b)
    return lib().oc_inner(a.size, _p(a), _p(b))


def cg(dims, mass, U, phi, tol=1e-12, max_iter=50000, raise_on_error=True):
    d = _dims(dims); U = _f64(U); phi = _f64(phi)
    X = np.zeros_like(phi)
    it = ctypes.c_int(0)
    rr = ctypes.c_double(0)
    st = lib().oc_cg(_p(d), mass, tol, max_iter, _p(U), _p(phi), _p(X), ctypes.byref(it), ctypes.byref(rr))
    if st and raise_on_error:
        raise OracleError(st, "cg")
    return X, it.value, rr.value, st


def time_cg_iterations(dims, mass, U, phi, n):
    d = _dims(dims); U = _f64(U); phi = _f64(phi)
    return lib().oc_time_cg_iterations(_p(d), mass, _p(U), _p(phi), int(n))


def staples(dims, U, site, mu):
    d = _dims(dims); U = _f64(U)
    A1 = np.empty(18); A2 = np.empty(18)
    lib().oc_staples(_p(d), _p(U), int(site), int(mu), _p(A1), _p(A2))
    return A1, A2


def force(dims, U, X, Y, beta, c11, c12, which=3):
    d = _dims(dims); U = _f64(U)
    V = volume(dims)
    X = _f64(X if X is not None else np.zeros(V * 6))
    Y = _f64(Y if Y is not None else np.zeros(V * 6))
    F = np.empty(V * 72)
    lib().oc_force(_p(d), _p(U), _p(X), _p(Y), beta, c11, c12, int(which), _p(F))
    return F


def action(dims, U, beta, c11, c12):
    d = _dims(dims); U = _f64(U)
    return lib().oc_action(_p(d), _p(U), beta, c11, c12)


def avg_plaq(dims, U):
    d = _dims(dims); U = _f64(U)
    return lib().oc_avg_plaq(_p(d), _p(U))


def tr_p2(dims, P):
    d = _dims(dims); P = _f64(P)
    return lib().oc_tr_p2(_p(d), _p(P))


def exp_iq(Q18):
    Q = _f64(Q18)
    out = np.empty(18)
    st = lib().oc_exp_iq(_p(Q), _p(out))
    if st:
        raise OracleError(st, "exp_iq")
    return out


def link_update(dims, eps, P, U, check=True):
    d = _dims(dims); P = _f64(P); U = _f64(U).copy()
    st = lib().oc_link_update(_p(d), eps, _p(P), _p(U), int(bool(check)))
    if st:
        raise OracleError(st, "link_update")
    return U


def leapfrog(dims, mass, U, P, phi, *, beta, c11, c12, eps, nsteps, which=3, tol=1e-12, max_iter=50000,
             check=True):
    d = _dims(dims)
    U = _f64(U).copy(); P = _f64(P).copy()
    V = volume(dims)
    phi = _f64(phi if phi is not None else np.zeros(V * 6))
    X = np.zeros(V * 6)
    st = lib().oc_leapfrog(_p(d), mass, tol, max_iter, beta, c11, c12, eps, int(nsteps), int(which),
                           int(bool(check)), _p(phi), _p(U), _p(P), _p(X))
    if st:
        raise OracleError(st, "leapfrog")
    return U, P, X


def hmc_trajectory(dims, mass, U, chi, P, r_uniform, *, beta, c11, c12, eps, nsteps, tol=1e-12,
                   max_iter=50000):
    d = _dims(dims)
    U = _f64(U).copy(); chi = _f64(chi); P = _f64(P)
    H = np.zeros(7)
    acc = ctypes.c_int(0)
    st = lib().oc_hmc_trajectory(_p(d), mass, tol, max_iter, beta, c11, c12, eps, int(nsteps), _p(chi), _p(P),
                                 float(r_uniform), _p(U), _p(H), ctypes.byref(acc))
    if st:
        raise OracleError(st, "hmc_trajectory")
    return U, H, bool(acc.value)


def polyakov_grid(dims, U):
    d = _dims(dims); U = _f64(U)
    out = np.empty(2 * int(d[0] * d[1] * d[2]))
    lib().oc_polyakov_grid(_p(d), _p(U), _p(out))
    return out.view(np.complex128)


def polyakov_correlator(spatial, grid, r_squared):
    """correlator_over_fixed_distance_fast (Lattice.cpp:886-948) of a complex
    grid [(x*L1 + y)*L2 + z] for one r^2, in the reference's summation order."""
    sp = np.ascontiguousarray(np.asarray(spatial, dtype=np.int32))
    g = np.ascontiguousarray(np.asarray(grid, dtype=np.complex128)).view(np.float64)
    assert sp.shape == (3,) and g.size == 2 * int(np.prod(sp))
    out = np.empty(2)
    lib().oc_polyakov_correlator(_p(sp), _p(g), int(r_squared), _p(out))
    return complex(out[0], out[1])


def philox(key, ctr):
    """Philox4x32-10 block of a 2-word key and 4-word counter (uint32)."""
    k = np.ascontiguousarray(np.asarray(key, dtype=np.uint32))
    c = np.ascontiguousarray(np.asarray(ctr, dtype=np.uint32))
    out = np.empty(4, dtype=np.uint32)
    lib().oc_philox(_p(k), _p(c), _p(out))
    return out


def heatbath_x0(a, beta, n, seed=1, sweep=0):
    """n draws of the SU(2) heatbath x0 at (a, beta): density
    sqrt(1 - x0^2) exp(a beta x0) (SU3_Sampling.cpp:17-29)."""
    out = np.empty(int(n))
    st = lib().oc_heatbath_x0(float(a), float(beta), int(seed), int(sweep), int(n), _p(out))
    if st:
        raise OracleError(st, "heatbath_x0")
    return out


def heatbath_subgroup(U18, A18, beta, sub, seed=1, link=0, sweep=0):
    """One subgroup update U <- embed(R) U of the Cabibbo-Marinari step (sub
    0, 1, 2 = rows/columns (0,1), (0,2), (1,2)); returns (U', a, x0)."""
    U = _f64(U18); A = _f64(A18)
    out = np.empty(18); ax = np.zeros(2)
    st = lib().oc_heatbath_subgroup(_p(U), _p(A), float(beta), int(sub), int(seed), int(link), int(sweep),
                                    _p(out), _p(ax))
    if st:
        raise OracleError(st, "heatbath_subgroup")
    return out, ax[0], ax[1]


def heatbath(dims, U, beta, c11, c12, seed, sweep0=0, nsweeps=1, order=1):
    """nsweeps Cabibbo-Marinari sweeps (Lattice.cpp:725-757).  order 1 is
    the device's colour order, 0 the reference's lexicographic order.
    Returns (U', per-sweep average plaquette, per-sweep action density)."""
    d = np.ascontiguousarray(np.asarray(dims, dtype=np.int32))
    U = _f64(U).copy()
    pl = np.empty(int(nsweeps)); ac = np.empty(int(nsweeps))
    st = lib().oc_heatbath(_p(d), _p(U), float(beta), float(c11), float(c12), int(seed), int(sweep0),
                           int(nsweeps), int(order), _p(pl), _p(ac))
    if st:
        raise OracleError(st, "heatbath")
    return U, pl, ac


HEX_ALPHA = (0.95, 0.76, 0.38)  # alpha1, alpha2, alpha3 of Smearing.h:7-9


def hex_smear(dims, U, alpha=HEX_ALPHA, nsteps=1):
    """nsteps HEX smearing steps (Smearing.cpp:76-318); returns new links."""
    d = _dims(dims)
    U = _f64(U).copy()
    a = np.ascontiguousarray(np.asarray(alpha, dtype=np.float64))
    st = lib().oc_hex_smear(_p(d), _p(U), _p(a), int(nsteps))
    if st:
        raise OracleError(st, "hex_smear")
    return U
