"""Field-level mirror of the reference's Fermions.h / Lattice.h API on the
MI355X library.

Same function names, argument meaning and error behaviour as the reference
(Fermions.h:31-70, :105, :130; Lattice.h:75-77; Smearing.h:29-31); host arrays are float64 in
the reference layouts.  The reference's compile-time constants
(Parameters.h:9-24, :51-52) are the fields of :class:`Parameters`, set once
with :func:`configure`.  Each call hands its host arrays to the device and
back, as the reference's signatures require; device-resident use goes through
:mod:`staggered_fermions_hmc_amd.device` instead.
"""
import dataclasses

import numpy as np

from staggered_fermions_hmc_amd import device as _dev


@dataclasses.dataclass
class Parameters:
    """Parameters.h:9-52 as runtime values."""
    Spatial_Size: int = 4
    temporal_size: int = 4
    fermion_mass: float = 1.0
    epsilon: float = 1e-2
    number_of_steps_fermion: int = 100
    tol: float = 1e-12
    max_iter: int = 50000
    c_11: float = 5.0 / 3.0
    c_12: float = -1.0 / 12.0
    alpha1: float = 0.95  # Smearing.h:7-9
    alpha2: float = 0.76
    alpha3: float = 0.38
    precision: str = "f64"

    @property
    def dims(self):
        L = self.Spatial_Size
        return (L, L, L, self.temporal_size)

    @property
    def V(self):
        return self.Spatial_Size ** 3 * self.temporal_size


params = Parameters()
_ctx = None
_seed = [20240601]


def configure(**kw):
    """Set the runtime parameters (the reference recompiles Parameters.h)."""
    global params
    params = dataclasses.replace(params, **kw)
    return params


def context():
    global _ctx
    if _ctx is None:
        _ctx = _dev.Context(0)
    return _ctx


def _gauge(arr):
    return _dev.Gauge(context(), params.dims, arr)


def _next_seed():
    _seed[0] += 1
    return _seed[0]


# ---------------------------------------------------------------- generation
def generate_chi_field():
    """Complex Gaussian, sigma 1/sqrt(2) per real part (Fermions.cpp:294-300)."""
    return _dev.random_spinor(params.dims, _next_seed())


def generate_phi_field(arr, chi=None):
    """phi = D chi (Fermions.cpp:302-314); chi may be given for reproducibility."""
    if chi is None:
        chi = generate_chi_field()
    return _gauge(arr).apply_d(params.fermion_mass, chi)


def generate_conjugate_field_configuration():
    """P = sum_a p_a lambda_a/2 per link (Fermions.cpp:342-360)."""
    return _dev.random_momenta(params.dims, _next_seed())


# ---------------------------------------------------------------- operator and solver
def field_inner_product_real(A, B):
    """Fermions.cpp:521-528."""
    return context().inner(A, B)


def apply_DDdagger(in_field, arr, out_field):
    """out = D(D^dagger in), written into the caller's array (Fermions.cpp:530-547)."""
    out_field[:] = _gauge(arr).apply_ddag(params.fermion_mass, in_field)


def solve_CG_DDdagger(phi_field, arr):
    """Fermions.cpp:553-617; raises RuntimeError on breakdown / non-convergence."""
    x, _ = _gauge(arr).cg(phi_field, params.fermion_mass, params.tol, params.max_iter, params.precision)
    return x


def solve_X_field(phi_field, arr):
    return solve_CG_DDdagger(phi_field, arr)


def solve_Y_field(X_field, arr):
    """Y = D^dagger X (Fermions.cpp:623-633)."""
    return _gauge(arr).apply_d(params.fermion_mass, X_field, dagger=True)


# ---------------------------------------------------------------- forces and actions
def compute_full_force_field(X_field, Y_field, arr, beta):
    """Fermions.cpp:640-659."""
    return _gauge(arr).force(beta, params.c_11, params.c_12, X_field, Y_field, which=3)


def compute_fermion_only_force_field(X_field, Y_field, arr):
    """Fermions.cpp:1114-1134."""
    return _gauge(arr).force(0.0, params.c_11, params.c_12, X_field, Y_field, which=2)


def compute_gauge_only_force_field(arr, beta):
    """Fermions.cpp:1278-1294."""
    return _gauge(arr).force(beta, params.c_11, params.c_12, which=1)


def compute_Tr_P2(P_field):
    """Fermions.cpp:662-678."""
    return _dev.Gauge(context(), params.dims).tr_p2(P_field)


def compute_action(arr, beta):
    """Lattice.cpp:759-781."""
    return _gauge(arr).action(beta, params.c_11, params.c_12)


def compute_avg_plaq(arr):
    """Lattice.cpp:784-804."""
    return _gauge(arr).avg_plaq()


_hb_sweep = [0]


def heatbath_update(arr, beta):
    """One Cabibbo-Marinari heatbath sweep of the Symanzik action in place
    (Lattice.cpp:725-757); returns (average plaquette, action density).  The
    links are visited in the device's colour order; sweeps take consecutive
    sweep indices of the Philox streams."""
    g = _gauge(arr)
    pl, ac = g.heatbath(beta, params.c_11, params.c_12, seed=_seed[0], sweep0=_hb_sweep[0], nsweeps=1)
    _hb_sweep[0] += 1
    arr[:] = g.download()
    return float(pl[0]), float(ac[0])


def _hex(arr, nsteps):
    g = _gauge(arr)
    g.hex_smear((params.alpha1, params.alpha2, params.alpha3), nsteps)
    return g


def Apply_HEX_smearing(U_array):
    """One HEX smearing step in place (Smearing.cpp:304-311)."""
    U_array[:] = _hex(U_array, 1).download()


def Apply_4HEX_smearing(U_array):
    """Four HEX smearing steps in place (Smearing.cpp:313-318)."""
    U_array[:] = _hex(U_array, 4).download()


def Apply_4HEX_smearing_with_history(U_thin):
    """The thin links and each of four HEX levels (Smearing.cpp:320-331)."""
    g = _gauge(U_thin)
    out = [np.array(U_thin, dtype=np.float64, copy=True)]
    for _ in range(4):
        g.hex_smear((params.alpha1, params.alpha2, params.alpha3), 1)
        out.append(g.download())
    return out


def numerically_stable_matrix_exponential(Q):
    """exp(iQ) of a hermitian traceless 3x3 (Fermions.cpp:843-902); raises
    ValueError where the reference throws std::invalid_argument."""
    Q = np.asarray(Q, dtype=np.complex128).reshape(3, 3)
    out = context().exp_iq(np.ascontiguousarray(Q).view(np.float64).reshape(-1))
    return out.view(np.complex128).reshape(3, 3)


# ---------------------------------------------------------------- integrator
def _md(beta, which=3, eps=None, nsteps=None):
    return dict(mass=params.fermion_mass, tol=params.tol, max_iter=params.max_iter, precision=params.precision,
                beta=beta, c11=params.c_11, c12=params.c_12,
                eps=params.epsilon if eps is None else eps,
                nsteps=params.number_of_steps_fermion if nsteps is None else nsteps, which=which)


def leapfrog_trajectory(P_field, beta, phi_field, arr, eps=None, n_steps=None, which=3, check=True):
    """The three leapfrog stages of Fermions.cpp:684-776 run back to back on the
    device: mutates P_field and arr in place like the reference and returns the
    final X field."""
    g = _gauge(arr)
    P, X = g.leapfrog(P_field, phi_field, check=check, **_md(beta, which, eps, n_steps))
    P_field[:] = P
    arr[:] = g.download()
    return X


def single_sweep_monte_carlo_update_WITH_FERMIONS(gauge_field, beta, chi=None, P_field=None, r_uniform=None):
    """One HMC update (Fermions.cpp:812-828).  gauge_field is replaced in place
    on accept.  Returns (H parts, accepted)."""
    chi = generate_chi_field() if chi is None else chi
    P = generate_conjugate_field_configuration() if P_field is None else P_field
    r = np.random.default_rng(_next_seed()).uniform() if r_uniform is None else r_uniform
    g = _gauge(gauge_field)
    H, acc = g.hmc_trajectory(chi, P, r, check=True, **_md(beta))
    if acc:
        gauge_field[:] = g.download()
    return H, acc
