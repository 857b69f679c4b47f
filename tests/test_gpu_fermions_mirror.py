"""The reference-named calls of tests/fermions_mirror.py against
the oracle on the reference's default 4^4 lattice: each function returns what
the reference function of the same name returns (Fermions.h:31-70, :105,
:130; Lattice.h:75-77), through the HIP library."""
import numpy as np
import pytest

from oracle import oracle as oc
from staggered_fermions_hmc_amd import device as dev
import fermions_mirror as fm

pytestmark = pytest.mark.gpu

DIMS = (4, 4, 4, 4)
BETA = 4.1


@pytest.fixture(scope="module", autouse=True)
def configured():
    fm.configure(Spatial_Size=4, temporal_size=4, fermion_mass=0.1, tol=1e-12, epsilon=0.05,
                 number_of_steps_fermion=4, precision="f64")
    yield


def rel(a, b):
    return np.abs(a - b).max() / max(1e-300, np.abs(b).max())


def test_operator_solve_and_y_field():
    U = dev.random_links(DIMS, 0.5, seed=1)
    chi = dev.random_spinor(DIMS, seed=2)
    out = np.empty_like(chi)
    fm.apply_DDdagger(chi, U, out)
    assert rel(out, oc.apply_ddag(DIMS, 0.1, U, chi)) < 1e-14
    phi = fm.generate_phi_field(U, chi)
    assert rel(phi, oc.apply_d(DIMS, 0.1, U, chi)) < 1e-14
    X = fm.solve_X_field(phi, U)
    r = phi - oc.apply_ddag(DIMS, 0.1, U, X)
    assert np.sqrt(r @ r / (phi @ phi)) < 1e-11
    Y = fm.solve_Y_field(X, U)
    assert rel(Y, oc.apply_d(DIMS, 0.1, U, X, dagger=True)) < 1e-14
    assert abs(fm.field_inner_product_real(phi, X) - oc.inner(phi, X)) < 1e-12 * abs(oc.inner(phi, X))


def test_forces_actions_and_observables():
    U = dev.random_links(DIMS, 0.5, seed=3)
    X = dev.random_spinor(DIMS, seed=4)
    Y = dev.random_spinor(DIMS, seed=5)
    c11, c12 = fm.params.c_11, fm.params.c_12
    assert rel(fm.compute_full_force_field(X, Y, U, BETA), oc.force(DIMS, U, X, Y, BETA, c11, c12, 3)) < 1e-13
    assert rel(fm.compute_fermion_only_force_field(X, Y, U), oc.force(DIMS, U, X, Y, 0.0, c11, c12, 2)) < 1e-13
    assert rel(fm.compute_gauge_only_force_field(U, BETA), oc.force(DIMS, U, None, None, BETA, c11, c12, 1)) < 1e-13
    P = dev.random_momenta(DIMS, seed=6)
    assert abs(fm.compute_Tr_P2(P) - oc.tr_p2(DIMS, P)) < 1e-12 * abs(oc.tr_p2(DIMS, P))
    S = oc.action(DIMS, U, BETA, c11, c12)
    assert abs(fm.compute_action(U, BETA) - S) < 1e-12 * abs(S)
    assert abs(fm.compute_avg_plaq(U) - oc.avg_plaq(DIMS, U)) < 1e-13


def test_exponential_and_its_guard():
    rng = np.random.default_rng(7)
    H = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    Q = 0.3 * (H + H.conj().T)
    Q -= np.trace(Q) / 3 * np.eye(3)
    E = fm.numerically_stable_matrix_exponential(Q)
    Eo = oc.exp_iq(np.ascontiguousarray(Q).view(np.float64).reshape(-1)).view(np.complex128).reshape(3, 3)
    assert np.abs(E - Eo).max() < 1e-14
    with pytest.raises(ValueError):
        fm.numerically_stable_matrix_exponential(1j * np.diag([1.0, 2.0, -3.0]))


def test_leapfrog_and_hmc_update_match_oracle():
    U = dev.random_links(DIMS, 0.3, seed=8)
    chi = dev.random_spinor(DIMS, seed=9)
    P = dev.random_momenta(DIMS, seed=10)
    phi = oc.apply_d(DIMS, 0.1, U, chi)
    W = U.copy()
    Pm = P.copy()
    X = fm.leapfrog_trajectory(Pm, BETA, phi, W)
    Uo, Po, Xo = oc.leapfrog(DIMS, 0.1, U, P, phi, beta=BETA, c11=fm.params.c_11, c12=fm.params.c_12, eps=0.05,
                             nsteps=4, tol=1e-12)
    assert rel(W, Uo) < 1e-12 and rel(Pm, Po) < 1e-10 and rel(X, Xo) < 1e-9
    G = U.copy()
    H, acc = fm.single_sweep_monte_carlo_update_WITH_FERMIONS(G, BETA, chi=chi, P_field=P, r_uniform=0.0)
    Uh, Ho, acco = oc.hmc_trajectory(DIMS, 0.1, U, chi, P, 0.0, beta=BETA, c11=fm.params.c_11, c12=fm.params.c_12,
                                     eps=0.05, nsteps=4, tol=1e-12)
    assert acc == acco
    assert abs(H[6] - Ho[6]) < 1e-8 * max(1.0, abs(Ho[6]))
    if acc:
        assert rel(G, Uh) < 1e-12


def test_hex_smearing_entry_points():
    U = dev.random_links(DIMS, 0.5, seed=11)
    alpha = (fm.params.alpha1, fm.params.alpha2, fm.params.alpha3)
    W = U.copy()
    fm.Apply_HEX_smearing(W)
    assert np.abs(W - oc.hex_smear(DIMS, U, alpha=alpha, nsteps=1)).max() < 1e-12
    W = U.copy()
    fm.Apply_4HEX_smearing(W)
    assert np.abs(W - oc.hex_smear(DIMS, U, alpha=alpha, nsteps=4)).max() < 1e-12
    hist = fm.Apply_4HEX_smearing_with_history(U)
    assert len(hist) == 5 and np.array_equal(hist[0], U) and np.array_equal(hist[4], W)


def test_heatbath_update_in_place():
    U = np.zeros(int(np.prod(DIMS)) * 72)
    U[0::18] = U[8::18] = U[16::18] = 1.0
    first = fm.heatbath_update(U, BETA)
    assert isinstance(first, tuple) and len(first) == 2
    V1 = U.copy()
    second = fm.heatbath_update(U, BETA)
    assert not np.array_equal(U, V1) and first != second
    M = (U[0::2] + 1j * U[1::2]).reshape(-1, 3, 3)
    assert np.abs(np.matmul(np.conj(np.swapaxes(M, 1, 2)), M) - np.eye(3)).max() < 1e-13
    assert 0 < second[0] < 1 and second[1] < 0
