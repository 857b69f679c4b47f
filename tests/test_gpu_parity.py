"""Parity of the HIP path (through the C ABI) with the FP64 CPU oracle on the
same seeded inputs, plus size-independent properties at the benchmark sizes.

Tolerances: the FP64 kernels agree with the oracle to rounding (1e-13
relative elementwise for D, forces, exp); the CG solution to the solve
tolerance; mixed precision to its FP64 true residual."""
import numpy as np
import pytest

from oracle import oracle as oc
from staggered_fermions_hmc_amd import device as dev
import fields as F

pytestmark = pytest.mark.gpu

DIMS_SET = [(4, 4, 4, 4), (4, 4, 4, 8), (6, 4, 4, 2), (2, 2, 2, 4), (8, 6, 4, 4), (12, 4, 6, 8)]


@pytest.fixture(scope="module")
def ctx():
    c = dev.Context(0)
    yield c
    c.close()


def rel(a, b):
    return np.abs(a - b).max() / max(1e-300, np.abs(b).max())


# ---------------------------------------------------------------- layout
@pytest.mark.parametrize("dims", DIMS_SET)
def test_layout_round_trip_exact(ctx, dims):
    U = dev.random_links(dims, 0.5, seed=1)
    g = dev.Gauge(ctx, dims, U)
    assert np.array_equal(g.download(), U)
    s = dev.random_spinor(dims, seed=2)
    f = dev.Field(g, s)
    assert np.array_equal(f.download(), s)


# ---------------------------------------------------------------- Dirac operator
@pytest.mark.parametrize("dims", DIMS_SET)
@pytest.mark.parametrize("dagger", [False, True])
def test_dslash_matches_oracle(ctx, dims, dagger):
    m = 0.13
    U = dev.random_links(dims, 0.4, seed=3)
    psi = dev.random_spinor(dims, seed=4)
    g = dev.Gauge(ctx, dims, U)
    out = g.apply_d(m, psi, dagger=dagger)
    ref = oc.apply_d(dims, m, U, psi, dagger=dagger)
    assert rel(out, ref) < 1e-14


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (8, 8, 8, 8)])
def test_ddag_matches_oracle(ctx, dims):
    U = dev.random_links(dims, 0.3, seed=5)
    psi = dev.random_spinor(dims, seed=6)
    g = dev.Gauge(ctx, dims, U)
    assert rel(g.apply_ddag(0.05, psi), oc.apply_ddag(dims, 0.05, U, psi)) < 1e-14


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (6, 4, 4, 2), (8, 6, 4, 4), (16, 16, 16, 16)])
@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
def test_hop_forms_agree(ctx, dims, prec):
    """All hop forms (one lane per site with four, one or two directions in
    flight; four lanes per site) give the same operator and CG solves."""
    m = 0.1
    U = dev.random_links(dims, 0.3, seed=41)
    phi = dev.random_spinor(dims, seed=42)
    g = dev.Gauge(ctx, dims, U)
    sol, op = [], []
    try:
        for v in (0, 1, 2, 3):
            dev.set_hop_variant(prec, v)
            sol.append(g.cg(phi, m, tol=1e-10, precision=prec)[0])
            if prec == "f64":
                op.append(g.apply_d(m, phi))
    finally:
        dev.set_hop_variant(prec, -1)
    for k in (1, 2, 3):
        if prec == "f64":
            assert rel(op[k], op[0]) < 1e-14, k  # D: the forms differ at most in summation order
            assert rel(sol[k], sol[0]) < 1e-9, k
        else:
            assert rel(sol[k], sol[0]) < 1e-8, k


def _stored_links(U, dtype, two_rows):
    """The low-precision link copy as the library stores and reads it: each
    real rounded to nearest from double to float or half (the compiled
    conversion to half rounds once, from double), and with SU(3) links (the
    two-row read) the third row of every link rebuilt as conj(row0 x row1)
    from the stored rows."""
    q = U.astype(dtype).astype(np.float64)
    if two_rows:
        M = (q[0::2] + 1j * q[1::2]).reshape(-1, 3, 3)
        M[:, 2, :] = np.conj(np.cross(M[:, 0, :], M[:, 1, :]))
        q = np.empty_like(U)
        q[0::2], q[1::2] = M.real.reshape(-1), M.imag.reshape(-1)
    return q


def _stored_f16_spinor(psi):
    """FP16 records as the library writes them: FP32 components, the site's
    max-norm n (FP32), halves of v * (1 / n) rounded to nearest; returns
    n * h per component."""
    v = psi.astype(np.float32).reshape(-1, 6)
    n = np.abs(v).max(axis=1)
    inv = np.where(n > 0, np.float32(1) / np.where(n > 0, n, np.float32(1)), np.float32(0)).astype(np.float32)
    h = (v * inv[:, None]).astype(np.float16)
    return (n[:, None].astype(np.float64) * h.astype(np.float64)).reshape(-1)


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (8, 6, 4, 4), (16, 16, 16, 16)])
@pytest.mark.parametrize("su3", [True, False])
def test_low_precision_hop_matches_oracle(ctx, dims, su3):
    """D and D^dag in the inner iteration's storage (the kernels the mixed
    CG runs: FP16 products from the stored halves, FP32 arithmetic).
    Against the FP64 oracle on the exact inputs: FP32 within 1e-6 and FP16
    within 4e-3 of max|D psi| (2^-11 relative rounding of every stored link
    element and spinor component, a few dozen terms per site).  Against the
    FP64 oracle applied to the stored (quantised) links and spinor, FP16 to
    within half an FP16 unit of the output's own storage (2^-12 of each
    site's norm) plus FP32 accumulation (measured: 2.443e-4 at most, 16^4),
    and FP32 to 5e-7 (measured 1.9e-7): this pins the arithmetic on the
    stored values, norms, signs and rebuilt rows, not only the precision.
    su3 False: links scaled by 0.9, which take the three-row reads."""
    m = 0.1
    U = dev.random_links(dims, 0.3, seed=61)
    if not su3:
        U = 0.9 * U
    psi = dev.random_spinor(dims, seed=62)
    g = dev.Gauge(ctx, dims, U)
    for dg in (False, True):
        exact = oc.apply_d(dims, m, U, psi, dagger=dg)
        d32 = g.apply_d(m, psi, dagger=dg, precision="f32")
        assert rel(d32, exact) < 1e-6, dg
        d16 = g.apply_d(m, psi, dagger=dg, precision="f16")
        assert rel(d16, exact) < 4e-3, dg
        emul = oc.apply_d(dims, m, _stored_links(U, np.float16, su3), _stored_f16_spinor(psi), dagger=dg)
        e = emul.reshape(-1, 6)
        site_norm = np.abs(e).max(axis=1)
        diff = np.abs(d16.reshape(-1, 6) - e).max(axis=1)
        assert np.all(diff <= (2.0 ** -12 + 2e-5) * site_norm), (dg, (diff / site_norm).max())
        emul32 = oc.apply_d(dims, m, _stored_links(U, np.float32, su3), psi.astype(np.float32).astype(np.float64),
                            dagger=dg)
        assert rel(d32, emul32) < 5e-7, dg


def test_dslash_delta_source_stencil(ctx):
    """The oracle's hand-derived stencil cases, on the device (to 4e-15: the
    FP64 hop rebuilds the third row of SU(3) links from the first two, which
    moves each entry by up to a few units of rounding)."""
    dims = (4, 4, 4, 4)
    U = F.random_links(dims, 0.6, seed=5)
    g = dev.Gauge(ctx, dims, U)
    for n0 in [(0, 0, 0, 3), (1, 1, 1, 0)]:
        psi = np.zeros((256, 3), complex)
        psi[F.site_index(dims, np.array(n0)), 1] = 0.3 - 0.8j
        flat = F.vec_to_flat(psi)
        for dg in (False, True):
            assert rel(g.apply_d(0.37, flat, dagger=dg), oc.apply_d(dims, 0.37, U, flat, dagger=dg)) < 4e-15


def test_free_field_closed_form_large(ctx):
    """U = 1 at 16^3 x 32: D D^dag of an antiperiodic plane wave is
    (m^2 + sum sin^2 p) times it."""
    dims = (16, 16, 16, 32)
    m = 0.01
    k = (3, 5, 1, 7)
    p = np.array([2 * np.pi * k[i] / dims[i] for i in range(3)] + [(2 * k[3] + 1) * np.pi / dims[3]])
    c = F.coords(dims)
    psi = np.exp(1j * (c @ p))[:, None] * np.array([0.6, -0.2j, 0.1 + 0.3j])[None, :]
    g = dev.Gauge(ctx, dims, F.cold_links(dims))
    out = F.flat_to_vec(g.apply_ddag(m, F.vec_to_flat(psi)))
    lam = m * m + np.sum(np.sin(p) ** 2)
    assert np.abs(out - lam * psi).max() < 1e-13


def test_hop_antihermitian_at_32_4(ctx):
    dims = (32, 32, 32, 32)
    U = dev.random_links(dims, 0.3, seed=9)
    g = dev.Gauge(ctx, dims, U)
    a = dev.random_spinor(dims, seed=10)
    b = dev.random_spinor(dims, seed=11)
    Ka = g.apply_d(0.0, a)
    Kb = g.apply_d(0.0, b)
    ca, cb = F.flat_to_vec(a), F.flat_to_vec(b)
    lhs = np.vdot(ca, F.flat_to_vec(Kb))
    rhs = -np.vdot(F.flat_to_vec(Ka), cb)
    assert abs(lhs - rhs) < 1e-10 * np.sqrt(a @ a * (b @ b))
    # D^dag = eps D eps
    es = np.repeat(F.eps_sign(dims), 6)
    assert rel(g.apply_d(0.01, a, dagger=True), es * g.apply_d(0.01, es * a)) < 1e-14


# ---------------------------------------------------------------- inner product
def test_inner_product(ctx):
    a = np.random.default_rng(1).standard_normal(100003)
    b = np.random.default_rng(2).standard_normal(100003)
    assert abs(ctx.inner(a, b) - oc.inner(a, b)) < 1e-10


# ---------------------------------------------------------------- CG
@pytest.mark.parametrize("dims,m", [((4, 4, 4, 4), 0.1), ((8, 8, 8, 8), 0.05), ((6, 4, 4, 2), 0.2)])
def test_cg_f64_matches_oracle(ctx, dims, m):
    U = dev.random_links(dims, 0.3, seed=12)
    chi = dev.random_spinor(dims, seed=13)
    phi = oc.apply_d(dims, m, U, chi)
    g = dev.Gauge(ctx, dims, U)
    x, st = g.cg(phi, m, tol=1e-10, precision="f64", true_residual=True)
    xr, it, rsq, _ = oc.cg(dims, m, U, phi, tol=1e-10)
    assert abs(st["iterations"] - it) <= 1
    assert rel(x, xr) < 1e-9
    r = phi - oc.apply_ddag(dims, m, U, x)
    assert abs(st["true_rel_resid"] - np.sqrt(r @ r / (phi @ phi))) < 1e-12
    # pseudofermion identity Y = D^dag X = chi
    assert rel(g.apply_d(m, x, dagger=True), chi) < 1e-8
    # stopped on the FP64 true residual: at most a restart past the
    # reference recurrence, and the true residual meets the tolerance
    xs, ss = g.cg(phi, m, tol=1e-10, precision="f64", true_residual="stop")
    r = phi - oc.apply_ddag(dims, m, U, xs)
    assert np.sqrt(r @ r / (phi @ phi)) < 1e-10 and ss["true_rel_resid"] < 1e-10
    assert it - 1 <= ss["iterations"] <= it + 32


@pytest.fixture(params=[0, 1], ids=["fused", "split"])
def cg_form(request):
    """Both forms of the mixed-precision iteration, restored to the default."""
    for prec in ("f32", "f16"):
        dev.set_cg_form(prec, request.param)
    yield request.param
    for prec in ("f32", "f16"):
        dev.set_cg_form(prec, -1)


@pytest.mark.parametrize("prec", ["f32", "f16"])
@pytest.mark.parametrize("dims,m", [((4, 4, 4, 4), 0.1), ((8, 8, 8, 8), 0.05), ((6, 4, 8, 10), 0.07)])
def test_cg_mixed_reaches_true_residual(ctx, dims, m, prec, cg_form):
    U = dev.random_links(dims, 0.3, seed=14)
    phi = oc.apply_d(dims, m, U, dev.random_spinor(dims, seed=15))
    g = dev.Gauge(ctx, dims, U)
    x, st = g.cg(phi, m, tol=1e-10, precision=prec)
    r = phi - oc.apply_ddag(dims, m, U, x)
    true = np.sqrt(r @ r / (phi @ phi))
    assert true < 1e-10, (true, st)
    assert abs(st["true_rel_resid"] - true) < 1e-11
    xr, *_ = oc.cg(dims, m, U, phi, tol=1e-12)
    assert rel(x, xr) < 1e-8
    assert st["reliable_updates"] >= 1


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_cg_forms_agree(ctx, prec):
    """The fused and split mixed iterations run the same recurrence: equal
    iteration and update counts, solutions within the solve tolerance."""
    dims, m = (8, 8, 8, 12), 0.03
    U = dev.random_links(dims, 0.3, seed=61)
    g = dev.Gauge(ctx, dims, U)
    phi = dev.random_spinor(dims, seed=62)
    out = {}
    try:
        for form in (0, 1):
            dev.set_cg_form(prec, form)
            out[form] = g.cg(phi, m, tol=1e-10, precision=prec)
    finally:
        dev.set_cg_form(prec, -1)
    (x0, s0), (x1, s1) = out[0], out[1]
    assert abs(s0["iterations"] - s1["iterations"]) <= 2, (s0, s1)
    assert s0["true_rel_resid"] < 1e-10 and s1["true_rel_resid"] < 1e-10
    assert rel(x0, x1) < 1e-8


def test_cg_error_paths(ctx):
    dims = (4, 4, 4, 4)
    U = dev.random_links(dims, 0.3, seed=16)
    g = dev.Gauge(ctx, dims, U)
    phi = dev.random_spinor(dims, seed=17)
    x, st = g.cg(np.zeros_like(phi), 0.1)
    assert st["iterations"] == 0 and not x.any()
    with pytest.raises(RuntimeError, match="failed to converge after 3 iterations"):
        g.cg(phi, 0.1, max_iter=3)
    bad = phi.copy()
    bad[7] = np.nan
    with pytest.raises(RuntimeError, match="CG breakdown"):
        g.cg(bad, 0.1)
    # the context stays usable after an error
    x, st = g.cg(phi, 0.1, tol=1e-10)
    assert st["iterations"] > 0
    # a tolerance below the FP64 rounding floor of the true residual (about
    # 1e-16 relative): mixed precision runs its iteration budget through
    # repeated reliable updates and reports non-convergence with the
    # reference's message (not a stalled state).  The source is scaled so
    # that <p, v> stays far above the breakdown threshold of 1e-30.
    big = 1e6 * phi
    for prec in ("f32", "f16"):
        with pytest.raises(RuntimeError, match="failed to converge after 600 iterations"):
            g.cg(big, 0.1, tol=1e-17, max_iter=600, precision=prec)
    # FP64 stopped on the true residual: meets it or reports non-convergence
    try:
        x, st = g.cg(big, 0.1, tol=1e-17, max_iter=600, precision="f64", true_residual="stop")
        assert st["true_rel_resid"] < 1e-17
    except RuntimeError as e:
        assert "failed to converge after 600 iterations" in str(e)


def test_cg_16_4_both_precisions(ctx):
    """The 16^4, m=0.05 solve: FP64 and FP32-inner reach a true 1e-10."""
    dims = (16, 16, 16, 16)
    m = 0.05
    U = dev.random_links(dims, 0.3, seed=21)
    g = dev.Gauge(ctx, dims, U)
    phi = dev.random_spinor(dims, seed=22)
    x64, s64 = g.cg(phi, m, tol=1e-10, precision="f64", true_residual="stop")
    x32, s32 = g.cg(phi, m, tol=1e-10, precision="f32")
    x16, s16 = g.cg(phi, m, tol=1e-10, precision="f16")
    r = phi - oc.apply_ddag(dims, m, U, x64)
    assert np.sqrt(r @ r / (phi @ phi)) < 1e-10
    assert s64["true_rel_resid"] < 1e-10
    assert s32["true_rel_resid"] < 1e-10
    assert s16["true_rel_resid"] < 1e-10
    assert rel(x32, x64) < 1e-8 and rel(x16, x64) < 1e-8


@pytest.mark.parametrize("dims", [(24, 24, 24, 24), (12, 12, 12, 20)])
def test_mixed_solves_on_non_power_of_two_lattices(ctx, dims):
    """Mixed-precision solves where V is not a power of two: 24^4 (331776
    sites, above the fused form's 2^18, so the split form with group sums)
    and 12^3 x 20 (34560 sites, the fused form with per-workgroup partials),
    in both storage precisions at their default update thresholds and chunk
    lengths.  The true residual is the oracle's FP64 D D^dag of the returned
    solution, not the solver's own report."""
    m = 0.02
    U = dev.random_links(dims, 0.3, seed=61)
    g = dev.Gauge(ctx, dims, U)
    phi = dev.random_spinor(dims, seed=62)
    for prec in ("f32", "f16"):
        x, st = g.cg(phi, m, tol=1e-10, precision=prec)
        r = phi - oc.apply_ddag(dims, m, U, x)
        true_res = np.sqrt(r @ r / (phi @ phi))
        assert true_res < 1e-10, (prec, st)
        assert abs(st["true_rel_resid"] - true_res) < 1e-12, (prec, st, true_res)
        assert st["reliable_updates"] >= 1, (prec, st)


# ---------------------------------------------------------------- forces / actions
@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (6, 4, 4, 2), (4, 6, 8, 4)])
@pytest.mark.parametrize("c11,c12", [(1.0, 0.0), (5.0 / 3.0, -1.0 / 12.0), (0.0, 1.0)])
def test_gauge_force_and_action_match_oracle(ctx, dims, c11, c12):
    U = dev.random_links(dims, 0.5, seed=23)
    g = dev.Gauge(ctx, dims, U)
    Fg = g.force(4.0, c11, c12, which=1)
    Fr = oc.force(dims, U, None, None, 4.0, c11, c12, which=1)
    assert np.abs(Fg - Fr).max() < 1e-13 * max(1.0, np.abs(Fr).max())
    S = g.action(4.0, c11, c12)
    assert abs(S - oc.action(dims, U, 4.0, c11, c12)) < 1e-11 * max(1.0, abs(S))
    assert abs(g.avg_plaq() - oc.avg_plaq(dims, U)) < 1e-13


@pytest.mark.parametrize("scale", [0.9, 1.0 + 1e-12])
def test_gauge_force_non_su3_links_match_oracle(ctx, scale):
    """Links off SU(3) (scaled by 0.9, or by 1 + 1e-12, above the 1e-14
    two-row threshold of the FP64 kernels) take the force's full-row path,
    which makes no use of unitarity, and still match the oracle."""
    dims = (4, 6, 4, 4)
    U = scale * dev.random_links(dims, 0.5, seed=27)
    g = dev.Gauge(ctx, dims, U)
    for c11, c12 in ((5.0 / 3.0, -1.0 / 12.0), (1.0, 0.0)):
        Fg = g.force(4.0, c11, c12, which=1)
        Fr = oc.force(dims, U, None, None, 4.0, c11, c12, which=1)
        assert np.abs(Fg - Fr).max() < 1e-13 * max(1.0, np.abs(Fr).max())


def test_staple_kernels_refuse_extents_above_their_packing(ctx):
    """The staple walks pack each coordinate into 8 bits: the force, action,
    plaquette and HEX smearing refuse an extent of 256 and the heatbath one
    above 252 (SFH_E_INVALID), instead of walking wrong sites; the hop and
    the solver take such lattices."""
    dims = (256, 2, 2, 2)
    U = dev.random_links(dims, 0.3, seed=28)
    g = dev.Gauge(ctx, dims, U)
    for call in (lambda: g.force(4.0, 5 / 3, -1 / 12, which=1), lambda: g.action(4.0, 5 / 3, -1 / 12),
                 lambda: g.avg_plaq(), lambda: g.hex_smear((0.95, 0.76, 0.38), 1)):
        with pytest.raises(dev.SfhError):
            call()
    b = dev.random_spinor(dims, seed=29)
    x, st = g.cg(b, 0.5, tol=1e-10, precision="f64")
    r = b - oc.apply_ddag(dims, 0.5, U, x)
    assert np.sqrt(r @ r / (b @ b)) < 1e-10
    dims = (256, 4, 4, 4)
    g = dev.Gauge(ctx, dims, dev.random_links(dims, 0.3, seed=30))
    with pytest.raises(dev.SfhError):
        g.heatbath(4.1, 5 / 3, -1 / 12, seed=1, sweep0=0, nsweeps=1)


def test_fermion_and_full_force_match_oracle(ctx):
    dims = (4, 4, 4, 8)
    m = 0.1
    U = dev.random_links(dims, 0.4, seed=24)
    X = dev.random_spinor(dims, seed=25)
    Y = dev.random_spinor(dims, seed=26)
    g = dev.Gauge(ctx, dims, U)
    for which in (2, 3):
        Fd = g.force(5.6, 5 / 3, -1 / 12, X, Y, which=which)
        Fr = oc.force(dims, U, X, Y, 5.6, 5 / 3, -1 / 12, which=which)
        assert np.abs(Fd - Fr).max() < 1e-13 * max(1.0, np.abs(Fr).max())


def test_cold_known_answers_on_device(ctx):
    dims = (4, 4, 4, 4)
    g = dev.Gauge(ctx, dims, F.cold_links(dims))
    V = 256
    assert abs(g.action(4.1, 5 / 3, -1 / 12) - (-4.1 * (6 * V * 5 / 3 + 12 * V * (-1 / 12)))) < 1e-10
    assert abs(g.avg_plaq() - 6.0) < 1e-13
    assert np.abs(g.force(4.1, 5 / 3, -1 / 12, which=1)).max() < 1e-14


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 6, 8, 6), (16, 16, 16, 32)])
def test_polyakov_grid_matches_oracle(ctx, dims):
    U = dev.random_links(dims, 0.4, seed=44)
    g = dev.Gauge(ctx, dims, U)
    assert np.abs(g.polyakov_grid() - oc.polyakov_grid(dims, U)).max() < 1e-12


def test_checkpoint_byte_format(ctx, tmp_path):
    """sfh_gauge_save writes the reference's format (Lattice.cpp:810-828):
    the AoS link arrays back to back as raw doubles; load restores them."""
    dims = (4, 4, 4, 6)
    U1 = dev.random_links(dims, 0.3, seed=45)
    U2 = dev.random_links(dims, 0.5, seed=46)
    g1, g2 = dev.Gauge(ctx, dims, U1), dev.Gauge(ctx, dims, U2)
    path = tmp_path / "final_lattice_configs.bin"
    dev.save_links(ctx, [g1, g2], path)
    raw = np.fromfile(path, dtype="<f8")
    assert np.array_equal(raw, np.concatenate([U1, U2]))
    h1, h2 = dev.Gauge(ctx, dims), dev.Gauge(ctx, dims)
    dev.load_links(ctx, [h1, h2], path)
    assert np.array_equal(h1.download(), U1) and np.array_equal(h2.download(), U2)
    with pytest.raises(ValueError):
        dev.load_links(ctx, [h1, h2, dev.Gauge(ctx, dims)], path)


def test_tr_p2(ctx):
    dims = (4, 4, 4, 6)
    P = dev.random_momenta(dims, seed=27)
    g = dev.Gauge(ctx, dims)
    assert abs(g.tr_p2(P) - oc.tr_p2(dims, P)) < 1e-10


# ---------------------------------------------------------------- exp and link update
def test_exp_matches_oracle(ctx):
    rng = np.random.default_rng(28)
    for scale in (1e-7, 1e-3, 0.05, 0.7, 3.0):
        for _ in range(5):
            Q = F.mat_to_flat(scale * F.random_algebra(rng, ()))
            assert np.abs(ctx.exp_iq(Q) - oc.exp_iq(Q)).max() < 1e-14
    for Q in [np.diag([0.4, 0.4, -0.8]), np.zeros((3, 3)), np.diag([-0.4, -0.4, 0.8])]:
        Qf = F.mat_to_flat(Q.astype(complex))
        assert np.abs(ctx.exp_iq(Qf) - oc.exp_iq(Qf)).max() < 1e-15
    with pytest.raises(ValueError, match="imaginary trace"):
        ctx.exp_iq(F.mat_to_flat(np.diag([0.5j, 0.1, -0.1])))


def test_link_update_matches_oracle(ctx):
    dims = (4, 4, 4, 4)
    U = dev.random_links(dims, 0.3, seed=29)
    P = dev.random_momenta(dims, seed=30)
    g = dev.Gauge(ctx, dims, U)
    g.link_update(0.1, P)
    assert np.abs(g.download() - oc.link_update(dims, 0.1, P, U)).max() < 1e-14
    bad = P.copy()
    bad[0] += 1.0
    with pytest.raises(RuntimeError, match="not Hermitian"):
        g.link_update(0.1, bad)


def test_tune_hop_picks_a_form(ctx):
    dims = (8, 8, 8, 8)
    g = dev.Gauge(ctx, dims, dev.random_links(dims, 0.3, seed=43))
    try:
        for prec in ("f64", "f32", "f16"):
            v, us = g.tune_hop(prec, n=5)
            assert v in (0, 1, 2, 3) and us > 0
    finally:
        for prec in ("f64", "f32", "f16"):
            dev.set_hop_variant(prec, -1)


def test_tune_cg_form_and_kernel_times(ctx):
    dims = (8, 8, 8, 8)
    g = dev.Gauge(ctx, dims, dev.random_links(dims, 0.3, seed=44))
    try:
        for prec in ("f32", "f16"):
            ua, ub = g.time_cg_kernels(prec, n=5)
            assert ua > 0 and ub > 0
            f, us = g.tune_cg_form(prec, n=5)
            assert f in (0, 1) and us > 0
    finally:
        for prec in ("f32", "f16"):
            dev.set_cg_form(prec, -1)
    with pytest.raises(dev.SfhError):
        g.time_cg_kernels("f64", n=5)


def test_dropin_selftest_binary():
    """The C++ drop-in layer passes the reference-style checks."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "staggered_fermions_hmc_amd",
                       "compat", "dropin_selftest")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


# ---------------------------------------------------------------- integrator
@pytest.mark.parametrize("which", [1, 2, 3])
def test_leapfrog_matches_oracle(ctx, which):
    dims = (4, 4, 4, 4)
    m = 0.2
    U = dev.random_links(dims, 0.3, seed=31)
    P = dev.random_momenta(dims, seed=32)
    phi = oc.apply_d(dims, m, U, dev.random_spinor(dims, seed=33))
    kw = dict(beta=4.0, c11=5 / 3, c12=-1 / 12, eps=0.05, nsteps=4, which=which, tol=1e-12)
    Ur, Pr, Xr = oc.leapfrog(dims, m, U, P, phi, **kw)
    g = dev.Gauge(ctx, dims, U)
    Pd, Xd = g.leapfrog(P, phi, mass=m, **kw)
    assert np.abs(g.download() - Ur).max() < 1e-11
    assert np.abs(Pd - Pr).max() < 1e-9 * max(1.0, np.abs(Pr).max())
    if which & 2:
        assert rel(Xd, Xr) < 1e-8


def test_leapfrog_reversible_on_device(ctx):
    dims = (8, 8, 8, 8)
    m = 0.1
    U0 = dev.random_links(dims, 0.3, seed=34)
    P0 = dev.random_momenta(dims, seed=35)
    g = dev.Gauge(ctx, dims, U0)
    phi = g.apply_d(m, dev.random_spinor(dims, seed=36))
    kw = dict(mass=m, beta=5.6, c11=1.0, c12=0.0, eps=0.05, nsteps=5, which=3, tol=1e-13)
    P1, _ = g.leapfrog(P0, phi, **kw)
    P2, _ = g.leapfrog(-P1, phi, **kw)
    assert np.abs(g.download() - U0).max() < 1e-10
    assert np.abs(-P2 - P0).max() < 1e-9


def test_hmc_trajectory_matches_oracle(ctx):
    dims = (4, 4, 4, 4)
    m = 0.1
    U = dev.random_links(dims, 0.3, seed=37)
    P = dev.random_momenta(dims, seed=38)
    chi = dev.random_spinor(dims, seed=39)
    kw = dict(beta=5.6, c11=1.0, c12=0.0, eps=0.1, nsteps=10, tol=1e-12)
    Ur, Hr, accr = oc.hmc_trajectory(dims, m, U, chi, P, 0.3, **kw)
    g = dev.Gauge(ctx, dims, U)
    H, acc = g.hmc_trajectory(chi, P, 0.3, mass=m, **kw)
    assert np.abs(H[:6] - Hr[:6]).max() < 1e-8 * np.abs(Hr[:6]).max()
    assert abs(H[6] - Hr[6]) < 1e-7
    assert acc == accr
    assert np.abs(g.download() - Ur).max() < 1e-10


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_mixed_restart_after_convergence_pause(ctx, prec):
    """Near the end of a solve the iterated residual can meet the target
    while the true one has not: the solve pauses, and the FP64 update
    restarts it.  In the fused form the pause can fall at either buffer
    parity (the restart moves p_k into the buffer the next chunk starts
    from); the split form keeps one buffer.  Both run the same recurrence, so
    over several right-hand sides the iteration counts agree and every solve
    reaches the true residual."""
    dims, m = (8, 8, 8, 12), 0.03
    U = dev.random_links(dims, 0.3, seed=71)
    g = dev.Gauge(ctx, dims, U)
    try:
        for seed in range(72, 80):
            phi = dev.random_spinor(dims, seed=seed)
            out = {}
            for form in (0, 1):
                dev.set_cg_form(prec, form)
                out[form] = g.cg(phi, m, tol=1e-10, precision=prec)
            (x0, s0), (x1, s1) = out[0], out[1]
            assert abs(s0["iterations"] - s1["iterations"]) <= 2, (seed, s0, s1)
            assert abs(s0["reliable_updates"] - s1["reliable_updates"]) <= 1, (seed, s0, s1)
            for x in (x0, x1):
                r = phi - oc.apply_ddag(dims, m, U, x)
                assert np.sqrt(r @ r / (phi @ phi)) < 1e-10
    finally:
        dev.set_cg_form(prec, -1)


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_mixed_without_due_updates_restarts_fresh(ctx, prec):
    """A reliable-update factor too small to ever fall due leaves the sloppy
    residual to drift until the iterated residual meets the target; the FP64
    update then finds the true residual far above it and restarts the
    iteration from p = r_true (keeping the old direction, scaled by
    |r_true|^2 / |r_k|^2, would stall).  Both forms converge alike."""
    dims, m = (8, 8, 8, 12), 0.03
    U = dev.random_links(dims, 0.3, seed=81)
    g = dev.Gauge(ctx, dims, U)
    phi = dev.random_spinor(dims, seed=82)
    _, s64 = g.cg(phi, m, tol=1e-10, precision="f64")
    try:
        out = {}
        for form in (0, 1):
            dev.set_cg_form(prec, form)
            out[form] = g.cg(phi, m, tol=1e-10, precision=prec, reliable_delta=1e-30, max_iter=20000)
    finally:
        dev.set_cg_form(prec, -1)
    (x0, s0), (x1, s1) = out[0], out[1]
    assert abs(s0["iterations"] - s1["iterations"]) <= 2, (s0, s1)
    assert s0["reliable_updates"] >= 2 and s0["iterations"] < 6 * s64["iterations"], (s0, s64)
    for x in (x0, x1):
        r = phi - oc.apply_ddag(dims, m, U, x)
        assert np.sqrt(r @ r / (phi @ phi)) < 1e-10


@pytest.mark.parametrize("scale,noise", [(1.0, 0.0), (1.0, 1e-12), (0.9, 0.0)], ids=["su3", "near_su3", "scaled"])
def test_fp64_hop_and_cg_two_row_threshold(ctx, scale, noise):
    """The FP64 hop reads two link rows only when the third matches
    conj(row0 x row1) to 1e-14: exact SU(3) links take that path, links off
    SU(3) by 1e-12 (two-row for the FP32/FP16 copies) and links scaled by 0.9
    read all three rows.  Every case matches the oracle to FP64 rounding;
    rebuilding the third row of the 1e-12 case would miss it by 1e-12."""
    dims, m = (8, 8, 8, 8), 0.05
    U = scale * dev.random_links(dims, 0.3, seed=91)
    U = U + noise * np.random.default_rng(92).standard_normal(U.shape)
    g = dev.Gauge(ctx, dims, U)
    psi = dev.random_spinor(dims, seed=93)
    for dagger in (False, True):
        assert rel(g.apply_d(m, psi, dagger=dagger), oc.apply_d(dims, m, U, psi, dagger=dagger)) < 1e-14
    assert rel(g.apply_ddag(m, psi), oc.apply_ddag(dims, m, U, psi)) < 1e-14
    x, st = g.cg(psi, m, tol=1e-12, precision="f64")
    xr, *_ = oc.cg(dims, m, U, psi, tol=1e-12)
    assert rel(x, xr) < 1e-10


@pytest.mark.parametrize("precision", ["f32", "f16"])
def test_mixed_cg_non_su3_links(ctx, precision):
    """The low-precision hops read two link rows and rebuild the third only
    for SU(3) links; links scaled by 0.9 take the three-row kernels and the
    mixed solve still reaches the true residual and the oracle's solution.
    The same lattice's SU(3) links then take the two-row kernels again."""
    dims = (8, 8, 8, 8)
    U = dev.random_links(dims, 0.3, seed=51)
    phi = dev.random_spinor(dims, seed=52)
    g = dev.Gauge(ctx, dims, 0.9 * U)
    for links in (0.9 * U, U):
        g.upload(links)
        x, st = g.cg(phi, 0.05, tol=1e-10, precision=precision)
        xr, *_ = oc.cg(dims, 0.05, links, phi, tol=1e-12)
        r = phi - oc.apply_ddag(dims, 0.05, links, x)
        assert np.sqrt(r @ r / (phi @ phi)) < 1e-10, st
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-8


# ---------------------------------------------------------------- Polyakov correlators
@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (6, 6, 6, 4), (8, 4, 6, 4), (32, 32, 32, 32)])
def test_polyakov_correlators_match_oracle(ctx, dims):
    """Device grid and correlators (Lattice.cpp:868-948) for r^2 = 1 ..
    3 (L/2)^2 + 1 against the oracle's serial sums: equal to rounding of the
    summation order, relative to the scale of the summands."""
    U = dev.random_links(dims, 0.6, seed=91)
    g = dev.Gauge(ctx, dims, U)
    L = np.asarray(dims[:3])
    r2_max = int(((L // 2) ** 2).sum()) + 1 if dims[0] < 32 else 15  # Main.cpp: 1 .. Spatial_Size/2 - 1
    got = g.polyakov_correlators(r2_max)
    P = oc.polyakov_grid(dims, U)
    scale = np.abs(P).max() ** 2
    for r2 in range(1, r2_max + 1):
        ref = oc.polyakov_correlator(dims[:3], P, r2)
        assert abs(got[r2 - 1] - ref) < 1e-13 * scale, (r2, got[r2 - 1], ref)
    # from a host grid, one r^2 at a time (the drop-in's call pattern)
    for r2 in (1, 2, r2_max):
        c = dev.polyakov_correlators_grid(ctx, dims[:3], P, r2, r2_first=r2)
        assert abs(c[0] - oc.polyakov_correlator(dims[:3], P, r2)) < 1e-13 * scale
