"""Chebyshev inner solves of the complex64 metric refinement (tuning key
"cheb_inner", on by default).  Where S is stored as CSR/ELL with the
diagonal or no preconditioner and its Gershgorin interval [lo, hi] has
lo > 0 and hi <= 4 lo, each FP32 inner solve is the Chebyshev semi-iteration
on that interval, one fused step per product with no reductions; the FP64
refinement accepts on the FP64 true residual exactly as with PCG inner
solves.  Checked against SciPy's solve of the complex64-rounded right-hand
side (one complex64 rounding of the output), against the PCG inner solves
(cheb_inner = 0), with real and complex Hermitian S, and through the
engine's moments of S^-1 H against the FP64 oracle.  A matrix whose
interval is not positive keeps the PCG inner solves, bit for bit."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from helpers import host, random_block, rel_err, to_dev
from kubomsd_amd.synthetic import honeycomb_flake
from oracle import native as O

pytestmark = pytest.mark.gpu


def flake_S(n=700, seed=11, complex_values=False):
    """S = I + 0.08 U on the bonds of a honeycomb flake (Gershgorin interval
    within [0.76, 1.24]); complex_values: Hermitian with complex bond values."""
    m = honeycomb_flake(n, s_scale=0.08, seed=seed)
    S = sp.csr_matrix(m.S)
    if complex_values:
        up = sp.triu(S, k=1).tocoo()
        rng = np.random.default_rng(seed)
        ph = np.exp(1j * rng.uniform(0, 2 * np.pi, up.nnz))
        U = sp.coo_matrix((up.data * ph, (up.row, up.col)), shape=S.shape)
        S = (sp.identity(S.shape[0], format="csr") + U + U.conj().T).tocsr()
    S.sort_indices()
    return m, S


def solve(ctx, S, B, flag):
    from kubomsd_amd.backends.hip import HipMetric
    ctx.set_tuning("cheb_inner", flag)
    try:
        met = HipMetric(ctx, S, S.diagonal().real)
        Y, rep = met.solve(to_dev(ctx, B, "c64"), 1e-10, 0.0, 1000)
        return host(Y), rep
    finally:
        ctx.set_tuning("cheb_inner", 1)


@pytest.mark.parametrize("complex_values", [False, True])
def test_cheb_inner_solves_reach_fp64_residual(hip_ctx, complex_values):
    _, S = flake_S(complex_values=complex_values)
    B = random_block(S.shape[0], 16, 7)
    Bq = B.astype(np.complex64).astype(np.complex128)
    ref = spla.spsolve(S.tocsc(), Bq)
    out = {}
    for flag in (1, 0):
        Y, rep = solve(hip_ctx, S, B, flag)
        assert all(rep["converged"]) and max(rep["relative_residual"]) <= 1e-10
        res = np.linalg.norm(Bq - S @ Y, axis=0) / np.linalg.norm(Bq, axis=0)
        assert res.max() < 1e-6  # the solution is rounded once to complex64
        assert rel_err(Y, ref) < 2e-7
        out[flag] = (Y, rep)
    # the series runs the same number of terms on every column; the two
    # inner solvers round differently
    assert len(set(out[1][1]["iterations"])) == 1 and out[1][1]["iterations"][0] > 0
    assert not np.array_equal(out[1][0], out[0][0])
    assert rel_err(out[1][0], out[0][0]) < 2e-7


def test_pcg_kept_where_the_interval_is_not_positive(hip_ctx):
    """A well-conditioned SPD matrix that is not diagonally dominant: its
    Gershgorin interval reaches below zero, so the inner solves stay PCG and
    the two settings give the same bits."""
    rng = np.random.default_rng(5)
    n = 300
    A = sp.random(n, n, density=0.03, random_state=7, format="csr")
    S = (A @ A.T + 0.5 * sp.identity(n)).astype(np.complex128).tocsr()
    S.sort_indices()
    d = S.diagonal().real
    off = np.asarray(abs(S).sum(axis=1)).ravel() - np.abs(S.diagonal())
    assert (d - off).min() <= 0  # not usable for the series
    B = rng.normal(size=(n, 8)) + 1j * rng.normal(size=(n, 8))
    Y1, rep1 = solve(hip_ctx, S, B, 1)
    Y0, rep0 = solve(hip_ctx, S, B, 0)
    assert all(rep1["converged"])
    np.testing.assert_array_equal(Y1, Y0)
    assert list(rep1["iterations"]) == list(rep0["iterations"])


def test_flake_moments_with_cheb_inner_match_oracle(hip_ctx):
    """Moments of S^-1 H (complex64 blocks, each order one H product and one
    refined S solve with Chebyshev inner solves) against the FP64 oracle
    engine, as the complex64 engine tests bound them."""
    from kubomsd_amd.backends.hip import HipEngine, HipMetric
    m, S = flake_S(n=400, seed=3)
    H = sp.csr_matrix(m.H)
    rt = 2e-13
    oe = O.Engine(O.Csr(H.toarray()), O.Metric(O.Csr(S.toarray()), S.diagonal().real, ()), [(-1, [])] * 3, rt,
                  0.0, 1000)
    he = HipEngine(hip_ctx, H, HipMetric(hip_ctx, S, S.diagonal().real), [(-1, [])] * 3, rt, 0.0, 1000)
    B = random_block(H.shape[0], 4, 13)
    got = host(he.moments(to_dev(hip_ctx, B, "c64"), 24, 0.0, 12.0))
    want = oe.moments(B, 24, 0.0, 12.0)
    assert rel_err(got, want) < 2e-4


@pytest.mark.parametrize("eligible", [True, False])
def test_mixed_complex128_solves_take_the_refinement(hip_ctx, eligible):
    """kmsd_metric_set_mixed: complex128 right-hand sides solved by FP32
    inner solves (Chebyshev where S allows it, else PCG) inside the FP64
    refinement reach the same FP64 residual as the FP64 PCG, and the two
    solutions agree to the solve tolerance."""
    from kubomsd_amd.backends.hip import HipMetric
    if eligible:
        _, S = flake_S(n=500, seed=17)
    else:
        n = 300
        A = sp.random(n, n, density=0.03, random_state=9, format="csr")
        S = (A @ A.T + 0.5 * sp.identity(n)).astype(np.complex128).tocsr()
        S.sort_indices()
    B = random_block(S.shape[0], 8, 23)
    out = {}
    for mixed in (True, False):
        met = HipMetric(hip_ctx, S, S.diagonal().real)
        met.set_mixed(mixed)
        Y, rep = met.solve(to_dev(hip_ctx, B, "c128"), 1e-10, 0.0, 1000)
        Y = host(Y)
        assert Y.dtype == np.complex128
        assert all(rep["converged"]) and max(rep["relative_residual"]) <= 1e-10
        res = np.linalg.norm(B - S @ Y, axis=0) / np.linalg.norm(B, axis=0)
        assert res.max() <= 2e-10  # the device's FP64 residual, recomputed on the host
        out[mixed] = (Y, rep)
    assert rel_err(out[True][0], out[False][0]) < 1e-9
    if eligible:
        assert len(set(out[True][1]["iterations"])) == 1


def test_block_jacobi_series_after_ritz_bounds(hip_ctx):
    """4x4-block S with atom-block Jacobi (the simple-cubic orbital model):
    the first complex64 inner solve runs PCG and records its CG coefficients;
    the Ritz values of that Lanczos tridiagonal, widened, give the interval of
    the block-Jacobi Chebyshev semi-iteration that later inner solves use
    (one matrix-core product with the block solve and the three-term update
    per step).  Every solve reaches the FP64 residual and SciPy's solution;
    the semi-iteration solves run the same steps on every column;
    with cheb_inner = 0 all solves stay PCG and agree to the rounding."""
    import scipy.linalg as la
    from kubomsd_amd.backends.hip import HipMetric
    from kubomsd_amd.synthetic import cubic_orbital_model
    model = cubic_orbital_model(6, seed=4411)
    S = sp.csr_matrix(model.S)
    blocks = [(a, b, la.cholesky(model.S[a:b, a:b].toarray(), lower=True)) for a, b in model.atom_slices]
    Bs = [random_block(model.n, 16, 40 + i) for i in range(3)]
    outs = {}
    for flag in (1, 0):
        hip_ctx.set_tuning("cheb_inner", flag)
        try:
            met = HipMetric(hip_ctx, S, S.diagonal().real, blocks)
            outs[flag] = [met.solve(to_dev(hip_ctx, B, "c64"), 1e-10, 0.0, 1000) for B in Bs]
        finally:
            hip_ctx.set_tuning("cheb_inner", 1)
    for flag in (1, 0):
        for B, (Y, rep) in zip(Bs, outs[flag]):
            Y = host(Y)
            Bq = B.astype(np.complex64).astype(np.complex128)
            assert all(rep["converged"]) and max(rep["relative_residual"]) <= 1e-10
            assert rel_err(Y, spla.spsolve(S.tocsc(), Bq)) < 2e-7
    later = outs[1][2][1]
    assert len(set(later["iterations"])) == 1 and later["iterations"][0] > 0
    # the first solve's first refinement round is PCG in both settings (its
    # coefficients give the bounds); later rounds differ by the inner solver
    assert not np.array_equal(host(outs[1][2][0]), host(outs[0][2][0]))
    for (Y1, _), (Y0, _) in zip(outs[1], outs[0]):
        assert rel_err(host(Y1), host(Y0)) < 2e-7


def test_block_series_that_misses_the_spectrum_returns_to_pcg(hip_ctx):
    """bcheb_lo_pct puts the block-Jacobi iteration interval's lower end at 75 % of the
    largest eigenvalue of M^-1 S (dense generalized eigenvalues on the host),
    so a series round leaves the low eigencomponents nearly untouched: the
    refinement still reaches the FP64 residual and SciPy's solution, that
    round (the first solve's second) makes the metric return to PCG, and the
    later solves are the PCG solves of cheb_inner = 0 bit for bit."""
    import scipy.linalg as la
    from kubomsd_amd.backends.hip import HipMetric
    from kubomsd_amd.synthetic import cubic_orbital_model
    model = cubic_orbital_model(6, seed=4411)
    S = sp.csr_matrix(model.S)
    Sd = model.S.toarray()
    Md = np.zeros_like(Sd)
    for a, b in model.atom_slices:
        Md[a:b, a:b] = Sd[a:b, a:b]
    w = la.eigh(Sd, Md, eigvals_only=True)
    assert w.max() / w.min() > 3
    pct = int(100 * 0.75 * w.max() / w.min())
    blocks = [(a, b, la.cholesky(model.S[a:b, a:b].toarray(), lower=True)) for a, b in model.atom_slices]
    Bs = [random_block(model.n, 16, 60 + i) for i in range(3)]
    outs = {}
    for flag in (1, 0):
        hip_ctx.set_tuning("cheb_inner", flag)
        hip_ctx.set_tuning("bcheb_lo_pct", pct)
        try:
            met = HipMetric(hip_ctx, S, S.diagonal().real, blocks)
            outs[flag] = [met.solve(to_dev(hip_ctx, B, "c64"), 1e-10, 0.0, 1000) for B in Bs]
        finally:
            hip_ctx.set_tuning("cheb_inner", 1)
            hip_ctx.set_tuning("bcheb_lo_pct", 96)
    for flag in (1, 0):
        for B, (Y, rep) in zip(Bs, outs[flag]):
            Bq = B.astype(np.complex64).astype(np.complex128)
            assert all(rep["converged"]) and max(rep["relative_residual"]) <= 1e-10
            assert rel_err(host(Y), spla.spsolve(S.tocsc(), Bq)) < 2e-7
    # first solve: a PCG round (its coefficients give the interval), one
    # series round, then PCG; the later solves are PCG only
    assert not np.array_equal(host(outs[1][0][0]), host(outs[0][0][0]))
    for i in (1, 2):
        np.testing.assert_array_equal(host(outs[1][i][0]), host(outs[0][i][0]))
        assert list(outs[1][i][1]["iterations"]) == list(outs[0][i][1]["iterations"])


def test_unrepresentable_complex64_solution_then_ordinary_solves(hip_ctx):
    """S = I / 4 (a one-point Gershgorin interval: the inner solves are
    exact divisions).  A complex64 right-hand side of 3e38 has a solution of
    1.2e39, which complex64 cannot hold: the solve raises.  The output's
    finiteness flag is checked in the conversion pass, alternating between
    two flags that each launch clears for the next: the two following
    solves on the same context, the second reusing the flag the failing
    one set, return 4 B."""
    from kubomsd_amd.backends.hip import HipMetric
    from kubomsd_amd.errors import NumericalError
    n = 64
    S = (0.25 * sp.identity(n, format="csr")).astype(np.complex128).tocsr()
    met = HipMetric(hip_ctx, S, S.diagonal().real)
    bad = np.ones((n, 2), np.complex64)
    bad[0, 0] = 3e38
    with pytest.raises(NumericalError, match="representable"):
        met.solve(to_dev(hip_ctx, bad, "c64"), 1e-10, 0.0, 100)
    for seed in (1, 2):
        B = random_block(n, 2, seed)
        Y, rep = met.solve(to_dev(hip_ctx, B, "c64"), 1e-10, 0.0, 100)
        assert all(rep["converged"])
        np.testing.assert_allclose(host(Y), 4 * B.astype(np.complex64), rtol=1e-6)


@pytest.mark.parametrize("rtol,atol", [(1e-6, 0.0), (1e-10, 1e-9)])
def test_point_series_rounds_sized_one_by_one(hip_ctx, rtol, atol):
    """Solves outside the two-round lookahead (rtol >= 1e-7, or atol > 0):
    each refinement round is sized from the previous residual, read on the
    host.  The FP64 residual meets max(rtol ||b||, atol) per column and the
    solution agrees with SciPy's to that accuracy."""
    from kubomsd_amd.backends.hip import HipMetric
    _, S = flake_S(n=600, seed=29)
    B = random_block(S.shape[0], 8, 31)
    met = HipMetric(hip_ctx, S, S.diagonal().real)
    Y, rep = met.solve(to_dev(hip_ctx, B, "c64"), rtol, atol, 1000)
    Y = host(Y)
    Bq = B.astype(np.complex64).astype(np.complex128)
    assert all(rep["converged"])
    thr = np.maximum(rtol * np.linalg.norm(Bq, axis=0), atol)
    res = np.linalg.norm(Bq - S @ Y.astype(np.complex128), axis=0)
    assert (res <= np.maximum(thr, 2e-7 * np.linalg.norm(Bq, axis=0))).all()  # plus one complex64 rounding
    assert len(set(rep["iterations"])) == 1 and rep["iterations"][0] > 0
    assert rel_err(Y, spla.spsolve(S.tocsc(), Bq)) < max(10 * rtol, 2e-7)
