"""kubomsd_amd.native, the drop-in for the reference's kubo_msd._native, used
the way the reference's adapter uses it (src/kubo_msd/backends/native.py:
42-133: CSR(indptr, indices, data, n, threads), Metric(S, diag, blocks,
threads), Engine(H, metric, definitions, rtol, atol, maxiter, hbar), NumPy in
and out), checked with the reference's own unit-test cases
(tests/unit/test_native.py) and against the FP64 oracle engine."""
import numpy as np
import pytest
import scipy.linalg as la
import scipy.sparse as sp

from helpers import dense_model
from oracle import native as O

pytestmark = pytest.mark.gpu


def csr(matrix, threads=1):
    """The reference adapter's csr() (backends/native.py:42-46) on the drop-in."""
    from kubomsd_amd import native
    host = sp.csr_matrix(matrix, dtype=np.complex128, copy=True)
    host.sum_duplicates()
    host.sort_indices()
    return native.CSR(host.indptr, host.indices, host.data, host.shape[0], threads)


def test_parallel_spmm_matches_scipy():
    n = 600
    off = np.linspace(.02, .04, n - 1) * (1 + 1j)
    S = sp.diags([off.conj(), 1.2 * np.ones(n), off], [-1, 0, 1], format="csr")
    B = np.random.default_rng(99).normal(size=(n, 8)) + 1j * np.random.default_rng(98).normal(size=(n, 8))
    np.testing.assert_allclose(csr(S, 2).matmul(B), S @ B, rtol=2e-15, atol=2e-15)


@pytest.mark.parametrize("blocks", [False, True], ids=["diagonal", "block-jacobi"])
def test_pcg_preserves_scaled_rhs_and_true_residual(blocks):
    from kubomsd_amd import native
    rng = np.random.default_rng(97)
    A = rng.normal(size=(11, 11)) + 1j * rng.normal(size=(11, 11))
    S = A.conj().T @ A + 2 * np.eye(11)
    B = (rng.normal(size=(11, 4)) + 1j * rng.normal(size=(11, 4))) * [0, 1e-180, 1, 1e180]
    fac = [(a, b, la.cholesky(S[a:b, a:b], lower=True)) for a, b in ((0, 3), (3, 7), (7, 11))] if blocks else []
    metric = native.Metric(csr(S), S.diagonal().real.copy(), fac, 1)
    out, report = metric.solve(B, 2e-13, 0.0, 1000)
    for j in (1, 2, 3):
        scale = np.max(np.abs(B[:, j]))
        np.testing.assert_allclose(out[:, j] / scale, la.solve(S, B[:, j] / scale), rtol=3e-11, atol=2e-12)
        assert la.norm(B[:, j] / scale - S @ (out[:, j] / scale)) / la.norm(B[:, j] / scale) <= 1.1 * 2e-13
    np.testing.assert_allclose(out[:, 0], 0, atol=0)
    assert all(report["converged"]) and max(report["relative_residual"]) <= 2e-13
    d = metric.diagnostics()
    assert d["calls"] == 1 and d["omp_threads_requested"] == 1


def test_failure_report_known_answer_and_rejections():
    from kubomsd_amd import native
    S = np.array([[1, .2j], [-.2j, 2]])
    metric = native.Metric(csr(S), S.diagonal().real.copy(), [], 1)
    with pytest.raises(native.MetricFailure) as failed:  # a RuntimeError, as the reference's
        metric.solve(np.array([[0, 1], [0, 1]], complex), 1e-12, 0.0, 1)
    assert isinstance(failed.value, RuntimeError)
    diag = failed.value.diagnostics
    assert diag["converged"] == [True, False] and diag["iterations"] == [0, 1]
    # one Jacobi-PCG iteration from p = (1, .5) (tests/unit/test_native.py:218-232)
    S2 = np.array([[1., .2], [.2, 2.]])
    out, rep = native.Metric(csr(S2), S2.diagonal().copy(), [], 1).solve(np.ones((2, 1)), .5, 0.0, 1)
    np.testing.assert_allclose(out[:, 0], [15 / 17, 15 / 34], rtol=1e-14, atol=1e-14)
    assert rep["converged"] == [True] and rep["iterations"] == [1]
    with pytest.raises(ValueError, match="finite"):
        metric.solve(np.array([[1], [np.nan]]), 1e-12, 0.0, 10)
    with pytest.raises(ValueError):
        native.CSR(np.array([0, 1]), np.array([3]), np.array([1 + 0j]), 1, 1)


@pytest.mark.parametrize("kind", ["velocity", "position", "composite"])
def test_engine_matches_oracle(kind):
    from kubomsd_amd import native
    m = dense_model(9, 903)
    H, S, X = m.H.toarray(), m.S.toarray(), m.velocity["x"].toarray()
    r = np.linspace(0, 2, 9)
    mats = {"velocity": (0, [X]), "position": (1, [X]),
            "composite": (2, [H * (r[None, :] - r[:, None]), X - r[:, None] * S, X - S * r[None, :]])}[kind]
    defs_n = [(mats[0], [csr(a) for a in mats[1]]), (-1, []), (-1, [])]
    defs_o = [(mats[0], [O.Csr(sp.csr_matrix(a, dtype=complex)) for a in mats[1]]), (-1, []), (-1, [])]
    eng = native.Engine(csr(H), native.Metric(csr(S), S.diagonal().real.copy(), [], 1), defs_n, 1e-13, 0.0, 1000,
                        O.HBAR_EV_FS)
    ref = O.Engine(O.Csr(sp.csr_matrix(H)), O.Metric(O.Csr(sp.csr_matrix(S)), S.diagonal().real, (), 1), defs_o,
                   1e-13, 0.0, 1000)
    rng = np.random.default_rng(18)
    B = rng.normal(size=(9, 3)) + 1j * rng.normal(size=(9, 3))
    d = [1.0, 0.0, 0.0]
    np.testing.assert_allclose(eng.v_cov(B, d), ref.v_cov(B, d), rtol=2e-11, atol=2e-12)
    np.testing.assert_allclose(eng.h_action(B), ref.h_action(B), rtol=2e-11, atol=2e-12)
    np.testing.assert_allclose(eng.chain(B, 12, .1, 2.0), ref.chain(B, 12, .1, 2.0), rtol=2e-10, atol=2e-11)
    np.testing.assert_allclose(eng.moments(B, 20, .1, 2.0), ref.moments(B, 20, .1, 2.0), rtol=2e-10, atol=2e-11)
    coeff = np.exp(-np.arange(8)) * (1 + .3j)
    Baug = np.hstack([B, np.zeros_like(B)])
    np.testing.assert_allclose(eng.series(Baug, coeff, .1, 2.0, [d]), ref.series(Baug, coeff, .1, 2.0, [d]),
                               rtol=2e-10, atol=2e-11)
    assert (eng.h_calls, eng.v_calls) == (ref.h_calls, ref.v_calls)
