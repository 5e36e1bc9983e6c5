"""Pins the oracle's IC(0), cluster-Jacobi and direct metric modes (CPU only).

1. Bit for bit against the reference's own CPU kernel (oracle/_ref, built
   from cpp/native.cpp) when that build exists: ic0_factor values and error
   rows, IC(0)/cluster/direct solves with their reports and counters.
2. Against what the mathematics fixes: IC(0) on a pattern with no fill is the
   exact Cholesky factor (numpy); every mode solves S Y = B to the dense
   answer; the IC(0) preconditioner is Hermitian positive definite; the
   reference's own known answers (tests/unit/test_metric_ic0.py:9-29 matrix,
   the shift-ladder matrix of :32-44).
"""
import numpy as np
import pytest
import scipy.linalg as la
import scipy.sparse as sp

from helpers import random_block
from oracle import native as O

S4 = sp.csr_matrix([[3, .2j, 0, .3], [-.2j, 2, .4j, 0], [0, -.4j, 1.5, .1], [.3, 0, .1, 2.]])
LADDER = sp.csr_matrix(np.eye(5) + .17 * np.array([[0, -4, -2, 0, 0], [-4, 0, 0, 0, 3], [-2, 0, 0, 2, -4],
                                                   [0, 0, 2, 0, 0], [0, 3, -4, 0, 0.]]))


@pytest.fixture(scope="module")
def ref():
    mod = O.load_reference_native()
    if mod is None:
        pytest.skip("reference CPU kernel build (oracle/_ref) not present")
    return mod


def _rcsr(ref, m, threads=1):
    m = sp.csr_matrix(m, dtype=complex)
    m.sum_duplicates()
    m.sort_indices()
    return ref.CSR(m.indptr.astype(np.int64), m.indices.astype(np.int64), m.data, m.shape[0], threads)


def banded_spd(n, seed):
    rng = np.random.default_rng(seed)
    off1 = .2 * np.exp(1j * rng.uniform(-np.pi, np.pi, n - 1))
    off5 = .1 * np.exp(1j * rng.uniform(-np.pi, np.pi, n - 5))
    return sp.diags([off5.conj(), off1.conj(), 1.2 + .3 * rng.random(n), off1, off5], [-5, -1, 0, 1, 5],
                    format="csr")


def ic_lower(S, ordering):
    """Normalized, permuted lower pattern (the reference's setup,
    metric/preconditioners.py:68-77): W = D S D, P by RCM or natural."""
    z = 1 / np.sqrt(S.diagonal().real)
    W = (sp.diags(z) @ S @ sp.diags(z)).tocsr()
    W.eliminate_zeros()
    W.sort_indices()
    p = (sp.csgraph.reverse_cuthill_mckee(W, symmetric_mode=True) if ordering == "rcm"
         else np.arange(S.shape[0]))
    lower = sp.tril(W[p][:, p], format="csr")
    lower.sort_indices()
    return lower, p.astype(np.int64), z


def ic_factor_csr(lower, shift=0.0):
    return sp.csr_matrix((O.ic0_factor(lower, shift), lower.indices.copy(), lower.indptr.copy()), shape=lower.shape)


@pytest.mark.parametrize("ordering", ["natural", "rcm"])
@pytest.mark.parametrize("S", [S4, banded_spd(200, 1), banded_spd(257, 2)], ids=["4x4", "band200", "band257"])
def test_ic0_factor_and_solve_bit_identical_to_reference(ref, S, ordering):
    lower, p, z = ic_lower(S, ordering)
    for shift in (0.0, 1e-4):
        np.testing.assert_array_equal(O.ic0_factor(lower, shift), np.asarray(ref.ic0_factor(_rcsr(ref, lower), shift)))
    L = ic_factor_csr(lower)
    n = S.shape[0]
    B = random_block(n, 5, 3)
    B[:, 1] *= 1e-150
    B[:, 2] = 0
    om = O.Metric(O.Csr(S), S.diagonal().real, ())
    rm = ref.Metric(_rcsr(ref, S), S.diagonal().real.copy(), [], 1)
    om.set_ic0(L, p, z)
    rm.set_ic0(_rcsr(ref, L), p, z)
    a, ra = om.solve(B, 1e-12, 0., 1000)
    b, rb = rm.solve(B, 1e-12, 0., 1000)
    np.testing.assert_array_equal(a, b)
    assert ra["iterations"] == list(rb["iterations"]) and ra["converged"] == list(rb["converged"])
    np.testing.assert_array_equal(ra["relative_residual"], rb["relative_residual"])
    np.testing.assert_allclose(a, la.solve(S.toarray(), B), rtol=1e-10, atol=1e-160)


def test_ic0_failure_rows_match_reference(ref):
    lower, _, _ = ic_lower(LADDER, "natural")
    with pytest.raises(O.OracleError) as mine:
        O.ic0_factor(lower, 0.0)
    with pytest.raises(RuntimeError) as theirs:
        ref.ic0_factor(_rcsr(ref, lower), 0.0)
    assert str(mine.value) == str(theirs.value) and "pivot at row" in str(mine.value)
    np.testing.assert_array_equal(O.ic0_factor(lower, 1e-2), np.asarray(ref.ic0_factor(_rcsr(ref, lower), 1e-2)))
    for bad, msg in ((-1.0, "invalid IC shift"), (np.nan, "invalid IC shift")):
        with pytest.raises(O.OracleError, match=msg):
            O.ic0_factor(lower, bad)
    full = (lower + sp.triu(lower.conj().T, 1)).tocsr()  # entries right of the diagonal
    full.sort_indices()
    with pytest.raises(O.OracleError, match="diagonal at row end"):
        O.ic0_factor(full, 0.0)


@pytest.mark.parametrize("threads", [1, 3])
def test_cluster_blocks_bit_identical_to_reference(ref, threads):
    n = 90
    S = banded_spd(n, 7).toarray()
    rng = np.random.default_rng(5)
    perm = rng.permutation(n)
    sizes = [1, 7, 20, 2, 30, 30]  # noncontiguous clusters of mixed size
    cuts = np.cumsum([0] + sizes)
    blocks = []
    for a, b in zip(cuts[:-1], cuts[1:]):
        idx = np.sort(perm[a:b]) if a % 2 else perm[a:b]
        blocks.append((idx.astype(np.int64), la.cholesky(S[np.ix_(idx, idx)], lower=True)))
    B = random_block(n, 4, 9)
    om = O.Metric(O.Csr(S, threads), np.real(np.diag(S)), (), threads)
    rm = ref.Metric(_rcsr(ref, S, threads), np.real(np.diag(S)).copy(), [], threads)
    om.set_index_blocks(blocks)
    rm.set_index_blocks(blocks)
    a, ra = om.solve(B, 1e-13, 0., 1000)
    b, rb = rm.solve(B, 1e-13, 0., 1000)
    np.testing.assert_array_equal(a, b)
    assert ra["iterations"] == list(rb["iterations"])
    np.testing.assert_allclose(a, la.solve(S, B), rtol=1e-11)
    with pytest.raises(O.OracleError, match="incomplete cluster coverage"):
        om.set_index_blocks(blocks[:-1])
    with pytest.raises(O.OracleError, match="invalid cluster coverage"):
        om.set_index_blocks(blocks + [blocks[0]])


def test_direct_solve_and_refinement_bit_identical_to_reference(ref):
    rng = np.random.default_rng(7201)
    A = rng.normal(size=(48, 48)) + 1j * rng.normal(size=(48, 48))
    S = A @ A.conj().T + np.eye(48) * 20
    L = la.cholesky(S, lower=True)
    B = rng.normal(size=(48, 6)) + 1j * rng.normal(size=(48, 6))
    B[:, 0] = 0
    B[:, 1] *= 1e-100
    B[:, 2] *= 1e100
    for factor, steps in ((L, 3), (L * (1 + 1e-6), 3), (L * (1 + 1e-6), 0)):
        om = O.Metric(O.Csr(S), np.real(np.diag(S)), [(0, 48, factor)])
        rm = ref.Metric(_rcsr(ref, S), np.real(np.diag(S)).copy(), [(0, 48, factor)], 1)
        om.set_direct(True, steps)
        rm.set_direct(True, steps)
        res = []
        for m in (om, rm):
            try:
                res.append(m.solve(B, 1e-12, 0., 1000))
            except (O.OracleError, RuntimeError) as e:
                res.append((None, dict(e.diagnostics)))
        (a, ra), (b, rb) = res
        if a is not None or b is not None:
            np.testing.assert_array_equal(a, b)
        assert ra["iterations"] == [0] * 6 and list(rb["iterations"]) == [0] * 6
        np.testing.assert_array_equal(ra["relative_residual"], rb["relative_residual"])
        assert om.diagnostics()["refinement_steps"] == rm.diagnostics()["refinement_steps"]
        assert om.diagnostics()["true_residual_refreshes"] == rm.diagnostics()["true_residual_refreshes"]
        if steps:
            np.testing.assert_allclose(a / np.maximum(abs(B).max(0), 1), la.solve(S, B) / np.maximum(abs(B).max(0), 1),
                                       rtol=1e-12, atol=1e-12)
    assert om.diagnostics()["refinement_steps"] == 0
    with pytest.raises(O.OracleError, match="invalid direct metric factor"):
        O.Metric(O.Csr(S), np.real(np.diag(S)), ()).set_direct(True, 3)


def test_ic0_without_fill_is_the_exact_cholesky_factor():
    """A tridiagonal (or any banded-with-full-band) pattern has no fill, so
    IC(0) is the Cholesky factor of the permuted normalized matrix."""
    n = 60
    rng = np.random.default_rng(11)
    off = .3 * np.exp(1j * rng.uniform(-np.pi, np.pi, n - 1))
    S = sp.diags([off.conj(), 1.5 + rng.random(n), off], [-1, 0, 1], format="csr")
    lower, p, z = ic_lower(S, "natural")
    W = (np.diag(z) @ S.toarray() @ np.diag(z))[np.ix_(p, p)]
    L = ic_factor_csr(lower).toarray()
    np.testing.assert_allclose(L, np.linalg.cholesky(W), rtol=0, atol=1e-14)
    # a wrong conjugation in the merge would break this Hermitian case
    assert np.abs(np.imag(off)).max() > .1


def test_ic0_known_answer_of_reference_tests():
    """tests/unit/test_metric_ic0.py:9-29: the IC preconditioner is Hermitian
    positive definite and the solve reproduces the dense answer."""
    for ordering in ("natural", "rcm"):
        lower, p, z = ic_lower(S4, ordering)
        L = ic_factor_csr(lower)
        m = O.Metric(O.Csr(S4), S4.diagonal().real, ())
        m.set_ic0(L, p, z)
        Ld = L.toarray()
        Minv = np.diag(z)[:, p] @ la.inv(Ld @ Ld.conj().T) @ np.diag(z)[:, p].T
        np.testing.assert_allclose(Minv, Minv.conj().T, atol=1e-14)
        assert np.linalg.eigvalsh(Minv)[0] > 0
        B = np.array([[1j, 0], [2, 0], [3j, 0], [1, 0]])
        Y, r = m.solve(B, 1e-13, 0., 100)
        np.testing.assert_allclose(Y, np.linalg.solve(S4.toarray(), B), atol=1e-12)
        assert all(r["converged"])


def test_ic0_shift_ladder_case():
    """tests/unit/test_metric_ic0.py:32-44: no pivot at shift 0, success at 1e-2."""
    lower, _, _ = ic_lower(LADDER, "natural")
    with pytest.raises(O.OracleError, match="IC nonpositive"):
        O.ic0_factor(lower, 0.0)
    L = O.ic0_factor(lower, 1e-2)
    assert np.isfinite(L).all()
