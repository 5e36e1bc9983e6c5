"""Edge cases the reference's compiled-kernel tests pin
(tests/unit/test_native.py:96-101, 161-194, 236-255), on the device path:
nonfinite and scalar right-hand sides, non-contiguous blocks, extreme but
representable scales, CSR input with duplicate and unsorted entries, empty
rows, ragged row counts and the smallest systems."""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import hip_engine, host, oracle_engine, phases, random_block, rel_err, to_dev
from kubomsd_amd.synthetic import Model

pytestmark = pytest.mark.gpu


def metric(ctx, S):
    from kubomsd_amd.backends.hip import HipMetric
    S = sp.csr_matrix(S, dtype=np.complex128)
    return HipMetric(ctx, S, S.diagonal().real)


def test_metric_rejects_nonfinite_rhs(hip_ctx):
    """test_native.py:96-101: a NaN right-hand side is a ValueError naming 'finite'."""
    from kubomsd_amd.errors import NumericalError
    m = metric(hip_ctx, np.eye(2))
    with pytest.raises(NumericalError, match="finite"):
        m.solve(to_dev(hip_ctx, np.array([[1], [np.nan]], complex)), 1e-10, 0.0, 100)
    assert issubclass(NumericalError, ValueError)


def test_scalar_and_vector_rhs_rejected_without_device_calls(hip_ctx):
    """test_native.py:161-181: a scalar (or 1-D) block is shape-checked before
    any device access."""
    import torch
    from kubomsd_amd.errors import KuboMSDError
    m = metric(hip_ctx, np.eye(2))
    for bad in (torch.tensor(1.0 + 0j, dtype=torch.complex128, device=hip_ctx.torch_device),
                torch.ones(2, dtype=torch.complex128, device=hip_ctx.torch_device)):
        with pytest.raises(KuboMSDError):
            m.solve(bad, 1e-10, 0.0, 100)


def test_noncontiguous_rhs_and_true_residual_refresh(hip_ctx):
    """test_native.py:183-194: a strided right-hand side (every other column)
    solved to 1e-11 with at least one true-residual refresh."""
    n = 96
    S = sp.diags([-np.ones(n - 1), 2.03 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csr", dtype=complex)
    rng = np.random.default_rng(88)
    full = to_dev(hip_ctx, rng.normal(size=(n, 6)) + 1j * rng.normal(size=(n, 6)))
    B = full[:, ::2]
    assert not B.is_contiguous()
    m = metric(hip_ctx, S)
    out, rep = m.solve(B, 1e-11, 0.0, 1000)
    np.testing.assert_allclose(S @ host(out), host(B), rtol=1e-9, atol=1e-9)
    assert all(r <= 1e-11 for r in rep["relative_residual"])
    assert m.diagnostics()["true_residual_refreshes"] >= 1


def test_extreme_representable_rhs_is_preserved(hip_ctx):
    """test_native.py:236-246: 1.3e308 (1 + i) survives the per-column scaling."""
    m = metric(hip_ctx, np.eye(2))
    B = np.array([[complex(1.3e308, 1.3e308)], [1.0 + 0j]])
    out, rep = m.solve(to_dev(hip_ctx, B), 1e-10, 0.0, 100)
    out = host(out)
    assert np.isfinite(out).all()
    np.testing.assert_allclose(out[0].real / 1.3e308, [1.0], rtol=1e-14)
    np.testing.assert_allclose(out[0].imag / 1.3e308, [1.0], rtol=1e-14)
    np.testing.assert_allclose(out[1], [1.0], rtol=1e-14, atol=0)
    assert rep["converged"] == [True]


def test_duplicate_and_unsorted_csr_entries_are_summed(hip_ctx):
    """CSR with repeated (row, column) entries and unsorted columns: the
    product equals SciPy's, which sums duplicates (the reference sums them
    before its kernel: src/kubo_msd/backends/native.py:42-46)."""
    from kubomsd_amd.backends.hip import HipMatrix
    rng = np.random.default_rng(5)
    n = 50
    rows = rng.integers(0, n, 400)
    cols = rng.integers(0, n, 400)
    vals = rng.normal(size=400) + 1j * rng.normal(size=400)
    ptr = np.zeros(n + 1, np.int64)
    np.add.at(ptr, rows + 1, 1)
    ptr = np.cumsum(ptr)
    order = np.argsort(rows, kind="stable")  # rows grouped, columns left unsorted within a row
    A = sp.csr_matrix((vals[order], cols[order], ptr), shape=(n, n))
    assert not A.has_sorted_indices
    B = random_block(n, 5, 6)
    want = A.toarray() @ B
    for fmt in (None, 0):
        H = HipMatrix(hip_ctx, A) if fmt is None else HipMatrix(hip_ctx, A, fmt=fmt)
        assert rel_err(host(H.matmul(to_dev(hip_ctx, B))), want) < 1e-12
    # the C ABI itself, fed the raw arrays (duplicates and unsorted columns kept)
    import ctypes as C
    import torch
    from kubomsd_amd.backends import hip as HH
    lib = HH.library()
    p_ = np.ascontiguousarray(A.indptr, np.int64)
    i_ = np.ascontiguousarray(A.indices, np.int64)
    v_ = np.ascontiguousarray(A.data, np.complex128)
    for fmt in (0, 2):  # KMSD_CSR, KMSD_AUTO
        h = C.c_void_p()
        hip_ctx.check(lib.kmsd_matrix_create(hip_ctx.h, fmt, n, len(v_), p_.ctypes.data_as(C.c_void_p),
                                             i_.ctypes.data_as(C.c_void_p), v_.ctypes.data_as(C.c_void_p),
                                             C.byref(h)))
        try:
            Bd = to_dev(hip_ctx, B)
            out = torch.empty_like(Bd)
            hip_ctx.bind_stream()
            hip_ctx.check(lib.kmsd_matrix_matmul(h, C.c_void_p(Bd.data_ptr()), C.c_void_p(out.data_ptr()), 5,
                                                 HH.KMSD_C128))
            assert rel_err(host(out), want) < 1e-12
        finally:
            lib.kmsd_matrix_destroy(h)


def ragged_model(n, seed):
    """Hermitian tridiagonal-plus-long-range model with one empty row."""
    rng = np.random.default_rng(seed)
    H = sp.lil_matrix((n, n), dtype=complex)
    for i in range(n - 1):
        if i == n // 2 or i + 1 == n // 2:
            continue  # row n//2 stays empty (no diagonal either)
        v = rng.normal() + 1j * rng.normal()
        H[i, i + 1] = v
        H[i + 1, i] = np.conj(v)
        H[i, i] = rng.uniform(-1, 1)
    for i in range(0, n - 37, 7):
        if n // 2 not in (i, i + 37):
            H[i, i + 37] = .3
            H[i + 37, i] = .3
    H = sp.csr_matrix(H) * .15
    return Model(H, sp.identity(n, dtype=complex, format="csr"), {}, np.zeros((n, 3)))


@pytest.mark.parametrize("n", [1, 2, 1001, 4099])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_ragged_and_tiny_systems_match_oracle(hip_ctx, n, dtype):
    """Row counts that fill no whole row block, the 1x1 and 2x2 systems and an
    empty row: moments and chains equal the FP64 oracle."""
    if n == 1:
        m = Model(sp.csr_matrix(np.array([[0.3 + 0j]])), sp.identity(1, dtype=complex, format="csr"), {},
                  np.zeros((1, 3)))
    elif n == 2:
        m = Model(sp.csr_matrix(np.array([[0.1, 0.4j], [-0.4j, -0.2]])), sp.identity(2, dtype=complex, format="csr"),
                  {}, np.zeros((2, 3)))
    else:
        m = ragged_model(n, n)
    B = phases(m.n, 3, 7)
    want = oracle_engine(m).moments(B, 33, 0.0, 1.0)
    eng = hip_engine(hip_ctx, m)
    got = host(eng.moments(to_dev(hip_ctx, B, dtype), 33, 0.0, 1.0))
    tol = 1e-11 if dtype == "c128" else 5e-6
    assert rel_err(got, want) < tol
    chain = host(eng.chain(to_dev(hip_ctx, B, dtype), 5, 0.0, 1.0))
    assert rel_err(chain, oracle_engine(m).chain(B, 5, 0.0, 1.0)) < tol
