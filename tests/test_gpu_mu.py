"""Double-moment tables table[b,j,q] = sum_i conj(L[j,i,b]) R[q,i,b] against an
FP64 einsum (the reference's own contraction, src/kubo_msd/solvers/spectral.py:37).
The FP64 tables (FP64 matrix cores) cover several row and column tiles, batch
widths that are and are not multiples of four, and N ranges split over
workgroups; the FP32 and bf16x3 square tables take one batch column."""
import numpy as np
import pytest

from helpers import host, to_dev

pytestmark = pytest.mark.gpu


def blocks(m, k, n, nb, seed):
    rng = np.random.default_rng(seed)
    L = rng.normal(size=(m, n, nb)) + 1j * rng.normal(size=(m, n, nb))
    R = rng.normal(size=(k, n, nb)) + 1j * rng.normal(size=(k, n, nb))
    return L, R


def dev3(ctx, a, dtype):
    return to_dev(ctx, a.reshape(a.shape[0], -1), dtype).reshape(a.shape)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("shape", [(37, 11, 500, 3), (5, 70, 129, 16), (16, 1, 33, 1), (300, 43, 2000, 64),
                                   (33, 130, 17, 8), (70, 17, 5000, 6)])
def test_fp64_table_matches_einsum(hip_ctx, dtype, shape):
    from kubomsd_amd.backends.hip import KMSD_MU_FP64, mu_table
    m, k, n, nb = shape
    L, R = blocks(m, k, n, nb, 1)
    if dtype == "c64":
        L = L.astype(np.complex64).astype(np.complex128)
        R = R.astype(np.complex64).astype(np.complex128)
    want = np.einsum("jnb,knb->bjk", L.conj(), R)
    got = host(mu_table(hip_ctx, dev3(hip_ctx, L, dtype), dev3(hip_ctx, R, dtype), KMSD_MU_FP64))
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    # accumulate mode adds into an existing table
    again = host(mu_table(hip_ctx, dev3(hip_ctx, L, dtype), dev3(hip_ctx, R, dtype), KMSD_MU_FP64,
                          table=to_dev(hip_ctx, got.reshape(nb, -1)).reshape(nb, m, k), accumulate=True))
    assert np.max(np.abs(again - 2 * want)) <= 2e-12 * np.max(np.abs(want))


@pytest.mark.parametrize("mode,tol", [("fp32", 2e-6), ("bf16x3", 1e-4)])
@pytest.mark.parametrize("shape", [(100, 70, 3000), (64, 64, 64), (129, 31, 1000)])
def test_mfma_table_matches_einsum(hip_ctx, mode, tol, shape):
    from kubomsd_amd.backends.hip import KMSD_MU_BF16X3, KMSD_MU_FP32, mu_table
    m, k, n = shape
    L, R = blocks(m, k, n, 1, 2)
    L = L.astype(np.complex64).astype(np.complex128)
    R = R.astype(np.complex64).astype(np.complex128)
    want = np.einsum("jnb,knb->bjk", L.conj(), R)
    code = KMSD_MU_FP32 if mode == "fp32" else KMSD_MU_BF16X3
    got = host(mu_table(hip_ctx, dev3(hip_ctx, L, "c64"), dev3(hip_ctx, R, "c64"), code))
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    assert err <= tol, err
