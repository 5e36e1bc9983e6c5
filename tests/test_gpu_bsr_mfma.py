"""Matrix-core 4x4-block products (bsr4_mfma_kernel): FP32 matrix values on
complex64 blocks, block products on v_mfma_f32_4x4x1_16b with FP32 row sums,
the kernel of the metric products inside complex64 inner PCG solves.

The library runs it for every FP32-valued 4x4-block product of a complex64
block whose width is a multiple of 4; the context flag "fp32_math" routes
plain products (kmsd_matrix_matmul) there too, which is how these tests reach
it directly.  Each product is compared with SciPy's FP64 product of the same
FP32-rounded values and inputs (tolerance 2e-6 of the largest entry: one
complex64 rounding of the output plus FP32 sums of at most 4 x 45 products)
and with the VALU kernel that "fp32_math" selects when "bsr_mfma" = 0: that
kernel also sums in FP32, block by block and column by column within a
block, and a 4x4x1 matrix-core step is one fused multiply-add per output, so
the two agree bit for bit.  In the metric solves (no "fp32_math") the VALU
kernel sums in FP64, so there the two differ in rounding only.

Cases: block rows with 0 to 45 blocks (LDS chunks of 10, 20 and 40 blocks for
4, 8 and 16 lanes per block row, so long rows take several chunks and empty
rows none), widths 4..96 (one or two column chunks, idle lanes past the last
column), real and imaginary-only stored values, and the simple-cubic orbital
model's S."""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import host, random_block, rel_err, to_dev
from kubomsd_amd.synthetic import cubic_orbital_model

pytestmark = pytest.mark.gpu


def ragged_bsr(nq=96, seed=17):
    """Real 4x4-block matrix whose block rows hold 0..45 blocks (row q % 46)."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for q in range(nq):
        k = q % 46
        c = np.sort(rng.choice(nq, size=min(k, nq), replace=False))
        rows += [q] * len(c)
        cols += list(c)
    blocks = rng.uniform(-1, 1, (len(rows), 4, 4))
    ptr = np.r_[0, np.cumsum(np.bincount(np.asarray(rows, dtype=int), minlength=nq))]
    return sp.bsr_matrix((blocks, np.asarray(cols, dtype=int), ptr), shape=(4 * nq, 4 * nq)).tocsr()


def f32_product(A, B):
    """FP64 product of the FP32-rounded matrix and complex64-rounded block."""
    A32 = A.copy()
    A32.data = A32.data.real.astype(np.float32).astype(np.float64) + 1j * A32.data.imag.astype(np.float32)
    return A32 @ B.astype(np.complex64).astype(np.complex128)


@pytest.fixture
def fp32_ctx(hip_ctx):
    hip_ctx.set_tuning("fp32_math", 1)
    yield hip_ctx
    hip_ctx.set_tuning("fp32_math", 0)
    hip_ctx.set_tuning("bsr_mfma", 1)


@pytest.mark.parametrize("w", [4, 8, 12, 16, 32, 48, 64, 96])
def test_mfma_products_match_scipy_and_valu_kernel(fp32_ctx, w):
    from kubomsd_amd.backends.hip import KMSD_BSR4, HipMatrix
    for A in (ragged_bsr(), 1j * ragged_bsr(seed=23), cubic_orbital_model(3, seed=4411).S):
        A = sp.csr_matrix(A)
        m = HipMatrix(fp32_ctx, A, fmt=KMSD_BSR4)
        B = random_block(A.shape[0], w, 100 + w)
        want = f32_product(A, B)
        fp32_ctx.set_tuning("bsr_mfma", 1)
        got = host(m.matmul(to_dev(fp32_ctx, B, "c64")))
        fp32_ctx.set_tuning("bsr_mfma", 0)
        valu = host(m.matmul(to_dev(fp32_ctx, B, "c64")))
        assert rel_err(got, want) < 2e-6, w
        np.testing.assert_array_equal(got, valu)
        # rows of empty block rows are exactly zero
        empty = np.flatnonzero(np.diff(A.indptr) == 0)
        assert np.all(got[empty] == 0)


def test_mfma_metric_solve_reaches_fp64_residual(hip_ctx):
    """complex64 metric solves (FP32 inner PCG on the matrix-core products,
    FP64 refinement) reach the FP64 residual with the same inner iteration
    counts (within 2) as the VALU products with FP64 row sums on the
    simple-cubic model; the solutions differ in rounding, which shows the
    flag switches the inner products' kernel."""
    import scipy.linalg as la
    from kubomsd_amd.backends.hip import HipMetric
    m = cubic_orbital_model(4, seed=4411)
    S = sp.csr_matrix(m.S)
    blocks = [(a, b, la.cholesky(m.S[a:b, a:b].toarray(), lower=True)) for a, b in m.atom_slices]
    B = random_block(S.shape[0], 16, 5)
    out = {}
    hip_ctx.set_tuning("cheb_inner", 0)  # PCG inner solves throughout (their S products are compared)
    try:
        for flag in (1, 0):
            hip_ctx.set_tuning("bsr_mfma", flag)
            met = HipMetric(hip_ctx, S, S.diagonal().real, blocks)
            Y, rep = met.solve(to_dev(hip_ctx, B, "c64"), 1e-10, 0.0, 1000)
            Yh = host(Y)
            res = np.linalg.norm(S @ Yh - B.astype(np.complex64), axis=0) / np.linalg.norm(B, axis=0)
            assert all(rep["converged"]) and res.max() < 1e-6
            out[flag] = (Yh, met.diagnostics()["max_iterations_sum"])
    finally:
        hip_ctx.set_tuning("bsr_mfma", 1)
        hip_ctx.set_tuning("cheb_inner", 1)
    assert rel_err(out[1][0], out[0][0]) < 1e-6
    assert not np.array_equal(out[1][0], out[0][0])
    assert abs(out[1][1] - out[0][1]) <= 2


def test_autotune_picks_tile_size_for_plane_matrices(hip_ctx, tmp_path):
    """kmsd_autotune with KMSD_TUNE_TILE also times the plane-tiled walk's tile
    size for a 4x4-block matrix with a plane stride (simple cubic 16^3: 256
    block rows per plane), applies it, and products with it still match
    SciPy; without a plane stride the context's tile size is reported."""
    import torch
    from kubomsd_amd.backends import hip as H
    m = cubic_orbital_model(16, seed=4411)
    S = H.HipMatrix(hip_ctx, m.S)
    assert S.info()["format"] == "bsr4" and S.info()["plane_stride"] == 256
    H._TUNED.clear()
    try:
        choice = hip_ctx.autotune(S, 16, torch.complex64, cache_path=str(tmp_path / "t.json"), f32_values=True,
                                  tile=True)
        assert choice["tile_bytes"] in (0, 1 << 19, 1 << 20, 1 << 21)
        B = random_block(m.n, 16, 3)
        got = host(S.matmul(to_dev(hip_ctx, B, "c64")))
        assert rel_err(got, m.S @ B.astype(np.complex64).astype(np.complex128)) < 2e-6
        flat = H.HipMatrix(hip_ctx, ragged_bsr(), fmt=H.KMSD_BSR4)
        assert flat.info()["plane_stride"] == 0
        H._TUNED.clear()
        other = hip_ctx.autotune(flat, 16, torch.complex64, cache_path=str(tmp_path / "t.json"), tile=True)
        assert other["tile_bytes"] == choice["tile_bytes"]  # unchanged: no plane to tile
    finally:
        H._TUNED.clear()
        hip_ctx.set_tuning("cpl", 0)
        hip_ctx.set_tuning("groups", 0)
        hip_ctx.set_tuning("tile_bytes", 1 << 20)


@pytest.mark.parametrize("dtype, w", [("c64", 8), ("c64", 32), ("c64", 48), ("c64", 64), ("c64", 128),
                                      ("c128", 16), ("c128", 24), ("c128", 32), ("c128", 64)])
def test_lds_staged_fp64_products_bitwise(hip_ctx, dtype, w):
    """FP64-valued 4x4-block products (bsr4_lds_kernel: block columns and
    values staged in LDS per chunk) sum in bsr4_step_kernel's order, so the
    two kernels ("bsr_lds" 1 / 0) agree bit for bit: plain products on ragged
    rows (0..45 blocks: several chunks of 8, 16 or 32 blocks) and real or
    imaginary-only values, and against SciPy."""
    from kubomsd_amd.backends.hip import KMSD_BSR4, HipMatrix
    try:
        for A in (ragged_bsr(), 1j * ragged_bsr(seed=29), cubic_orbital_model(3, seed=4411).H):
            A = sp.csr_matrix(A)
            m = HipMatrix(hip_ctx, A, fmt=KMSD_BSR4)
            B = random_block(A.shape[0], w, 300 + w)
            out = []
            for flag in (1, 0):
                hip_ctx.set_tuning("bsr_lds", flag)
                out.append(host(m.matmul(to_dev(hip_ctx, B, dtype))))
            np.testing.assert_array_equal(out[0], out[1])
            Bq = B.astype(np.complex64).astype(np.complex128) if dtype == "c64" else B
            assert rel_err(out[0], A @ Bq) < (1e-13 if dtype == "c128" else 2e-6)
    finally:
        hip_ctx.set_tuning("bsr_lds", 1)


@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_lds_staged_chain_and_solve_bitwise(hip_ctx, dtype):
    """Moment chains of S^-1 H (H products with the recursion's BETA/P terms
    and moment partials, FP64 refinement residuals of the metric solves) are
    bit for bit the same with either FP64-valued 4x4-block kernel."""
    import scipy.linalg as la
    from kubomsd_amd.backends.hip import HipEngine, HipMetric
    m = cubic_orbital_model(4, seed=4411)
    blocks = [(a, b, la.cholesky(m.S[a:b, a:b].toarray(), lower=True)) for a, b in m.atom_slices]
    B = random_block(m.n, 32, 9)
    outs = []
    try:
        for flag in (1, 0):
            hip_ctx.set_tuning("bsr_lds", flag)
            met = HipMetric(hip_ctx, m.S, m.S.diagonal().real, blocks)
            eng = HipEngine(hip_ctx, m.H, met, [(-1, [])] * 3, 1e-10, 0.0, 1000)
            outs.append(host(eng.moments(to_dev(hip_ctx, B, dtype), 24, 0.0, 40.0)))
    finally:
        hip_ctx.set_tuning("bsr_lds", 1)
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("policy, interval", [("fixed", 3), ("adaptive", 4)])
def test_mfma_products_in_refresh_and_adaptive_iterations(hip_ctx, policy, interval):
    """complex64 solves whose inner iterations take the staged path (true-
    residual refreshes every `interval` iterations, or the adaptive policy):
    the refresh's T = Q - S Y product also runs on the matrix cores (F_P
    epilogue), and the solves still reach the FP64 residual."""
    import scipy.linalg as la
    from kubomsd_amd.backends.hip import HipMetric
    m = cubic_orbital_model(4, seed=4411)
    S = sp.csr_matrix(m.S)
    blocks = [(a, b, la.cholesky(m.S[a:b, a:b].toarray(), lower=True)) for a, b in m.atom_slices]
    B = random_block(S.shape[0], 12, 21)
    met = HipMetric(hip_ctx, S, S.diagonal().real, blocks)
    met.set_residual_policy(policy, interval, 0.1)
    Y, rep = met.solve(to_dev(hip_ctx, B, "c64"), 1e-10, 0.0, 1000)
    Bq = B.astype(np.complex64).astype(np.complex128)
    res = np.linalg.norm(Bq - S @ host(Y), axis=0) / np.linalg.norm(Bq, axis=0)
    assert all(rep["converged"]) and max(rep["relative_residual"]) <= 1e-10
    assert res.max() < 1e-6  # the solution is rounded once to complex64
    assert met.diagnostics()["true_residual_checks"] >= 1
