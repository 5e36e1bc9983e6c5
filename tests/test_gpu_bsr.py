"""4x4 block-sparse (BSR4) device storage: products, Chebyshev chains,
moments, metric solves and the sampler's congruence copy, each against SciPy
or the FP64 oracle and against the same operator stored as CSR.

Tolerances: complex128 blocks agree with SciPy to 1e-13 of the largest entry
(FP64 products in a different summation order); complex64 blocks to 2e-6
(FP64 row sums, one complex64 rounding of inputs and outputs; FP32 values only
inside the metric's refinement, which still reaches its FP64 residual)."""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import host, oracle_engine, random_block, rel_err, to_dev
from kubomsd_amd.synthetic import cubic_orbital_model, honeycomb_lattice

pytestmark = pytest.mark.gpu


def cubic(L=3):
    return cubic_orbital_model(L, seed=4411)


def block_complex(nq=16, density=.3, seed=5):
    """Random complex matrix of 4x4 blocks (Hermitian), block density `density`."""
    rng = np.random.default_rng(seed)
    mask = sp.random(nq, nq, density=density, random_state=seed, format="csr")
    mask = ((mask + mask.T + sp.identity(nq)) != 0).astype(float)
    coo = mask.tocoo()
    blocks = rng.normal(size=(coo.nnz, 4, 4)) + 1j * rng.normal(size=(coo.nnz, 4, 4))
    M = sp.bsr_matrix((blocks, coo.col, np.r_[0, np.cumsum(np.bincount(coo.row, minlength=nq))]),
                      shape=(4 * nq, 4 * nq)).tocsr()
    return ((M + M.conj().T) / 2).tocsr()


@pytest.mark.parametrize("w", [1, 3, 8, 20, 24, 32, 64])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_bsr_matmul_matches_scipy(hip_ctx, w, dtype):
    from kubomsd_amd.backends.hip import HipMatrix
    for A in (cubic().H, block_complex()):
        m = HipMatrix(hip_ctx, A)
        assert m.info()["format"] == "bsr4"
        B = random_block(A.shape[0], w, 7 + w)
        got = host(m.matmul(to_dev(hip_ctx, B, dtype)))
        Bq = B.astype(np.complex64).astype(np.complex128) if dtype == "c64" else B
        assert rel_err(got, A @ Bq) < (1e-13 if dtype == "c128" else 2e-6)


def test_auto_keeps_csr_for_sparse_rows_and_bsr_needs_multiple_of_four(hip_ctx):
    from kubomsd_amd.backends.hip import KMSD_BSR4, HipMatrix
    from kubomsd_amd.errors import KuboMSDError
    assert HipMatrix(hip_ctx, honeycomb_lattice(6).H).info()["format"] == "csr"
    A = sp.random(10, 10, density=.5, random_state=3, format="csr") + sp.identity(10)
    with pytest.raises(KuboMSDError, match="multiple of 4"):
        HipMatrix(hip_ctx, A, fmt=KMSD_BSR4)
    m = HipMatrix(hip_ctx, sp.identity(12, format="csr"), fmt=KMSD_BSR4)  # forced: still exact
    got = host(m.matmul(to_dev(hip_ctx, random_block(12, 2, 1))))
    np.testing.assert_allclose(got, random_block(12, 2, 1), rtol=0, atol=1e-15)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_bsr_chain_and_moments_equal_csr_and_oracle(hip_ctx, dtype):
    """Orthogonal block model: moments of BSR H vs CSR H vs the oracle."""
    from kubomsd_amd.backends.hip import KMSD_CSR, HipEngine, HipMatrix, HipMetric
    A = block_complex(nq=24, seed=9)
    n = A.shape[0]
    S = sp.identity(n, dtype=np.complex128, format="csr")
    bound = float(abs(A).sum(axis=1).max()) * 1.01
    B = random_block(n, 5, 3)
    outs = []
    for fmt in (None, KMSD_CSR):
        H = HipMatrix(hip_ctx, A) if fmt is None else HipMatrix(hip_ctx, A, fmt=fmt)
        eng = HipEngine(hip_ctx, H, HipMetric(hip_ctx, S, np.ones(n)), [(-1, [])] * 3, 1e-10, 0.0, 1000)
        outs.append(host(eng.moments(to_dev(hip_ctx, B, dtype), 40, 0.0, bound)))
    from types import SimpleNamespace
    model = SimpleNamespace(H=A, S=S, velocity={})
    want = oracle_engine(model).moments(B.astype(np.complex64).astype(complex) if dtype == "c64" else B, 40, 0.0,
                                        bound)
    tol = 1e-12 if dtype == "c128" else 5e-6
    assert rel_err(outs[0], want) < tol
    assert rel_err(outs[0], outs[1]) < tol


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_bsr_metric_solve_and_nonorthogonal_action(hip_ctx, dtype):
    """Block-Jacobi PCG on a BSR S (simple cubic, 4 orbitals per atom) and the
    S^-1 H action against the oracle."""
    import scipy.linalg as la
    from kubomsd_amd.backends.hip import HipEngine, HipMetric
    m = cubic()
    n = m.n
    blocks = [(a, b, la.cholesky(m.S[a:b, a:b].toarray(), lower=True)) for a, b in m.atom_slices]
    met = HipMetric(hip_ctx, m.S, m.S.diagonal().real, blocks)
    assert met.S.info()["format"] == "bsr4"
    B = random_block(n, 6, 11)
    Bd = to_dev(hip_ctx, B, dtype)
    Y, rep = met.solve(Bd, 1e-10, 0.0, 1000)
    Bq = host(Bd)
    res = np.linalg.norm(Bq - m.S @ host(Y), axis=0) / np.linalg.norm(Bq, axis=0)
    assert all(rep["converged"]) and res.max() < (1e-10 if dtype == "c128" else 1e-6)
    eng = HipEngine(hip_ctx, m.H, met, [(-1, [])] * 3, 1e-12, 0.0, 1000)
    want = oracle_engine(m, rtol=1e-12, blocks=blocks).h_action(Bq)
    assert rel_err(host(eng.h_action(Bd)), want) < (1e-10 if dtype == "c128" else 2e-6)


def test_bsr_sampler_congruence_matches_csr(hip_ctx):
    from kubomsd_amd.backends.hip import KMSD_CSR, HipMatrix, HipMetric
    m = cubic()
    xi = np.exp(2j * np.pi * np.random.default_rng(2).random((m.n, 3)))
    coeff = np.array([1.1, -.2, .05, -.01])
    outs = []
    for fmt in (None, KMSD_CSR):
        S = HipMatrix(hip_ctx, m.S) if fmt is None else HipMatrix(hip_ctx, m.S, fmt=fmt)
        met = HipMetric(hip_ctx, S, m.S.diagonal().real)
        outs.append(host(met.sample(to_dev(hip_ctx, xi), coeff, 1.0, 0.9)))
    assert rel_err(outs[0], outs[1]) < 1e-13


def test_autotune_applies_and_caches_shape(hip_ctx, tmp_path):
    """kmsd_autotune picks a valid shape, the wrapper stores it per matrix
    signature, a second call reuses it, and products with the tuned shape
    still match SciPy."""
    import json
    import torch
    from kubomsd_amd.backends import hip as H
    m = H.HipMatrix(hip_ctx, cubic().H)
    path = tmp_path / "tuning.json"
    H._TUNED.clear()
    first = hip_ctx.autotune(m, 8, torch.complex64, cache_path=str(path))
    assert first["cpl"] in (1, 2, 4) and first["groups"] in (0, 1024, 2048, 4096)  # 0 = resident capacity
    table = json.loads(path.read_text())
    assert list(table.values()) == [first]
    H._TUNED.clear()
    assert hip_ctx.autotune(m, 8, torch.complex64, cache_path=str(path)) == first  # from the file
    B = random_block(m.n, 8, 4)
    got = host(m.matmul(to_dev(hip_ctx, B, "c64")))
    assert rel_err(got, m.host @ B.astype(np.complex64).astype(np.complex128)) < 2e-6
    hip_ctx.set_tuning("cpl", 0)  # restore the session defaults for later tests
    hip_ctx.set_tuning("groups", 0)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_plane_tiled_walk_equals_natural_order(hip_ctx, dtype):
    """A 16^3 simple-cubic block matrix has a plane stride of 256 block rows;
    products walk it in tiles across planes.  Tiles of 48 block rows (five
    full tiles and a 16-row remainder per plane) and the natural order give
    bitwise the same products (each row's sum keeps its order) and S^-1 H
    moments within rounding of each other and of the oracle."""
    from types import SimpleNamespace
    from kubomsd_amd.backends.hip import HipEngine, HipMatrix, HipMetric
    m = cubic_orbital_model(16, seed=4412)
    n, w = m.n, 8
    eb = 16 if dtype == "c128" else 8
    B = random_block(n, w, 21)
    H = HipMatrix(hip_ctx, m.H)
    assert H.info()["format"] == "bsr4" and H.info()["plane_stride"] == 256
    bound = 1.02 * float(abs(m.H).sum(axis=1).max()) / 0.25  # lambda_min(S) > 0.3 for this model
    outs, moms = [], []
    try:
        for tile in (0, 48 * 4 * w * eb, 1 << 20):
            hip_ctx.set_tuning("tile_bytes", tile)
            outs.append(host(H.matmul(to_dev(hip_ctx, B, dtype))))
            S = HipMatrix(hip_ctx, m.S)
            eng = HipEngine(hip_ctx, H, HipMetric(hip_ctx, S, m.S.diagonal().real), [(-1, [])] * 3, 1e-12, 0.0,
                            1000)
            moms.append(host(eng.moments(to_dev(hip_ctx, B[:, :4], dtype), 24, 0.0, bound)))
    finally:
        hip_ctx.set_tuning("tile_bytes", 1 << 20)
    np.testing.assert_array_equal(outs[1], outs[0])
    np.testing.assert_array_equal(outs[2], outs[0])
    tol = 1e-12 if dtype == "c128" else 5e-6
    assert rel_err(moms[1], moms[0]) < tol and rel_err(moms[2], moms[0]) < tol
    Bq = B[:, :4].astype(np.complex64).astype(complex) if dtype == "c64" else B[:, :4]
    want = oracle_engine(SimpleNamespace(H=m.H, S=m.S, velocity={})).moments(Bq, 24, 0.0, bound)
    assert rel_err(moms[1], want) < tol
