"""The prefetching fixed-width ELL product kernel (widths 3, 4, 6) against
the generic row kernel on the same launch shape: bitwise equal moments,
chains and actions (same summation order), for real, imaginary and complex
matrix values in complex64 and complex128 blocks, with the single value of
a uniform matrix read once or the value array streamed; with the automatic
(occupancy-derived) workgroup count, moments stay within rounding of the
FP64 oracle."""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import hip_engine, host, oracle_engine, phases, rel_err, to_dev
from kubomsd_amd.synthetic import Model, honeycomb_lattice, square_lattice

pytestmark = pytest.mark.gpu


def circulant6(n=3000, seed=41):
    """Hermitian matrix with exactly six off-diagonal entries per row
    (offsets +-1, +-7, +-31 mod n), complex values, real diagonal: an ELL
    width-6 product with complex values."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    i = np.arange(n)
    for off in (1, 7, 31):
        v = rng.normal(size=n) + 1j * rng.normal(size=n)
        j = (i + off) % n
        rows += [i, j]
        cols += [j, i]
        vals += [v, v.conj()]
    rows.append(i)
    cols.append(i)
    vals.append(rng.uniform(-1, 1, n).astype(complex))
    H = sp.csr_matrix((np.concatenate(vals) * .3, (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    H.sum_duplicates()
    H.sort_indices()
    return Model(H, sp.identity(n, dtype=complex, format="csr"), {}, np.zeros((n, 3)))


MODELS = {
    "honeycomb_w3": (lambda: honeycomb_lattice(24, t=-2.7, W=1.0), 3, 8.9),
    "square_w4": (lambda: square_lattice(30, W=2.0), 4, 5.1),
    "circulant_w6": (circulant6, 6, 4.0),
}


def _stored_per_row(ctx, H):
    from kubomsd_amd.backends.hip import HipMatrix
    info = HipMatrix(ctx, H).info()
    assert info["format"] == "csr"
    return info["stored"] / info["n"]


UNIFORM = {"honeycomb_w3": True, "square_w4": True, "circulant_w6": False}


@pytest.mark.parametrize("name", sorted(MODELS))
def test_uniform_value_detection(hip_ctx, name):
    """Periodic lattices with one hopping store full ELL rows of one value,
    which the prefetching kernel reads once (the bitwise comparison with the
    generic kernel, which streams the value array, is the test above)."""
    from kubomsd_amd.backends.hip import HipMatrix
    assert HipMatrix(hip_ctx, MODELS[name][0]().H).info()["uniform_values"] is UNIFORM[name]


def test_one_changed_bond_is_not_uniform(hip_ctx):
    """One bond of a periodic square lattice changed: the values are no longer
    uniform, and the prefetching kernel (streamed values) still equals the
    generic kernel bitwise and SciPy; an open lattice (padded ELL rows) is
    never uniform."""
    from kubomsd_amd.backends.hip import HipMatrix
    m = square_lattice(30, W=2.0)
    H = m.H.tolil()
    H[0, 1] = H[1, 0] = -1.25
    H = H.tocsr()
    A = HipMatrix(hip_ctx, H)
    assert not A.info()["uniform_values"]
    B = phases(m.n, 4, 9)
    hip_ctx.set_tuning("groups", 64)
    try:
        outs = []
        for ell in (0, 1):
            hip_ctx.set_tuning("ell_kernel", ell)
            outs.append(host(A.matmul(to_dev(hip_ctx, B))))
    finally:
        hip_ctx.set_tuning("ell_kernel", 1)
        hip_ctx.set_tuning("groups", 0)
    np.testing.assert_array_equal(outs[0], outs[1])
    assert rel_err(outs[1], H @ B) < 1e-13
    n = 30 * 30
    i = np.arange(n)
    right = i[(i % 30) < 29]
    open_H = sp.csr_matrix((-np.ones(2 * len(right)), (np.r_[right, right + 1], np.r_[right + 1, right])),
                           shape=(n, n)) + sp.identity(n, format="csr")
    info = HipMatrix(hip_ctx, open_H).info()
    assert info["stored"] > open_H.nnz - n and not info["uniform_values"]


@pytest.mark.parametrize("name", sorted(MODELS))
@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_ell_kernel_bitwise_equals_generic(hip_ctx, name, dtype):
    build, width, half = MODELS[name]
    m = build()
    assert _stored_per_row(hip_ctx, m.H) == width  # ELL of this width, no padding
    eng = hip_engine(hip_ctx, m)
    B = to_dev(hip_ctx, phases(m.n, 6, 17), dtype)
    hip_ctx.set_tuning("groups", 64)  # same partial-sum partition for both kernels
    try:
        runs = []
        for ell in (0, 1):
            hip_ctx.set_tuning("ell_kernel", ell)
            runs.append((host(eng.moments(B, 97, 0.0, half)), host(eng.chain(B, 9, 0.0, half)),
                         host(eng.h_action(B))))
    finally:
        hip_ctx.set_tuning("ell_kernel", 1)
        hip_ctx.set_tuning("groups", 0)
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("name", sorted(MODELS))
def test_ell_kernel_auto_groups_matches_oracle(hip_ctx, name):
    build, width, half = MODELS[name]
    m = build()
    Bh = phases(m.n, 5, 23)
    want = oracle_engine(m).moments(Bh, 129, 0.0, half)
    got = host(hip_engine(hip_ctx, m).moments(to_dev(hip_ctx, Bh), 129, 0.0, half))
    assert rel_err(got, want) < 1e-12
    got32 = host(hip_engine(hip_ctx, m).moments(to_dev(hip_ctx, Bh, "c64"), 129, 0.0, half))
    assert rel_err(got32, want) < 2e-5


def test_velocity_bond_matrix_imaginary_values(hip_ctx):
    """Lattice velocity V = i H o D has purely imaginary values (stored as one
    real array): its ELL product through the prefetching kernel equals the
    generic kernel and SciPy."""
    from kubomsd_amd.backends.hip import HipMatrix
    m = square_lattice(30, W=2.0)
    V = m.velocity["x"]
    A = HipMatrix(hip_ctx, V)
    B = phases(m.n, 8, 5)
    hip_ctx.set_tuning("groups", 64)
    try:
        outs = []
        for ell in (0, 1):
            hip_ctx.set_tuning("ell_kernel", ell)
            outs.append(host(A.matmul(to_dev(hip_ctx, B))))
    finally:
        hip_ctx.set_tuning("ell_kernel", 1)
        hip_ctx.set_tuning("groups", 0)
    np.testing.assert_array_equal(outs[0], outs[1])
    assert rel_err(outs[1], V @ B) < 1e-13
