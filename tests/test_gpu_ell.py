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
    "honeycomb_open_w3": (lambda: honeycomb_lattice(24, t=-2.7, W=1.0, periodic=False), 3, 8.9),
    "square_w4": (lambda: square_lattice(30, W=2.0), 4, 5.1),
    "circulant_w6": (circulant6, 6, 4.0),
}


def _stored_per_row(ctx, H):
    from kubomsd_amd.backends.hip import HipMatrix
    info = HipMatrix(ctx, H).info()
    assert info["format"] == "csr"
    return info["stored"] / info["n"]


UNIFORM = {"honeycomb_w3": True, "honeycomb_open_w3": True, "square_w4": True, "circulant_w6": False}


@pytest.mark.parametrize("name", sorted(MODELS))
def test_uniform_value_detection(hip_ctx, name):
    """Lattices with one hopping store ELL rows of one value (an open
    lattice's short rows are padded), which the prefetching kernel reads once,
    with zero in padding slots (the bitwise comparison with the generic
    kernel, which streams the value array, is the test above)."""
    from kubomsd_amd.backends.hip import HipMatrix
    assert HipMatrix(hip_ctx, MODELS[name][0]().H).info()["uniform_values"] is UNIFORM[name]


def test_one_changed_bond_is_not_uniform(hip_ctx):
    """One bond of a periodic square lattice changed: the values are no longer
    uniform, and the prefetching kernel (streamed values) still equals the
    generic kernel bitwise and SciPy; the padding of an open lattice's short
    rows does not count as a second value."""
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
    assert info["stored"] > open_H.nnz - n and info["uniform_values"]


def test_sampler_scaled_copy_of_one_value_overlap(hip_ctx):
    """An overlap stored as ELL with one off-diagonal value and a varying
    diagonal: the sampler's scaled copy Z S Z (Z = diag(S)^-1/2) has a value
    per bond, must not keep the one-value storage of S, and its samples equal
    the oracle's."""
    from helpers import O
    from kubomsd_amd.backends.hip import HipMatrix, HipMetric
    m = square_lattice(20, W=0.0)
    A = (m.H - sp.diags(m.H.diagonal())).tocsr()
    rng = np.random.default_rng(12)
    S = (sp.diags(1.0 + 0.2 * rng.random(m.n)) + 0.05 * A).astype(np.complex128).tocsr()
    S.sort_indices()
    assert HipMatrix(hip_ctx, S).info()["uniform_values"]
    xi = phases(m.n, 3, 4)
    coeff = np.array([1.1, -.2, .05, -.01])
    got = host(HipMetric(hip_ctx, S, S.diagonal().real).sample(to_dev(hip_ctx, xi), coeff, 1.0, 0.9))
    want = O.Metric(O.Csr(S, 1), S.diagonal().real, (), 1).sample(xi, coeff, 1.0, 0.9)
    assert rel_err(got, want) < 1e-13


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


def _bondwise_rescaled(m):
    """The lattice's velocities with every bond scaled by its own factor in
    [1, 1.1) (the same for (i, j) and (j, i)): about one distinct value per
    bond instead of +-t a / hbar."""
    for a in ("x", "y"):
        V = sp.coo_matrix(m.velocity[a])
        key = np.minimum(V.row, V.col).astype(np.int64) * m.n + np.maximum(V.row, V.col)
        f = 1 + 0.1 * ((key * 2654435761) % 1000003) / 1000003.0
        m.velocity[a] = sp.csr_matrix((V.data * f, (V.row, V.col)), shape=V.shape)
    return m


@pytest.mark.parametrize("dtype", ["c64", "c128"])
@pytest.mark.parametrize("nb", [1, 3, 4])
@pytest.mark.parametrize("direction", [(1.0, 0.0, 0.0), (0.0, -0.5, 0.0)])
@pytest.mark.parametrize("values", ["lattice", "bondwise", "honeycomb"])
def test_fused_velocity_chain_bitwise_equals_regrouped(hip_ctx, dtype, nb, direction, values):
    """kmsd_chain_v forms V T_k inside the chain's steps when the velocity's
    entries lie in H's ELL slots: bitwise the same table as the regrouping
    path it replaces (chain, one velocity product over all orders, regrouping
    back), and it allocates none of that path's two L x N x nb buffers; on
    the square and honeycomb lattices (ELL widths 4 and 3) and with a
    bond-wise rescaled velocity (a value per bond)."""
    m = honeycomb_lattice(16, t=-2.7, W=1.0) if values == "honeycomb" else square_lattice(30, W=2.0)
    if values == "bondwise":
        m = _bondwise_rescaled(m)
        assert len(np.unique(m.velocity["x"].data)) > 256
    L = 40
    B = to_dev(hip_ctx, phases(m.n, nb, 31), dtype)
    eng = hip_engine(hip_ctx, m)
    before = hip_ctx.memory()["used_bytes"]
    fused = host(eng.chain_v(B, L, 0.0, 5.1, list(direction)))
    grown = hip_ctx.memory()["used_bytes"] - before
    hip_ctx.set_tuning("fuse_velocity", 0)
    try:
        before = hip_ctx.memory()["used_bytes"]
        plain = host(eng.chain_v(B, L, 0.0, 5.1, list(direction)))
        grown_plain = hip_ctx.memory()["used_bytes"] - before
    finally:
        hip_ctx.set_tuning("fuse_velocity", 1)
    np.testing.assert_array_equal(fused, plain)
    assert np.abs(fused).max() > 0
    block = L * m.n * nb * (8 if dtype == "c64" else 16)
    assert grown < block <= grown_plain / 2


def test_velocity_outside_hamiltonian_pattern_is_not_fused(hip_ctx):
    """A velocity entry where H has none (and two axes at once) keep the
    regrouping path; the table still equals chain + per-order velocity
    products."""
    m = square_lattice(16, W=2.0)
    V = m.velocity["x"].tolil()
    V[0, 2] = 0.5j
    V[2, 0] = -0.5j
    m.velocity["x"] = V.tocsr()
    eng = hip_engine(hip_ctx, m)
    B = to_dev(hip_ctx, phases(m.n, 2, 8))
    for direction in ([1.0, 0, 0], [0.5, 0.25, 0]):
        ch = eng.chain(B, 9, 0.0, 5.1)
        want = np.stack([host(eng.v_cov(ch[k].contiguous(), direction)) for k in range(9)])
        got = host(eng.chain_v(B, 9, 0.0, 5.1, direction))
        assert rel_err(got, want) < 1e-14
