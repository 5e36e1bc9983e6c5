"""IC(0), cluster-block and direct metric solves on the device against the
oracle (reference cpp/native.cpp:92-121, 163-205, 237-297, 331-347) and
against dense answers, as the reference's own tests check them
(tests/unit/test_metric_ic0.py, test_metric_cluster.py,
test_metric_cholesky.py):
  - kmsd_ic0_factor equals the oracle's factor bit for bit (values and the
    failing row);
  - every mode solves S Y = B to the dense answer at rtol 1e-12 (complex128)
    and to the FP64-refined tolerance in complex64;
  - cluster blocks of mixed sizes (register and workgroup kernels) and the
    direct factor reproduce the oracle's solves to 1e-12, the direct mode's
    refinement steps and zero iterations included.
"""
import numpy as np
import pytest
import scipy.linalg as la
import scipy.sparse as sp

from helpers import host, random_block, rel_err, to_dev
from oracle import native as O
from test_oracle_metric_modes import LADDER, S4, banded_spd, ic_factor_csr, ic_lower

pytestmark = pytest.mark.gpu


def hip_metric(ctx, S, blocks=()):
    from kubomsd_amd.backends.hip import HipMetric
    S = sp.csr_matrix(S, dtype=np.complex128)
    return HipMetric(ctx, S, S.diagonal().real, blocks)


@pytest.mark.parametrize("ordering", ["natural", "rcm"])
@pytest.mark.parametrize("name", ["4x4", "band200", "band1001"])
def test_ic0_factor_bit_identical_to_oracle(hip_ctx, name, ordering):
    from kubomsd_amd.backends.hip import ic0_factor
    S = {"4x4": S4, "band200": banded_spd(200, 1), "band1001": banded_spd(1001, 4)}[name]
    lower, _, _ = ic_lower(S, ordering)
    for shift in (0.0, 1e-4):
        np.testing.assert_array_equal(ic0_factor(hip_ctx, lower, shift), O.ic0_factor(lower, shift))


def test_ic0_factor_failures_match_oracle(hip_ctx):
    from kubomsd_amd.backends.hip import ic0_factor
    from kubomsd_amd.errors import KuboMSDError, NumericalError
    lower, _, _ = ic_lower(LADDER, "natural")
    with pytest.raises(O.OracleError) as want:
        O.ic0_factor(lower, 0.0)
    with pytest.raises(NumericalError) as got:
        ic0_factor(hip_ctx, lower, 0.0)
    assert str(got.value) == str(want.value)
    np.testing.assert_array_equal(ic0_factor(hip_ctx, lower, 1e-2), O.ic0_factor(lower, 1e-2))
    with pytest.raises(KuboMSDError, match="invalid IC shift"):
        ic0_factor(hip_ctx, lower, -1.0)
    full = (lower + sp.triu(lower.conj().T, 1)).tocsr()
    full.sort_indices()
    with pytest.raises(KuboMSDError, match="diagonal at row end"):
        ic0_factor(hip_ctx, full, 0.0)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("ordering", ["natural", "rcm"])
def test_ic0_solves_match_dense_and_oracle(hip_ctx, dtype, ordering):
    n = 1001
    S = banded_spd(n, 4)
    lower, p, z = ic_lower(S, ordering)
    L = ic_factor_csr(lower)
    B = random_block(n, 6, 3)
    if dtype == "c128":
        B[:, 1] *= 1e-150
    B[:, 2] = 0
    om = O.Metric(O.Csr(S), S.diagonal().real, ())
    om.set_ic0(L, p, z)
    want, wrep = om.solve(B, 1e-12, 0., 1000)
    m = hip_metric(hip_ctx, S)
    m.set_ic0(L, p, z)
    got, rep = m.solve(to_dev(hip_ctx, B, dtype), 1e-12, 0., 1000)
    got = host(got)
    assert all(rep["converged"])
    dense = la.solve(S.toarray(), B)
    scale = np.maximum(np.abs(B).max(0), 1e-300)
    tol = 1e-10 if dtype == "c128" else 5e-6  # complex64 storage of the solution
    assert np.max(np.abs(got - want) / scale) < tol
    assert np.max(np.abs(got - dense) / scale) < tol
    if dtype == "c128":  # same preconditioner, same Krylov space: the same counts
        assert np.max(np.abs(np.array(rep["iterations"]) - np.array(wrep["iterations"]))) <= 1


def cluster_blocks(S, sizes, seed):
    n = S.shape[0]
    perm = np.random.default_rng(seed).permutation(n)
    cuts = np.cumsum([0] + list(sizes))
    assert cuts[-1] == n
    out = []
    for a, b in zip(cuts[:-1], cuts[1:]):
        rows = np.sort(perm[a:b]) if a % 2 else perm[a:b]
        out.append((rows.astype(np.int64), la.cholesky(S[np.ix_(rows, rows)], lower=True)))
    return out


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_cluster_blocks_of_mixed_size_match_oracle(hip_ctx, dtype):
    n = 400
    S = banded_spd(n, 7).toarray()
    blocks = cluster_blocks(S, [1, 7, 20, 64, 65, 150, 93], 5)  # register and workgroup kernels
    B = random_block(n, 5, 9)
    om = O.Metric(O.Csr(S), np.real(np.diag(S)), ())
    om.set_index_blocks(blocks)
    want, wrep = om.solve(B, 1e-12, 0., 1000)
    m = hip_metric(hip_ctx, S)
    m.set_index_blocks(blocks)
    got, rep = m.solve(to_dev(hip_ctx, B, dtype), 1e-12, 0., 1000)
    tol = 1e-10 if dtype == "c128" else 5e-6
    assert rel_err(host(got), want) < tol
    assert rel_err(host(got), la.solve(S, B)) < tol
    assert all(rep["converged"])
    if dtype == "c128":
        assert np.max(np.abs(np.array(rep["iterations"]) - np.array(wrep["iterations"]))) <= 1
    from kubomsd_amd.errors import KuboMSDError
    with pytest.raises(KuboMSDError, match="incomplete cluster coverage"):
        m.set_index_blocks(blocks[:-1])
    with pytest.raises(KuboMSDError, match="invalid cluster coverage"):
        m.set_index_blocks(blocks + [blocks[0]])


def test_large_contiguous_blocks(hip_ctx):
    """Atom blocks above the register kernel's 64 rows (the limit it had)."""
    n = 300
    S = banded_spd(n, 8).toarray()
    cuts = [0, 4, 100, 180, 300]
    blocks = [(a, b, la.cholesky(S[a:b, a:b], lower=True)) for a, b in zip(cuts[:-1], cuts[1:])]
    B = random_block(n, 4, 2)
    want, _ = O.Metric(O.Csr(S), np.real(np.diag(S)), blocks).solve(B, 1e-12, 0., 1000)
    got, rep = hip_metric(hip_ctx, S, blocks).solve(to_dev(hip_ctx, B), 1e-12, 0., 1000)
    assert rel_err(host(got), want) < 1e-10 and all(rep["converged"])


@pytest.mark.parametrize("n", [48, 200])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_direct_solves_and_refinement_match_oracle(hip_ctx, n, dtype):
    rng = np.random.default_rng(7201 + n)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    S = A @ A.conj().T + np.eye(n) * 20
    L = la.cholesky(S, lower=True)
    B = rng.normal(size=(n, 6)) + 1j * rng.normal(size=(n, 6))
    B[:, 0] = 0
    if dtype == "c128":
        B[:, 1] *= 1e-100
        B[:, 2] *= 1e100
    for factor, steps in ((L, 3), (L * (1 + 1e-6), 3)):
        om = O.Metric(O.Csr(S), np.real(np.diag(S)), [(0, n, factor)])
        om.set_direct(True, steps)
        want, wrep = om.solve(B, 1e-12, 0., 1000)
        m = hip_metric(hip_ctx, S, [(0, n, factor)])
        m.set_direct(True, steps)
        got, rep = m.solve(to_dev(hip_ctx, B, dtype), 1e-12, 0., 1000)
        scale = np.maximum(np.abs(B).max(0), 1)
        tol = 1e-11 if dtype == "c128" else 5e-7
        assert np.max(np.abs(host(got) - want) / scale) < tol
        assert np.max(np.abs(host(got) - la.solve(S, B)) / scale) < tol
        assert rep["iterations"] == [0] * 6 and all(rep["converged"])
        assert m.diagnostics()["refinement_steps"] == om.diagnostics()["refinement_steps"]
        assert max(rep["relative_residual"]) <= 1e-12


def test_direct_rejections_and_unrepresentable(hip_ctx):
    from kubomsd_amd.errors import KuboMSDError, NumericalError
    S = np.eye(2) * .5
    m = hip_metric(hip_ctx, S)
    with pytest.raises(KuboMSDError, match="invalid direct metric factor"):
        m.set_direct(True, 3)
    m = hip_metric(hip_ctx, S, [(0, 2, la.cholesky(S, lower=True))])
    m.set_direct(True, 3)
    with pytest.raises(NumericalError, match="representable"):
        m.solve(to_dev(hip_ctx, np.full((2, 1), 1.5e308 + 0j)), 1e-12, 0., 100)


def test_metric_modes_through_run_config(hip_ctx):
    """DeviceMetric built from RunConfig for each mode solves the original
    system (tests/unit/test_metric_ic0.py:9-29, test_metric_cholesky.py:15-37
    of the reference); the partition-file mode is rejected."""
    from kubomsd_amd.config import RunConfig
    from kubomsd_amd.errors import KuboMSDError
    from kubomsd_amd.operators.velocity import DeviceMetric
    base = dict(solver="stochastic-kpm", kernel="jackson", backend="hip", metric_rtol=1e-13)
    B = np.array([[1j, 0], [2, 0], [3j, 0], [1, 0]])
    for kw in (dict(preconditioner="ic0", metric_ic_ordering="natural"), dict(preconditioner="ic0"),
               dict(metric_solver="cholesky", metric_rtol=1e-12)):
        cfg = RunConfig([0], [0], **{**base, **kw})
        m = DeviceMetric(hip_ctx, sp.csr_matrix(S4), cfg)
        Y, rep = m.solve_s(to_dev(hip_ctx, B))
        np.testing.assert_allclose(host(Y), np.linalg.solve(S4.toarray(), B), atol=1e-12)
        assert np.all(rep.converged)
        d = m.diagnostics()
        if "metric_solver" in kw:
            assert d["solver"] == "cholesky" and d["resolved_preconditioner"] == "not-applicable"
            assert np.all(rep.iterations == 0) and d["factorizations"] == 1
        else:
            assert d["preconditioner_setup"]["shift"] == 0
    with pytest.raises(KuboMSDError, match="cluster-jacobi"):
        RunConfig([0], [0], **base, preconditioner="cluster-jacobi", metric_partition_file="p.json")


@pytest.mark.parametrize("mode", ["ic0", "cluster", "direct"])
def test_nonorthogonal_moments_with_each_mode_match_oracle(hip_ctx, mode):
    """The Chebyshev moments of S^-1 H are the same whichever metric mode
    solves S (oracle in the same mode), to the solve tolerance."""
    from helpers import definitions, phases
    from kubomsd_amd.backends.hip import HipEngine
    from kubomsd_amd.synthetic import square_lattice
    from kubomsd_amd.units import HBAR_EV_FS
    m = square_lattice(12, seed=3)
    n = m.n
    S = (sp.identity(n, format="csr") + .08 * (sp.random(n, n, density=.02, random_state=4) > 0)).astype(complex)
    S = ((S + S.conj().T) / 2).tocsr()
    S = S + sp.identity(n) * (abs(S).sum(1).max())  # diagonally dominant
    S = sp.csr_matrix(S)
    om = O.Metric(O.Csr(S), S.diagonal().real, ())
    hm = hip_metric(hip_ctx, S) if mode != "direct" else None
    if mode == "ic0":
        lower, p, z = ic_lower(S, "rcm")
        L = ic_factor_csr(lower)
        om.set_ic0(L, p, z)
        hm.set_ic0(L, p, z)
    elif mode == "cluster":
        blocks = cluster_blocks(S.toarray(), [70, 40, 34], 2)
        om.set_index_blocks(blocks)
        hm.set_index_blocks(blocks)
    else:
        F = la.cholesky(S.toarray(), lower=True)
        om = O.Metric(O.Csr(S), S.diagonal().real, [(0, n, F)])
        om.set_direct(True, 3)
        hm = hip_metric(hip_ctx, S, [(0, n, F)])
        hm.set_direct(True, 3)
    defs = definitions(m)
    oe = O.Engine(O.Csr(m.H), om, [(k, [O.Csr(x) for x in ms]) for k, ms in defs], 1e-12, 0., 1000)
    he = HipEngine(hip_ctx, m.H, hm, defs, 1e-12, 0., 1000, HBAR_EV_FS)
    B = phases(n, 4, 11)
    want = oe.moments(B, 24, 0.0, 6.0)
    got = host(he.moments(to_dev(hip_ctx, B), 24, 0.0, 6.0))
    assert rel_err(got, want) < 1e-9


@pytest.mark.parametrize("width", [4, 1, 7])
def test_cluster_sizes_strided_rows_and_extreme_columns(hip_ctx, width):
    """The reference GPU edition's grouped-triangular case
    (tests/gpu/test_metric_preconditioners.py:93-120): cluster sizes 1 … 257
    twice over a random row order, columns scaled by 0, 1e-180, 1e180, 1;
    here through the solve, which must return the dense answer per column."""
    sizes = [1, 13, 25, 64, 128, 256, 257] * 2
    n = sum(sizes)
    S = banded_spd(n, 731).toarray()
    blocks = cluster_blocks(S, sizes, 731)
    scales = np.resize(np.array([0, 1e-180, 1e180, 1.]), width)
    B = random_block(n, width, 12) * scales
    m = hip_metric(hip_ctx, S)
    m.set_index_blocks(blocks)
    got, rep = m.solve(to_dev(hip_ctx, B), 1e-12, 0., 1000)
    got = host(got)
    s = np.where(scales == 0, 1, scales)
    np.testing.assert_allclose(got / s, la.solve(S, B / s), rtol=1e-10, atol=1e-10)
    np.testing.assert_array_equal(got[:, scales == 0], 0)
    assert all(rep["converged"])


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_ic0_graph_replays_equal_eager_launches(hip_ctx, dtype):
    """The IC sweeps are captured once per argument set and replayed: every
    replay (many per solve, several solves with other device work between)
    gives bitwise the eager launches' solution."""
    n = 700
    S = banded_spd(n, 21)
    lower, p, z = ic_lower(S, "rcm")
    L = ic_factor_csr(lower)
    out = {}
    try:
        for graphs in (0, 1):
            hip_ctx.set_tuning("graphs", graphs)
            m = hip_metric(hip_ctx, S)
            m.set_ic0(L, p, z)
            res = []
            for seed in (1, 2, 1):
                B = to_dev(hip_ctx, random_block(n, 5, seed), dtype)
                Y, rep = m.solve(B, 1e-12, 0., 1000)
                res.append((host(Y), rep["iterations"]))
                (B * 2).sum().item()  # unrelated device work between solves
            out[graphs] = res
    finally:
        hip_ctx.set_tuning("graphs", 1)
    for (a, ia), (b, ib) in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)
        assert ia == ib
    np.testing.assert_array_equal(out[1][0][0], out[1][2][0])


def test_metric_solver_class_as_the_reference_uses_it(hip_ctx):
    """kubomsd_amd.metric.solve.MetricSolver with the reference's call
    patterns (tests/unit/test_metric_ic0.py:9-29, test_metric_cholesky.py:15-37,
    tests/gpu/test_metric_preconditioners.py:53-84)."""
    from types import SimpleNamespace
    from kubomsd_amd.config import RunConfig
    from kubomsd_amd.errors import NumericalError
    from kubomsd_amd.metric.solve import MetricSolver
    base = dict(solver="stochastic-kpm", kernel="jackson", backend="hip")
    for ordering in ("natural", "rcm"):
        m = MetricSolver(S4, RunConfig([0], [0], **base, preconditioner="ic0", metric_ic_ordering=ordering,
                                       metric_rtol=1e-13), hip_ctx)
        B = np.array([[1j, 0], [2, 0], [3j, 0], [1, 0]])
        Y, r = m.solve_s(B)
        np.testing.assert_allclose(host(Y), np.linalg.solve(S4.toarray(), B), atol=1e-12)
        assert np.all(r.converged)
    S = np.array([[2, .3j], [-.3j, .7]])
    m = MetricSolver(S, RunConfig([0], [0], **base, metric_solver="cholesky", metric_rtol=1e-12), hip_ctx)
    for _ in range(3):
        B = np.array([[1 + 1j, 0, 2e-180j, 1e180], [2j, 0, 1e-180, 2e180j]])
        Y, r = m.solve_s(B)
        s = B.max(axis=0, initial=1)
        np.testing.assert_allclose(host(Y) / s, np.linalg.solve(S, B) / s, rtol=1e-12, atol=1e-12)
        assert np.all(r.iterations == 0) and np.all(r.relative_residual <= 1e-12)
    d = m.diagnostics()
    assert d["factorizations"] == 1 and d["solver"] == "cholesky" and d["calls"] == 3
    with pytest.raises(NumericalError, match="cannot change prepared"):
        m.solve_s(B, SimpleNamespace(preconditioner="ic0"))
    # grouped block-Jacobi application (sizes 2, 3 and 5), and its negation
    rng = np.random.default_rng(318)
    sizes = [2, 3, 2, 3, 2, 3, 2, 3, 5]
    n = sum(sizes)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    S = A @ A.conj().T + np.eye(n)
    edges = np.r_[0, np.cumsum(sizes)]
    slices = list(zip(edges[:-1], edges[1:]))
    m = MetricSolver(S, RunConfig([0], [0], **base, preconditioner="block-jacobi", metric_rtol=1e-13), hip_ctx,
                     atom_slices=slices)
    B = rng.normal(size=(n, 8)) + 1j * rng.normal(size=(n, 8))
    want = np.empty_like(B[:, ::2])
    for a, b in slices:
        want[a:b] = la.solve(S[a:b, a:b], B[a:b, ::2], assume_a="pos")
    dev = to_dev(hip_ctx, B)[:, ::2]
    np.testing.assert_allclose(host(m.precondition(dev)), want, rtol=3e-13, atol=2e-14)
    np.testing.assert_allclose(host(m.precondition(-dev)), -want, rtol=3e-13, atol=2e-14)
    Y, r = m.solve_s(dev)
    np.testing.assert_allclose(host(Y), la.solve(S, B[:, ::2]), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("solver", ["pcg", "cholesky"])
def test_final_true_residual_acceptance_is_counted(hip_ctx, solver):
    """tests/unit/test_metric_adaptive.py:90-97: one final check, no refresh."""
    from kubomsd_amd.config import RunConfig
    from kubomsd_amd.metric.solve import MetricSolver
    m = MetricSolver(np.eye(2), RunConfig([0], [0], solver="stochastic-kpm", kernel="jackson", backend="hip",
                                          metric_solver=solver, metric_refine_maxiter=0), hip_ctx)
    m.solve_s(np.ones((2, 2)))
    d = m.diagnostics()
    assert d["true_residual_checks"] == 1 and d["true_residual_refreshes"] == 0


def test_adaptive_extremes_override_and_false_convergence(hip_ctx):
    """tests/unit/test_metric_adaptive.py:45-86: adaptive policy with column
    scales 0 / 1 / 1e-180 / 1e180 and a per-solve rtol override; and on an
    S of condition 1e4 with no periodic refresh, candidate convergence is
    checked against the true residual."""
    from types import SimpleNamespace
    from kubomsd_amd.config import RunConfig
    from kubomsd_amd.metric.solve import MetricSolver
    base = dict(solver="stochastic-kpm", kernel="jackson", backend="hip", metric_residual_policy="adaptive")
    rng = np.random.default_rng(4)
    U, _ = np.linalg.qr(rng.normal(size=(48, 48)) + 1j * rng.normal(size=(48, 48)))
    S = (U * np.geomspace(.01, 2, 48)) @ U.conj().T
    S = (S + S.conj().T) / 2
    B = (rng.normal(size=(48, 8)) + 1j * rng.normal(size=(48, 8)))[:, ::2] * [0, 1, 1e-180, 1e180]
    c = RunConfig([0], [0], **base, metric_check_interval=4, metric_rtol=1e-12, metric_maxiter=400)
    m = MetricSolver(S, c, hip_ctx)
    y, r = m.solve_s(B, SimpleNamespace(rtol=2e-13))
    y = host(y)
    assert np.all(r.relative_residual <= 2e-13)
    np.testing.assert_array_equal(y[:, 0], 0)
    for j in (1, 2, 3):
        s = np.max(abs(B[:, j]))
        np.testing.assert_allclose(y[:, j] / s, np.linalg.solve(S, B[:, j] / s), rtol=3e-11, atol=3e-11)
    assert m.diagnostics()["true_residual_checks"] > 0
    rng = np.random.default_rng(8)
    U, _ = np.linalg.qr(rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32)))
    S = (U * np.geomspace(1e-4, 1, 32)) @ U.conj().T
    S = (S + S.conj().T) / 2
    B = rng.normal(size=(32, 3)) + 1j * rng.normal(size=(32, 3))
    B[:, 0] = 0
    m = MetricSolver(S, RunConfig([0], [0], **base, metric_check_interval=10000, metric_maxiter=1000,
                                  metric_rtol=1e-12), hip_ctx)
    Y, r = m.solve_s(B)
    Y = host(Y)
    d = m.diagnostics()
    assert d["true_residual_checks"] > 1 and d["true_residual_refreshes"] > 0  # candidate checks, no periodic one
    assert np.all(r.converged) and r.iterations[0] == 0
    np.testing.assert_array_equal(Y[:, 0], 0)
    for j in (1, 2):
        assert np.linalg.norm(B[:, j] - S @ Y[:, j]) / np.linalg.norm(B[:, j]) <= 1e-12
