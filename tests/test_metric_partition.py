"""Partition files and metric-mode configuration (host only): the
reference's known partition of tests/unit/test_metric_cluster.py:12-21,
identity and coverage validation (metric/partition.py:90-110), the canonical
hashes, and the RunConfig rules for the direct solver
(tests/unit/test_metric_cholesky.py:41-56)."""
import hashlib
import json

import numpy as np
import pytest
import scipy.sparse as sp

from kubomsd_amd.errors import KuboMSDError, NumericalError
from kubomsd_amd.metric.partition import (SCHEMA, content_hash, load_metric_partition, make_spatial_partition,
                                          matrix_fingerprint, periodic_distances)

S3 = sp.csr_matrix([[2, .1j, .4], [-.1j, 1, .1j], [.4, -.1j, 1.5]])
SLICES = [(0, 1), (1, 2), (2, 3)]
STRUCTURE = dict(cell_A=np.eye(3) * 10, positions_A=[[0, 0, 0], [5, 0, 0], [.2, 0, 0]], symbols=["A"] * 3)
IDENTITY = dict(physical_key="p", orbital_order_hash="o")


def partition(max_orbitals=2):
    return make_spatial_partition(STRUCTURE, SLICES, max_orbitals=max_orbitals,
                                  s_fingerprint=matrix_fingerprint(S3), **IDENTITY)


def test_known_partition_and_round_trip(tmp_path):
    d = partition()
    assert d["schema"] == SCHEMA and d["algorithm"] == "nearest-periodic-seed-v1"
    assert [b["atoms"] for b in d["blocks"]] == [[0, 2], [1]]
    assert [b["orbitals"] for b in d["blocks"]] == [[0, 2], [1]]
    path = tmp_path / "p.json"
    path.write_text(json.dumps(d))
    got = load_metric_partition(path, s_fingerprint=matrix_fingerprint(S3), atom_slices=SLICES, n_orbitals=3,
                                **IDENTITY)
    assert got["blocks"] == d["blocks"]
    # one block per atom when the cap is one orbital
    assert [b["atoms"] for b in partition(1)["blocks"]] == [[0], [1], [2]]


@pytest.mark.parametrize("edit, message", [
    (lambda d: d.update(physical_key="q"), "identity mismatch: physical_key"),
    (lambda d: d.update(s_fingerprint="0" * 64), "identity mismatch: s_fingerprint"),
    (lambda d: d["blocks"][0].update(orbitals=[0, 1]), "invalid whole-atom partition block"),
    (lambda d: d["blocks"].pop(), "cover each atom"),
    (lambda d: d.update(structure_hash="x"), "structure hash"),
])
def test_partition_validation(tmp_path, edit, message):
    d = partition()
    edit(d)
    path = tmp_path / "p.json"
    path.write_text(json.dumps(d))
    with pytest.raises(NumericalError, match=message):
        load_metric_partition(path, s_fingerprint=matrix_fingerprint(S3), atom_slices=SLICES, n_orbitals=3,
                              **IDENTITY)


def test_hashes_are_canonical():
    v = {"b": [1, 2], "a": np.arange(3)}
    want = hashlib.sha256(b'{"a":[0,1,2],"b":[1,2]}').hexdigest()
    assert content_hash(v) == want
    A = sp.csr_matrix(S3)
    h = hashlib.sha256()
    for part in (np.array(A.shape), A.indptr, A.indices, A.data):
        h.update(part.tobytes())
    assert matrix_fingerprint(S3) == h.hexdigest()
    # duplicates and unsorted columns give the same fingerprint once summed
    coo = sp.coo_matrix(S3)
    dup = sp.csr_matrix((np.r_[coo.data, 0 * coo.data[:1]], (np.r_[coo.row, coo.row[:1]], np.r_[coo.col, coo.col[:1]])),
                        shape=S3.shape)
    assert matrix_fingerprint(dup) == matrix_fingerprint(S3)


def test_periodic_distances_minimum_image_in_skew_cell():
    cell = np.array([[4.0, 0, 0], [3.9, .5, 0], [0, 0, 6]])  # strongly skewed
    rng = np.random.default_rng(1)
    d = rng.uniform(-10, 10, size=(40, 3))
    got = periodic_distances(d, cell)
    g = np.arange(-60, 61)  # the short vector a2 - a1 needs large multiples of both
    shifts = np.array(np.meshgrid(g, g, np.arange(-3, 4), indexing="ij")).reshape(3, -1).T @ cell
    brute = np.sqrt(((d[:, None, :] - shifts[None]) ** 2).sum(-1)).min(1)
    np.testing.assert_allclose(got, brute, rtol=1e-12)


@pytest.mark.parametrize("options", [
    {"metric_residual_policy": "adaptive"}, {"metric_check_interval": 4}, {"metric_gap_ratio": .2},
    {"preconditioner": "block-jacobi"}, {"metric_refine_maxiter": -1}])
def test_direct_solver_rejects_pcg_options(options):
    from kubomsd_amd.config import RunConfig
    with pytest.raises(KuboMSDError):
        RunConfig([0], [0], solver="stochastic-kpm", kernel="jackson", metric_solver="cholesky", **options)


def test_full_cholesky_is_cpu_exact_only():
    from kubomsd_amd.config import RunConfig
    with pytest.raises(KuboMSDError, match="CPU EXACT"):
        RunConfig([0], [0], solver="stochastic-kpm", kernel="jackson", preconditioner="full-cholesky")


def test_prepare_cluster_factors_each_block():
    """metric/preconditioners.py:37-46 of the reference: a lower Cholesky
    factor of S[rows][:, rows] per partition block, rows as given."""
    import scipy.linalg as la
    from kubomsd_amd.metric.preconditioners import prepare_cluster
    d = partition()
    data = prepare_cluster(S3, d)
    assert data.kind == "cluster-jacobi" and data.report["block_sizes"] == [2, 1]
    for (rows, L), block in zip(data.blocks, d["blocks"]):
        np.testing.assert_array_equal(rows, block["orbitals"])
        sub = S3.toarray()[np.ix_(rows, rows)]
        np.testing.assert_allclose(L @ L.conj().T, sub, rtol=1e-14, atol=1e-15)
        np.testing.assert_array_equal(L, np.tril(L))
    bad = dict(d, blocks=[{"atoms": [0, 1, 2], "orbitals": [0, 1, 2]}])
    indefinite = sp.csr_matrix(np.array([[1, 2, 0], [2, 1, 0], [0, 0, 1.]]))
    with pytest.raises(NumericalError, match="positive definite"):
        prepare_cluster(indefinite, bad)


def test_atom_block_cholesky_matches_per_block_factorization():
    """The batched set-up equals scipy.linalg.cholesky of every diagonal
    block (metric/setup.py:73-76 of the reference), for mixed block sizes."""
    import scipy.linalg as la
    from kubomsd_amd.operators.velocity import atom_block_cholesky
    rng = np.random.default_rng(1)
    n = 40
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    S = A @ A.conj().T + n * np.eye(n)
    S[np.abs(S) < 3] = 0
    S = (S + S.conj().T) / 2
    slices = [(0, 1), (1, 4), (4, 8), (8, 9), (9, 13), (13, 40)]
    got = atom_block_cholesky(sp.csr_matrix(S), slices)
    assert [(a, b) for a, b, _ in got] == slices
    for a, b, L in got:
        np.testing.assert_allclose(L, la.cholesky(S[a:b, a:b], lower=True), rtol=1e-13, atol=1e-13)
    bad = S.copy()
    bad[2, 2] = -50
    with pytest.raises(NumericalError, match="not positive definite"):
        atom_block_cholesky(sp.csr_matrix(bad), slices)
    with pytest.raises(NumericalError, match="partition"):
        atom_block_cholesky(sp.csr_matrix(S), [(0, 4), (5, 40)])
