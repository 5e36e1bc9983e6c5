"""Set-up of the block factors and the IC(0) preconditioner of the metric
solve (reference src/kubo_msd/metric/preconditioners.py:11-106).  The
physical S is never modified: both act on fixed Hermitian positive definite
approximations.

  block factors: dense lower Cholesky factors of S's diagonal blocks (atom
      blocks), factored on the host in batched LAPACK calls, solved on the
      device;
  IC(0): W = D S D with D = diag(S)^-1/2, permuted by reverse Cuthill-McKee
      (or kept in natural order), incomplete Cholesky on W's lower pattern,
      computed on the device (kmsd_ic0_factor); on a pivot failure the next
      diagonal shift of 0, 1e-12, 1e-10, ..., 1e-2 up to `shift_max` is tried.
"""
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from ..errors import KuboMSDError, NumericalError

IC_SHIFTS = (0, 1e-12, 1e-10, 1e-8, 1e-6, 1e-4, 1e-2)


@dataclass
class PreconditionerData:
    kind: str
    factor: object = None                         # ic0: CSR factor of the permuted W
    permutation: object = None                    # ic0: original row of each factor row
    scaling: object = None                        # ic0: D per original row
    report: dict = field(default_factory=dict)


def block_cholesky(S, index_blocks, what="block"):
    """Lower Cholesky factors of S[rows][:, rows] for disjoint row lists
    covering S, in the given row order: S's entries are gathered into dense
    blocks at once and factored in one batched LAPACK call per block size."""
    S = sp.coo_matrix(S)
    n = S.shape[0]
    sizes = np.array([len(r) for r in index_blocks], np.int64)
    rows_all = np.concatenate([np.asarray(r, np.int64) for r in index_blocks]) if len(index_blocks) else \
        np.zeros(0, np.int64)
    if sizes.sum() != n or not np.array_equal(np.sort(rows_all), np.arange(n)):
        raise NumericalError(f"{what}s must partition the orbitals")
    owner = np.empty(n, np.int64)
    local = np.empty(n, np.int64)
    owner[rows_all] = np.repeat(np.arange(len(index_blocks)), sizes)
    local[rows_all] = np.arange(n) - np.repeat(np.cumsum(sizes) - sizes, sizes)
    inside = owner[S.row] == owner[S.col]
    r, c, v = S.row[inside], S.col[inside], S.data[inside]
    factors = [None] * len(index_blocks)
    for m in np.unique(sizes):
        qs = np.flatnonzero(sizes == m)
        slot = np.full(len(index_blocks), -1, np.int64)
        slot[qs] = np.arange(len(qs))
        sel = sizes[owner[r]] == m
        blk = np.zeros((len(qs), m, m), np.complex128)
        np.add.at(blk, (slot[owner[r[sel]]], local[r[sel]], local[c[sel]]), v[sel])
        try:
            L = np.linalg.cholesky(blk)
        except np.linalg.LinAlgError as exc:
            raise NumericalError(f"{what} of the overlap is not positive definite") from exc
        for k, q in enumerate(qs):
            factors[q] = L[k]
    return factors


def prepare_ic0(ctx, S, *, ordering, shift_max):
    """IC(0) data for the overlap S; S must be Hermitian positive definite
    (a shift in the preconditioner never admits an indefinite S)."""
    from .bounds import spd_bounds
    S = sp.csr_matrix(S, dtype=np.complex128)
    spd_bounds(S)
    return prepare_ic0_validated(ctx, S, ordering=ordering, shift_max=shift_max)


def prepare_ic0_validated(ctx, S, *, ordering, shift_max):
    from ..backends.hip import ic0_factor
    if ordering not in ("rcm", "natural") or not np.isfinite(shift_max) or shift_max < 0:
        raise NumericalError("invalid IC ordering/shift limit")
    start = time.perf_counter()
    z = 1 / np.sqrt(S.diagonal().real)
    W = (sp.diags(z) @ S @ sp.diags(z)).tocsr()
    W.eliminate_zeros()
    W.sort_indices()
    p = (sp.csgraph.reverse_cuthill_mckee(W, symmetric_mode=True) if ordering == "rcm"
         else np.arange(S.shape[0], dtype=np.int64))
    lower = sp.tril(W[p][:, p], format="csr")
    lower.sort_indices()
    attempts = []
    for shift in IC_SHIFTS:
        if shift > shift_max:
            continue
        try:
            values = ic0_factor(ctx, lower, shift)
        except KuboMSDError as exc:
            attempts.append(dict(shift=shift, status="FAILED", reason=str(exc)))
            continue
        attempts.append(dict(shift=shift, status="PASS"))
        L = sp.csr_matrix((values, lower.indices.copy(), lower.indptr.copy()), shape=S.shape)
        import scipy
        report = dict(ordering=ordering, shift=shift, shift_applies_to="M only, normalized coordinates",
                      attempts=attempts, nnz=L.nnz, factor_bytes=L.data.nbytes + L.indices.nbytes + L.indptr.nbytes,
                      minimum_pivot=float(np.min(L.diagonal().real ** 2)), seconds=time.perf_counter() - start,
                      permutation=p.tolist(), scipy_version=scipy.__version__, factor_device="hip")
        for a in (L.data, L.indices, L.indptr, z, p):
            a.setflags(write=False)
        return PreconditionerData("ic0", factor=L, permutation=np.asarray(p, np.int64), scaling=z, report=report)
    failure = NumericalError("IC(0) failed all allowed shifts; original S was not modified")
    failure.diagnostics = dict(attempts=attempts, ordering=ordering, shift_max=shift_max)
    raise failure
