"""Cluster-Jacobi partition files: disjoint whole-atom neighbourhoods of at
most `max_orbitals` orbitals, tied to one overlap matrix and one model
(reference src/kubo_msd/metric/partition.py:1-118).

The file format (schema "kubomsd-metric-partition-v1"), its identity fields
and the content and matrix hashes are the reference's, so partition files
written by either implementation load in the other.  The clustering is the
reference's "nearest-periodic-seed-v1": the lowest unassigned atom seeds a
block that takes the nearest unassigned atoms (periodic minimum-image
distance, ties by atom index) while the orbital count stays within the cap.
"""
import hashlib
import itertools
import json
from pathlib import Path

import numpy as np
import scipy.sparse as sp

from ..errors import NumericalError

SCHEMA = "kubomsd-metric-partition-v1"


def content_hash(value):
    """SHA-256 of the canonical JSON of `value` (sorted keys, no spaces,
    arrays as lists)."""
    text = json.dumps(value, sort_keys=True, separators=(",", ":"), default=lambda a: np.asarray(a).tolist())
    return hashlib.sha256(text.encode()).hexdigest()


def matrix_fingerprint(A):
    """SHA-256 over shape, pointers, indices and values of the sorted,
    duplicate-summed CSR form (reference src/kubo_msd/validation.py:20-23)."""
    A = sp.csr_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    h = hashlib.sha256()
    for part in (np.array(A.shape), A.indptr, A.indices, A.data):
        h.update(part.tobytes())
    return h.hexdigest()


def _check_slices(atom_slices, n):
    rows = []
    for a, b in atom_slices:
        if isinstance(a, bool) or isinstance(b, bool) or int(a) != a or int(b) != b or not 0 <= a < b <= n:
            raise NumericalError("invalid atom slices")
        rows.extend(range(a, b))
    if rows != list(range(n)):
        raise NumericalError("atom slices must partition all orbitals")


def periodic_distances(displacements, cell):
    """Minimum-image lengths of displacement rows in a (possibly skew) cell:
    the rounded fractional image bounds the answer, and the cell's smallest
    singular value bounds the lattice translations that could still improve
    it, which are all tried."""
    frac = displacements @ np.linalg.inv(cell)
    frac -= np.rint(frac)
    best = np.sum((frac @ cell) ** 2, axis=1)
    reach = np.sqrt(best.max(initial=0)) / np.linalg.svd(cell, compute_uv=False)[-1]
    span = int(np.ceil(reach + .5))
    for shift in itertools.product(range(-span, span + 1), repeat=3):
        d = (frac - np.asarray(shift)) @ cell
        best = np.minimum(best, np.sum(d * d, axis=1))
    return np.sqrt(best)


def make_spatial_partition(structure, atom_slices, *, max_orbitals, s_fingerprint, physical_key,
                           orbital_order_hash):
    """Partition dictionary (the file's content) for a periodic structure."""
    from scipy.spatial import cKDTree
    cell = np.asarray(structure["cell_A"], float)
    pos = np.asarray(structure["positions_A"], float)
    nat = len(list(structure["symbols"]))
    if (cell.shape != (3, 3) or pos.shape != (nat, 3) or len(atom_slices) != nat or not np.isfinite(cell).all()
            or not np.isfinite(pos).all() or abs(np.linalg.det(cell)) < 1e-12):
        raise NumericalError("invalid partition structure/atom mapping")
    n = atom_slices[-1][1]
    _check_slices(atom_slices, n)
    if not isinstance(max_orbitals, int) or max_orbitals < 1:
        raise NumericalError("invalid orbital cap")
    sizes = np.array([b - a for a, b in atom_slices])
    if sizes.max() > max_orbitals:
        raise NumericalError("atom exceeds orbital cap")
    frac = (pos @ np.linalg.inv(cell)) % 1
    tree = cKDTree(frac, boxsize=1.)
    sigma_min = np.linalg.svd(cell, compute_uv=False)[-1]
    free = set(range(nat))
    blocks = []
    while free:
        seed = min(free)
        k = min(32, nat)
        while True:
            fdist, near = tree.query(frac[seed], k=k, workers=1)
            cand = np.array([int(i) for i in np.atleast_1d(near) if i in free])
            dist = periodic_distances(pos[cand] - pos[seed], cell)
            order = np.lexsort((cand, dist))
            ranked, rdist = cand[order], dist[order]
            used, edge = int(sizes[seed]), None
            for atom, d in zip(ranked, rdist):
                if atom == seed:
                    continue
                if used + sizes[atom] > max_orbitals:
                    edge = d
                    break
                used += int(sizes[atom])
            # Atoms not yet queried lie at least sigma_min times their torus
            # distance away: stop once that exceeds the block's boundary.
            if k == nat or (edge is not None and np.atleast_1d(fdist)[-1] * sigma_min > edge + 1e-12):
                break
            k = min(nat, 2 * k)
        atoms, used = [seed], int(sizes[seed])
        for atom in ranked:
            atom = int(atom)
            if atom == seed:
                continue
            if used + sizes[atom] > max_orbitals:
                break
            atoms.append(atom)
            used += int(sizes[atom])
        free.difference_update(atoms)
        blocks.append(dict(atoms=atoms, orbitals=[j for i in atoms for j in range(*atom_slices[i])]))
    return dict(schema=SCHEMA, algorithm="nearest-periodic-seed-v1", s_fingerprint=s_fingerprint,
                physical_key=physical_key, orbital_order_hash=orbital_order_hash,
                structure_hash=content_hash(structure), n_orbitals=n, n_atoms=nat, max_orbitals=max_orbitals,
                blocks=blocks)


def load_metric_partition(path, *, s_fingerprint, physical_key, orbital_order_hash, atom_slices, n_orbitals):
    """Reads and validates a partition file against the overlap matrix and
    model it must belong to; every atom and orbital exactly once."""
    d = json.loads(Path(path).read_text())
    for key, want in dict(schema=SCHEMA, s_fingerprint=s_fingerprint, physical_key=physical_key,
                          orbital_order_hash=orbital_order_hash, n_orbitals=n_orbitals).items():
        if d.get(key) != want:
            raise NumericalError("partition identity mismatch: " + key)
    _check_slices(atom_slices, n_orbitals)
    if not isinstance(d.get("structure_hash"), str) or len(d["structure_hash"]) != 64:
        raise NumericalError("partition requires structure hash")
    atoms, orbitals = [], []
    for block in d["blocks"]:
        ids, rows = block["atoms"], block["orbitals"]
        if (not ids or any(type(i) != int or not 0 <= i < len(atom_slices) for i in ids)
                or rows != [j for i in ids for j in range(*atom_slices[i])] or len(rows) > d["max_orbitals"]):
            raise NumericalError("invalid whole-atom partition block")
        atoms.extend(ids)
        orbitals.extend(rows)
    if sorted(atoms) != list(range(len(atom_slices))) or sorted(orbitals) != list(range(n_orbitals)):
        raise NumericalError("partition must cover each atom/orbital exactly once")
    return d


def model_identity(model):
    """The identity a partition file is bound to: the model's operator-cache
    physical key and the hash of its orbital order."""
    cache = (getattr(model, "provenance", None) or {}).get("operator_cache") or {}
    return dict(physical_key=cache.get("physical_key"),
                orbital_order_hash=content_hash(list(getattr(model, "orbital_order", []) or [])))
