"""kubomsd-model-v1 input reader (reference src/kubo_msd/io/model.py:11-77,
docs/INPUT_FORMAT.md).  Same schema, units and validation; NPZ containers
(dense or CSR) are read with allow_pickle=False.  HDF5 containers need h5py,
which this image lacks: they are rejected with a message naming it."""
import json
from pathlib import Path

import numpy as np
import scipy.sparse as sp

from ..contracts import ModelBundle
from ..errors import KuboMSDError
from ..operators.velocity import hermitian
from ..units import BOHR_A, HARTREE_EV, HBAR_EV_FS, RYDBERG_EV

_REQUIRED = {"path", "shape", "dtype", "storage", "index_base", "representation"}


def read_matrix(entry, root, n):
    if not _REQUIRED <= entry.keys():
        raise KuboMSDError(f"matrix declaration missing {sorted(_REQUIRED - entry.keys())}")
    if (entry["shape"] != [n, n] or entry["index_base"] not in (0, 1) or entry["representation"] != "covariant"
            or entry["dtype"] not in ("float64", "complex128")):
        raise KuboMSDError("invalid matrix shape/index/representation/dtype")
    path = (root / entry["path"]).resolve()
    if not path.is_file():
        raise KuboMSDError(f"missing matrix: {path}")
    storage = entry["storage"]
    if path.suffix == ".npz":
        with np.load(path, allow_pickle=False) as f:
            if storage == "dense":
                A = np.array(f[entry.get("key", "matrix")])
            elif storage == "csr":
                shape = tuple(f["shape"])
                idx, ptr, base = f["indices"], f["indptr"], entry["index_base"]
                if (shape != (n, n) or idx.dtype.kind not in "iu" or ptr.dtype.kind not in "iu" or len(ptr) != n + 1
                        or ptr[0] != base or ptr[-1] - base != len(idx) or len(f["data"]) != len(idx)
                        or np.any(np.diff(ptr) < 0) or np.any(idx < base) or np.any(idx >= n + base)):
                    raise KuboMSDError("malformed CSR arrays")
                A = sp.csr_matrix((f["data"], idx - base, ptr - base), shape=shape)
            else:
                raise KuboMSDError("matrix storage must be dense/csr")
    elif path.suffix in (".h5", ".hdf5"):
        try:
            import h5py  # noqa: F401
        except ImportError as exc:
            raise KuboMSDError("HDF5 matrix containers need h5py, which is not installed; use NPZ") from exc
        raise KuboMSDError("HDF5 matrix containers are not read by this backend; use NPZ")
    else:
        raise KuboMSDError("matrix container must be NPZ/HDF5")
    if A.shape != (n, n) or str(A.dtype) != entry["dtype"]:
        raise KuboMSDError("matrix shape/dtype disagrees with declaration")
    return sp.csr_matrix(A, dtype=complex)


def load_model(path):
    path = Path(path).resolve()
    if not path.is_file():
        raise KuboMSDError(f"missing model: {path}")
    d = json.loads(path.read_text())
    root = path.parent
    if d.get("format") != "kubomsd-model-v1":
        raise KuboMSDError("unknown input schema (this backend reads kubomsd-model-v1)")
    for key in ("n_orbitals", "basis", "kpoint", "energy_unit", "length_unit", "orbital_order", "hamiltonian",
                "velocity_source"):
        if key not in d:
            raise KuboMSDError("missing model field " + key)
    n, order = d["n_orbitals"], d["orbital_order"]
    if not isinstance(n, int) or n < 1 or len(order) != n or len(set(map(str, order))) != n:
        raise KuboMSDError("invalid orbital count/order")
    if d["kpoint"] != [0, 0, 0]:
        raise KuboMSDError("v1 supports only one Gamma point")
    if d["basis"] not in ("orthogonal", "nonorthogonal"):
        raise KuboMSDError("basis must be explicit")
    ef = {"eV": 1., "hartree": HARTREE_EV, "rydberg": RYDBERG_EV}.get(d["energy_unit"])
    lf = {"angstrom": 1., "bohr": BOHR_A}.get(d["length_unit"])
    if ef is None or lf is None:
        raise KuboMSDError("unsupported energy/length units")
    H = hermitian(read_matrix(d["hamiltonian"], root, n) * ef, "Hamiltonian")
    if "overlap" in d:
        S = hermitian(read_matrix(d["overlap"], root, n), "overlap")
    elif d["basis"] == "orthogonal":
        S = sp.eye(n, dtype=complex, format="csr")
    else:
        raise KuboMSDError("nonorthogonal model requires overlap")
    if d["basis"] == "orthogonal" and sp.linalg.norm(S - sp.eye(n)) > 1e-12:
        raise KuboMSDError("orthogonal model supplied nonidentity S")
    source = d["velocity_source"]
    kind = source.get("type")
    velocity, position, composite = {}, {}, {}
    if kind == "matrix":
        factor = {"angstrom/fs": 1., "eV*angstrom": 1 / HBAR_EV_FS}.get(source.get("unit"))
        if factor is None:
            raise KuboMSDError("explicit physical velocity unit required")
        velocity = {a: hermitian(read_matrix(e, root, n) * factor, "velocity " + a)
                    for a, e in source["matrices"].items()}
    elif kind == "position":
        if "coordinates" in source:
            if d["basis"] != "orthogonal":
                raise KuboMSDError("atomic coordinates cannot replace nonorthogonal dipole matrices")
            r = np.asarray(source["coordinates"], float) * lf
            if r.shape != (n, 3) or not np.isfinite(r).all():
                raise KuboMSDError("invalid orbital positions")
            position = {a: sp.diags(r[:, k], format="csr").astype(complex) for k, a in enumerate("xyz")}
        else:
            position = {a: hermitian(read_matrix(e, root, n) * lf, "position " + a)
                        for a, e in source["matrices"].items()}
    elif kind == "composite":
        for a, parts in source["axes"].items():
            composite[a] = (read_matrix(parts["bond"], root, n) * ef * lf, read_matrix(parts["xbar"], root, n) * lf,
                            read_matrix(parts["delta"], root, n) * lf)
    else:
        raise KuboMSDError("unknown velocity source")
    return ModelBundle(H, S, velocity, position, composite, order, d.get("atom_slices", []),
                       {"input_format": d["format"], "internal_units": ["eV", "fs", "angstrom"],
                        "velocity_source": kind}, d["basis"])


def save_model(model, target, velocity_unit="angstrom/fs"):
    """Writes a kubomsd-model-v1 directory (CSR NPZ matrices) for a ModelBundle."""
    target = Path(target)
    target.mkdir(parents=True, exist_ok=True)
    n = model.n_orbitals

    def entry(name, matrix):
        m = sp.csr_matrix(matrix, dtype=np.complex128)
        np.savez(target / f"{name}.npz", data=m.data, indices=m.indices.astype(np.int64),
                 indptr=m.indptr.astype(np.int64), shape=np.array(m.shape))
        return dict(path=f"{name}.npz", shape=[n, n], dtype="complex128", storage="csr", index_base=0,
                    representation="covariant")
    doc = dict(format="kubomsd-model-v1", n_orbitals=n,
               basis="orthogonal" if sp.linalg.norm(sp.csr_matrix(model.S) - sp.eye(n)) == 0 else "nonorthogonal",
               kpoint=[0, 0, 0], energy_unit="eV", length_unit="angstrom", orbital_order=list(range(n)),
               hamiltonian=entry("H", model.H_cov), atom_slices=[list(s) for s in model.atom_slices],
               velocity_source=dict(type="matrix", unit=velocity_unit,
                                    matrices={a: entry("V" + a, v) for a, v in model.velocity.items()}))
    if doc["basis"] == "nonorthogonal":
        doc["overlap"] = entry("S", model.S)
    (target / "model.json").write_text(json.dumps(doc, indent=2) + "\n")
    return target / "model.json"
