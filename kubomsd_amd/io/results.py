"""Results writer and reader with the reference's schema names
(src/kubo_msd/io/results.py:20-130, docs/OUTPUTS.md), stored as NPZ because
h5py is not available: results.partial.npz + metadata.json while running,
renamed to results.npz with run_status 'complete' only by finalize().
Arrays keep the reference's names and (axis, time, energy) order; aliases
msd_<axis>_A2 / diffusion_<axis>_A2_per_fs and direction averages are added."""
from dataclasses import asdict, is_dataclass
import json
import os
from pathlib import Path

import numpy as np

from ..contracts import Reason
from ..errors import KuboMSDError
from ..units import HBAR_EV_FS


def json_safe(v):
    if is_dataclass(v):
        return json_safe(asdict(v))
    if isinstance(v, dict):
        return {str(k): json_safe(x) for k, x in v.items()}
    if isinstance(v, (list, tuple, np.ndarray)):
        return [json_safe(x) for x in v]
    if isinstance(v, np.generic):
        return json_safe(v.item())
    if isinstance(v, float) and not np.isfinite(v):
        return None
    if isinstance(v, Path):
        return str(v)
    return v


class ResultsWriter:
    def __init__(self, path, config):
        self.path = Path(path).resolve()
        self.path.mkdir(parents=True, exist_ok=False)
        self.cfg = config
        self.next = 0
        self.keys = None
        self.done = False
        self.arrays = {"energy_eV": np.asarray(config.energy_eV), "time_fs": np.asarray(config.time_fs),
                       "axes_vectors": np.asarray(config.axes_vectors)}
        self.metadata = {"schema_version": "1.0", "run_status": "partial", "resolved_config": config.as_dict(),
                         "axis_order": "axis,time,energy; statistical sums: block,axis,time,energy",
                         "units": {"energy": "eV", "time": "fs", "msd": "angstrom^2", "diffusion": "angstrom^2/fs",
                                   "vac": "angstrom^2/fs^2", "dos": "states/eV/system"},
                         "spin_degeneracy_multiplier": 1, "reason_codes": {e.name: int(e) for e in Reason},
                         "hbar_eV_fs": HBAR_EV_FS, "container": "npz"}
        self._save_meta()

    def _save_meta(self):
        tmp = self.path / "metadata.json.tmp"
        tmp.write_text(json.dumps(json_safe(self.metadata), ensure_ascii=False, indent=2, allow_nan=False) + "\n")
        os.replace(tmp, self.path / "metadata.json")

    def __enter__(self):
        return self

    def __exit__(self, typ, value, tb):
        if not self.done:
            self.metadata["run_status"] = "failed" if typ else "partial"
            self.metadata["written_time_points"] = self.next
            if typ:
                self.metadata["failure"] = {"type": typ.__name__, "message": str(value)}
                if hasattr(value, "diagnostics"):
                    self.metadata["failure"]["numerics"] = value.diagnostics
            np.savez(self.path / "results.partial.npz", **self.arrays)
            self._save_meta()

    def append(self, block):
        if self.done:
            raise KuboMSDError("cannot append finalized results")
        t = np.asarray(block.time_fs)
        end = block.start + len(t)
        if block.start != self.next or end > len(self.cfg.time_fs) or \
                not np.array_equal(t, self.cfg.time_fs[block.start:end]):
            raise KuboMSDError("TransportBlock must contain contiguous configured times")
        if not len(t) or not block.arrays:
            raise KuboMSDError("empty TransportBlock")
        keys = set(block.arrays)
        if self.keys is not None and keys != self.keys:
            raise KuboMSDError("estimator schema changed during append")
        nt = len(self.cfg.time_fs)
        for key, value in block.arrays.items():
            a = np.asarray(value)
            axis = 2 if key.startswith("numerator_block_sums/") else 1
            if (a.ndim != (4 if axis == 2 else 3) or a.shape[axis] != len(t) or a.shape[-1] != len(self.cfg.energy_eV)
                    or a.shape[axis - 1] != len(self.cfg.axes)):
                raise KuboMSDError(f"invalid shape for {key}: {a.shape}")
            if key not in self.arrays:
                shape = list(a.shape)
                shape[axis] = nt
                self.arrays[key] = np.full(shape, np.nan if a.dtype.kind == "f" else 0, dtype=a.dtype)
            index = [slice(None)] * a.ndim
            index[axis] = slice(block.start, end)
            self.arrays[key][tuple(index)] = a
        if self.next and (block.static or block.metadata):
            raise KuboMSDError("static fields/metadata are allowed only in first block")
        for key, value in block.static.items():
            if key in self.arrays:
                raise KuboMSDError("duplicate static field " + key)
            self.arrays[key] = np.asarray(value)
        self.metadata.update(json_safe(block.metadata))
        self.keys = keys
        self.next = end
        self.metadata["written_time_points"] = self.next
        self._save_meta()

    def _derive_gradient(self):
        """diffusion_A2_per_fs/msd_gradient = 1/2 d/dt of the primary MSD on the
        output times when the estimator did not supply it (with its numerator
        and masks), as the reference's writer does (io/results.py:56-72);
        numpy.gradient over the whole time axis equals the reference's
        per-time-block evaluation with one-point halos."""
        primary = self.metadata["primary_estimator"]
        key = "diffusion_A2_per_fs/msd_gradient"
        nt = len(self.cfg.time_fs)
        if nt < 2 or key in self.arrays or "msd_A2/" + primary not in self.arrays:
            return
        edge = 2 if nt > 2 else 1
        t = np.asarray(self.cfg.time_fs)
        self.arrays[key] = .5 * np.gradient(self.arrays["msd_A2/" + primary], t, axis=1, edge_order=edge)
        num = self.arrays.get("numerator_msd_A2/" + primary)
        if num is not None:
            self.arrays["numerator_diffusion_A2_per_fs/msd_gradient"] = .5 * np.gradient(num, t, axis=1,
                                                                                        edge_order=edge)
        for prefix in ("valid/", "reason_code/"):
            if prefix + primary in self.arrays:
                self.arrays[prefix + "msd_gradient"] = self.arrays[prefix + primary].copy()
        self.metadata["msd_gradient"] = {"source": primary, "edge_order": edge, "time_array": "actual output time_fs"}

    def finalize(self, summary):
        status = getattr(summary, "run_status", "complete")
        if self.done or self.next != len(self.cfg.time_fs) or status != "complete":
            raise KuboMSDError("cannot mark incomplete computation complete")
        if "dos_states_per_eV" not in self.arrays or "primary_estimator" not in self.metadata:
            raise KuboMSDError("required result metadata missing")
        total = sum(np.asarray(a).nbytes for a in self.arrays.values())
        if total > self.cfg.memory_limit_gb * 1024 ** 3 * .5:
            raise KuboMSDError("results exceed half the RAM budget (memory_limit_gb) for the NPZ container")
        self._derive_gradient()
        primary = self.metadata["primary_estimator"]
        dp = self.metadata.get("primary_diffusion_estimator", primary)
        aliases = {}
        for quantity, unit, est in (("msd", "A2", primary), ("diffusion", "A2_per_fs", dp)):
            key = f"{quantity}_{unit}/{est}"
            if key not in self.arrays:
                continue
            data = self.arrays[key]
            for i, axis in enumerate(self.cfg.axes):
                if axis in ("x", "y", "z"):
                    self.arrays[f"{quantity}_{axis}_{unit}"] = data[i]
                    aliases[f"{quantity}_{axis}_{unit}"] = key
            self.arrays[f"{quantity}_direction_average_{unit}"] = np.mean(data, axis=0)
        self.metadata["legacy_aliases"] = aliases
        self.metadata["direction_average"] = ("arithmetic mean over declared orthonormal directions; no volume "
                                              "normalization")
        self.metadata["summary"] = json_safe(summary)
        self.metadata["run_status"] = "complete"
        arrays = dict(self.arrays)
        arrays["schema_version"] = np.array("1.0")
        arrays["run_status"] = np.array("complete")
        arrays["primary_estimator"] = np.array(primary)
        arrays["metadata_json"] = np.array(json.dumps(json_safe(self.metadata), allow_nan=False))
        tmp = self.path / "results.partial.npz"
        np.savez_compressed(tmp, **arrays)
        os.replace(tmp, self.path / "results.npz")
        self._save_meta()
        self.done = True


def read_results(path, allow_partial=False):
    p = Path(path)
    if p.is_dir():
        p = p / ("results.npz" if (p / "results.npz").exists() else "results.partial.npz")
    with np.load(p, allow_pickle=False) as f:
        if "schema_version" not in f or str(f["schema_version"]) != "1.0" or "metadata_json" not in f:
            if not allow_partial:
                raise KuboMSDError("result is failed/partial; use allow_partial explicitly")
            return {"metadata": json.loads((p.parent / "metadata.json").read_text()),
                    "arrays": {k: np.array(f[k]) for k in f.files}}
        metadata = json.loads(str(f["metadata_json"]))
        arrays = {k: np.array(f[k]) for k in f.files if k != "metadata_json"}
    if metadata.get("run_status") != "complete" and not allow_partial:
        raise KuboMSDError("result is failed/partial; use allow_partial explicitly")
    return {"metadata": metadata, "arrays": arrays}
