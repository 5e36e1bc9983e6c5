"""Model input (kubomsd-model-v1, docs/INPUT_FORMAT.md:5-45 of the reference)
and the results schema (docs/OUTPUTS.md) on the CPU.

Input: files written by hand with NumPy, compared with the matrices read
directly from them: CSR with index_base 1, dense storage under a custom key,
hartree/bohr units, explicit velocities in eV*angstrom (divided by hbar),
positions from orbital coordinates, the composite bond/xbar/delta form, and
the reference's rejections.  Output: a stochastic run's arrays written by
ResultsWriter carry the documented names, shapes ((axis, time, energy);
block sums (block, axis, time, energy)), reason bit values, units and aliases.
"""
import json

import numpy as np
import pytest
import scipy.sparse as sp

from kubomsd_amd.config import RunConfig
from kubomsd_amd.contracts import ModelBundle, RunSummary, TransportBlock
from kubomsd_amd.errors import KuboMSDError
from kubomsd_amd.io.model import load_model, save_model
from kubomsd_amd.io.results import ResultsWriter, read_results
from kubomsd_amd.units import BOHR_A, HARTREE_EV, HBAR_EV_FS

RNG = np.random.default_rng(77)
N = 6


def herm(n=N, complex_=True):
    A = RNG.normal(size=(n, n)) + (1j * RNG.normal(size=(n, n)) if complex_ else 0)
    return (A + A.conj().T) / 2


def write_csr(path, M, base=0):
    m = sp.csr_matrix(M)
    np.savez(path, data=m.data, indices=m.indices.astype(np.int64) + base, indptr=m.indptr.astype(np.int64) + base,
             shape=np.array(m.shape))


def decl(name, storage="csr", base=0, dtype="complex128", **kw):
    return dict(path=name, shape=[N, N], dtype=dtype, storage=storage, index_base=base,
                representation="covariant", **kw)


def model_json(tmp_path, **fields):
    doc = dict(format="kubomsd-model-v1", n_orbitals=N, basis="orthogonal", kpoint=[0, 0, 0], energy_unit="eV",
               length_unit="angstrom", orbital_order=[f"o{i}" for i in range(N)])
    doc.update(fields)
    p = tmp_path / "model.json"
    p.write_text(json.dumps(doc))
    return p


def dense_of(path, base=0):
    """The matrix read directly from a hand-written CSR NPZ file."""
    with np.load(path) as f:
        return sp.csr_matrix((f["data"], f["indices"] - base, f["indptr"] - base), shape=tuple(f["shape"])).toarray()


def test_csr_index_base_one_and_velocity_in_eV_angstrom(tmp_path):
    H, V = herm(), herm()
    write_csr(tmp_path / "H.npz", H, base=1)
    write_csr(tmp_path / "V.npz", V, base=1)
    p = model_json(tmp_path, hamiltonian=decl("H.npz", base=1),
                   velocity_source=dict(type="matrix", unit="eV*angstrom", matrices={"x": decl("V.npz", base=1)}))
    m = load_model(p)
    np.testing.assert_allclose(m.H_cov.toarray(), dense_of(tmp_path / "H.npz", 1), rtol=0, atol=1e-15)
    np.testing.assert_allclose(m.velocity["x"].toarray(), dense_of(tmp_path / "V.npz", 1) / HBAR_EV_FS, rtol=1e-15)
    assert m.basis == "orthogonal" and sp.linalg.norm(m.S - sp.eye(N)) == 0


def test_dense_storage_custom_key_and_hartree_bohr_units(tmp_path):
    H = herm(complex_=False)
    np.savez(tmp_path / "H.npz", ham=H)
    X = herm()
    write_csr(tmp_path / "X.npz", X)
    p = model_json(tmp_path, energy_unit="hartree", length_unit="bohr",
                   hamiltonian=decl("H.npz", storage="dense", dtype="float64", key="ham"),
                   velocity_source=dict(type="position", matrices={"x": decl("X.npz")}))
    m = load_model(p)
    np.testing.assert_allclose(m.H_cov.toarray(), H * HARTREE_EV, rtol=1e-15)
    np.testing.assert_allclose(m.position["x"].toarray(), X * BOHR_A, rtol=1e-15)


def test_position_from_orbital_coordinates(tmp_path):
    H = herm()
    write_csr(tmp_path / "H.npz", H)
    r = RNG.normal(size=(N, 3))
    p = model_json(tmp_path, length_unit="bohr", hamiltonian=decl("H.npz"),
                   velocity_source=dict(type="position", coordinates=r.tolist()))
    m = load_model(p)
    for k, a in enumerate("xyz"):
        np.testing.assert_allclose(m.position[a].toarray(), np.diag(r[:, k] * BOHR_A), rtol=1e-15)


def test_composite_velocity_parts(tmp_path):
    H, S = herm(), np.eye(N) + 0.05 * herm()
    bond, xbar, delta = herm(), herm(), herm()
    for name, M in (("H", H), ("S", S), ("b", bond), ("xb", xbar), ("d", delta)):
        write_csr(tmp_path / f"{name}.npz", M)
    p = model_json(tmp_path, basis="nonorthogonal", energy_unit="rydberg", hamiltonian=decl("H.npz"),
                   overlap=decl("S.npz"),
                   velocity_source=dict(type="composite", axes={"x": dict(bond=decl("b.npz"), xbar=decl("xb.npz"),
                                                                         delta=decl("d.npz"))}))
    m = load_model(p)
    from kubomsd_amd.units import RYDBERG_EV
    b, xb, d = (x.toarray() for x in m.composite["x"])
    np.testing.assert_allclose(b, bond * RYDBERG_EV, rtol=1e-15)
    np.testing.assert_allclose(xb, xbar, rtol=1e-15)
    np.testing.assert_allclose(d, delta, rtol=1e-15)
    np.testing.assert_allclose(m.S.toarray(), S, rtol=1e-15)


@pytest.mark.parametrize("change, match", [
    (dict(index_base=2), "index"),
    (dict(representation="contravariant"), "representation"),
    (dict(shape=[N, N + 1]), "shape"),
])
def test_declaration_rejections(tmp_path, change, match):
    write_csr(tmp_path / "H.npz", herm())
    d = decl("H.npz")
    d.update(change)
    p = model_json(tmp_path, hamiltonian=d, velocity_source=dict(type="position", coordinates=np.zeros((N, 3)).tolist()))
    with pytest.raises(KuboMSDError, match=match):
        load_model(p)


def test_model_rejections(tmp_path):
    write_csr(tmp_path / "H.npz", herm())
    write_csr(tmp_path / "S.npz", np.eye(N) + 0.05 * herm())
    coords = dict(type="position", coordinates=np.zeros((N, 3)).tolist())
    with pytest.raises(KuboMSDError, match="coordinates cannot replace"):
        load_model(model_json(tmp_path, basis="nonorthogonal", hamiltonian=decl("H.npz"), overlap=decl("S.npz"),
                              velocity_source=coords))
    with pytest.raises(KuboMSDError, match="nonidentity S"):
        load_model(model_json(tmp_path, hamiltonian=decl("H.npz"), overlap=decl("S.npz"), velocity_source=coords))
    with pytest.raises(KuboMSDError, match="requires overlap"):
        load_model(model_json(tmp_path, basis="nonorthogonal", hamiltonian=decl("H.npz"), velocity_source=coords))
    with pytest.raises(KuboMSDError, match="velocity unit"):
        load_model(model_json(tmp_path, hamiltonian=decl("H.npz"),
                              velocity_source=dict(type="matrix", matrices={"x": decl("H.npz")})))
    with pytest.raises(KuboMSDError, match="Gamma"):
        load_model(model_json(tmp_path, kpoint=[0, 0, .5], hamiltonian=decl("H.npz"), velocity_source=coords))
    # malformed CSR: pointers not ending at nnz
    m = sp.csr_matrix(herm())
    np.savez(tmp_path / "bad.npz", data=m.data, indices=m.indices, indptr=m.indptr[:-1].tolist() + [m.nnz + 1],
             shape=np.array(m.shape))
    with pytest.raises(KuboMSDError, match="malformed CSR"):
        load_model(model_json(tmp_path, hamiltonian=decl("bad.npz"), velocity_source=coords))


def test_save_model_round_trip(tmp_path):
    H, S, V = herm(), np.eye(N) + 0.05 * herm(), herm()
    p = save_model(ModelBundle(sp.csr_matrix(H), sp.csr_matrix(S), velocity={"x": sp.csr_matrix(V)},
                               atom_slices=[(0, 3), (3, 6)]), tmp_path / "m")
    m = load_model(p)
    np.testing.assert_array_equal(m.H_cov.toarray(), H)
    np.testing.assert_array_equal(m.S.toarray(), S)
    np.testing.assert_array_equal(m.velocity["x"].toarray(), V)
    assert m.basis == "nonorthogonal" and [tuple(s) for s in m.atom_slices] == [(0, 3), (3, 6)]


# ---- results schema (docs/OUTPUTS.md) ----------------------------------------

DOC_REASONS = {"LOW_DOS": 1, "UNCERTAIN_TRACE": 2, "NO_SELECTED_STATES": 4, "TRUNCATION_UNVERIFIED": 8,
               "INTEGRATION_UNVERIFIED": 16, "INSUFFICIENT_BLOCKS": 32}


def stochastic_blocks(cfg, nbk, with_gradient=True):
    """Time slabs with the array names and shapes iter_stochastic_blocks yields."""
    na, ne, nt = len(cfg.axes), len(cfg.energy_eV), len(cfg.time_fs)
    keys = ["msd_A2/commutator"] + (["diffusion_A2_per_fs/msd_gradient"] if with_gradient else [])
    out = []
    for start in range(0, nt, cfg.time_block):
        stop = min(start + cfg.time_block, nt)
        arrays = {}
        for key in keys:
            v = RNG.random((na, stop - start, ne))
            arrays[key] = v
            arrays["numerator_" + key] = v * 2
            arrays["numerator_block_sums/" + key] = RNG.random((nbk, na, stop - start, ne))
            arrays["stderr_" + key] = v * .1
        for est in {k.split("/")[-1] for k in keys}:
            arrays["valid/" + est] = np.ones((na, stop - start, ne), bool)
            arrays["reason_code/" + est] = np.zeros((na, stop - start, ne), np.uint32)
        static = metadata = {}
        if start == 0:
            static = {"dos_states_per_eV": RNG.random(ne), "dos_block_sums": RNG.random((nbk, ne)),
                      "block_counts": np.full(nbk, 2), "sample_count": np.array(2 * nbk),
                      "sample_ids": np.arange(2 * nbk), "stderr_dos_states_per_eV": RNG.random(ne)}
            metadata = {"primary_estimator": "commutator", "primary_diffusion_estimator": "msd_gradient"}
        out.append(TransportBlock(start, cfg.time_fs[start:stop], arrays, static, metadata))
    return out


def test_results_schema_follows_outputs_doc(tmp_path):
    cfg = RunConfig(energy_eV=np.linspace(-1, 1, 7), time_fs=np.arange(5) * .5, solver="stochastic-kpm",
                    kernel="jackson", backend="hip", axes=("x", "y"), nsamples=6, statistics_block=2, time_block=2)
    with ResultsWriter(tmp_path / "r", cfg) as w:
        for b in stochastic_blocks(cfg, 3):
            w.append(b)
        w.finalize(RunSummary(str(tmp_path / "r"), "complete", "stepwise", "hip"))
    out = read_results(tmp_path / "r")
    a, meta = out["arrays"], out["metadata"]
    na, nt, ne = 2, 5, 7
    for key in ("msd_A2/commutator", "diffusion_A2_per_fs/msd_gradient", "numerator_msd_A2/commutator",
                "stderr_msd_A2/commutator", "valid/commutator", "reason_code/commutator", "valid/msd_gradient"):
        assert a[key].shape == (na, nt, ne), key  # (axis, time, energy)
    assert a["numerator_block_sums/msd_A2/commutator"].shape == (3, na, nt, ne)  # (block, axis, time, energy)
    assert a["dos_states_per_eV"].shape == (ne,) and a["dos_block_sums"].shape == (3, ne)
    np.testing.assert_array_equal(a["energy_eV"], cfg.energy_eV)
    np.testing.assert_array_equal(a["time_fs"], cfg.time_fs)
    for axis in "xy":  # aliases of the primary estimators
        assert a[f"msd_{axis}_A2"].shape == (nt, ne) and a[f"diffusion_{axis}_A2_per_fs"].shape == (nt, ne)
    np.testing.assert_array_equal(a["msd_direction_average_A2"], a["msd_A2/commutator"].mean(0))
    assert meta["legacy_aliases"]["diffusion_x_A2_per_fs"] == "diffusion_A2_per_fs/msd_gradient"
    assert meta["reason_codes"] == DOC_REASONS
    assert meta["units"] == {"energy": "eV", "time": "fs", "msd": "angstrom^2", "diffusion": "angstrom^2/fs",
                             "vac": "angstrom^2/fs^2", "dos": "states/eV/system"}
    assert meta["spin_degeneracy_multiplier"] == 1 and meta["run_status"] == "complete"
    assert meta["axis_order"] == "axis,time,energy; statistical sums: block,axis,time,energy"
    assert str(a["run_status"]) == "complete" and str(a["primary_estimator"]) == "commutator"
    assert json.loads((tmp_path / "r" / "metadata.json").read_text())["run_status"] == "complete"


def test_writer_derives_msd_gradient_like_the_reference(tmp_path):
    """Without an estimator-supplied gradient the writer adds
    diffusion = 1/2 d(MSD)/dt on the output times (numpy.gradient, second-order
    edges), its numerator and masks (reference io/results.py:56-72)."""
    cfg = RunConfig(energy_eV=np.linspace(-1, 1, 4), time_fs=[0, .3, .5, 1.2, 1.3], solver="stochastic-kpm",
                    kernel="jackson", backend="hip", axes=("x",), time_block=2)
    blocks = stochastic_blocks(cfg, 2, with_gradient=False)
    for b in blocks:
        b.metadata.pop("primary_diffusion_estimator", None)
    with ResultsWriter(tmp_path / "r", cfg) as w:
        for b in blocks:
            w.append(b)
        w.finalize(RunSummary("r", "complete", "stepwise", "hip"))
    out = read_results(tmp_path / "r")
    a = out["arrays"]
    msd = a["msd_A2/commutator"]
    np.testing.assert_array_equal(a["diffusion_A2_per_fs/msd_gradient"],
                                  .5 * np.gradient(msd, np.asarray(cfg.time_fs), axis=1, edge_order=2))
    np.testing.assert_array_equal(a["valid/msd_gradient"], a["valid/commutator"])
    assert out["metadata"]["msd_gradient"]["edge_order"] == 2


def test_writer_rejects_wrong_axes_energies_and_duplicate_static(tmp_path):
    cfg = RunConfig(energy_eV=np.linspace(-1, 1, 4), time_fs=[0, .5], solver="stochastic-kpm", kernel="jackson",
                    backend="hip", axes=("x",))
    b = stochastic_blocks(cfg, 2)[0]
    bad = TransportBlock(0, b.time_fs, {"msd_A2/commutator": np.zeros((2, 2, 4))}, b.static, b.metadata)
    with pytest.raises(KuboMSDError, match="invalid shape"):
        with ResultsWriter(tmp_path / "r", cfg) as w:
            w.append(bad)
    cfg2 = RunConfig(energy_eV=np.linspace(-1, 1, 4), time_fs=[0, .5], solver="stochastic-kpm", kernel="jackson",
                     backend="hip", axes=("x",))
    b2 = stochastic_blocks(cfg2, 2)[0]
    b2.static["energy_eV"] = np.zeros(4)
    with pytest.raises(KuboMSDError, match="duplicate static"):
        with ResultsWriter(tmp_path / "r2", cfg2) as w:
            w.append(b2)
    with pytest.raises(KuboMSDError, match="incomplete"):
        with ResultsWriter(tmp_path / "r3", cfg2) as w:
            w.append(stochastic_blocks(cfg2, 2)[0])
            w.finalize(RunSummary("r3", "failed", "stepwise", "hip"))
