"""CLI argument mapping and the NPZ results writer (CPU only).

Mirrors the reference's writer contracts (src/kubo_msd/io/results.py:20-130;
kubo_msd_gpu/tests/gpu/test_failure_contracts.py:46-62): partial until
finalize, 'failed' with the exception on error, contiguous time blocks only,
aliases msd_<axis>_A2 after finalize."""
import json
from pathlib import Path

import numpy as np
import pytest

from kubomsd_amd.cli import config_from_args, main, parser
from kubomsd_amd.config import RunConfig
from kubomsd_amd.contracts import TransportBlock
from kubomsd_amd.errors import KuboMSDError
from kubomsd_amd.io.results import ResultsWriter, read_results

EXAMPLES = Path(__file__).parent / "golden" / "examples"


def _cfg(**kw):
    base = dict(energy_eV=np.linspace(-1, 1, 5), time_fs=[0, .1, .2, .3], solver="stochastic-kpm", kernel="jackson",
                backend="hip", axes=("x", "y"), nsamples=4, statistics_block=2)
    base.update(kw)
    return RunConfig(**base)


def _block(cfg, start, stop, rng, first=True):
    na, ne, nt = len(cfg.axes), len(cfg.energy_eV), stop - start
    arrays = {"msd_A2/commutator": rng.random((na, nt, ne)),
              "numerator_msd_A2/commutator": rng.random((na, nt, ne)),
              "numerator_block_sums/msd_A2/commutator": rng.random((2, na, nt, ne)),
              "valid/commutator": np.ones((na, nt, ne), bool),
              "reason_code/commutator": np.zeros((na, nt, ne), np.uint32)}
    static = {"dos_states_per_eV": rng.random(ne)} if first else {}
    meta = {"primary_estimator": "commutator"} if first else {}
    return TransportBlock(start, cfg.time_fs[start:stop], arrays, static, meta)


def test_parser_maps_reference_flags():
    args = parser().parse_args(["run", "--input", "m.json", "--out", "o", "--solver", "stochastic-kpm", "--backend",
                                "gpu", "--kernel", "jackson", "--emin", "-1", "--emax", "1", "--n-energy", "3",
                                "--tmax-fs", ".2", "--dt-fs", ".1", "--spectral-min", "-4", "--spectral-max", "4",
                                "--axes", "x,y", "--no-vac", "--precision", "mixed", "--device-memory-limit-gb", "2"])
    cfg = config_from_args(args)
    assert cfg.backend == "hip" and cfg.solver == "stochastic-kpm" and cfg.precision == "mixed"
    assert cfg.spectral_bounds == (-4.0, 4.0) and cfg.axes == ("x", "y") and not cfg.include_vac
    np.testing.assert_allclose(cfg.energy_eV, [-1, 0, 1])
    np.testing.assert_allclose(cfg.time_fs, [0, .1, .2])
    assert cfg.device_memory_limit_gb == 2.0


def test_block_width_autotune_flag_and_validation():
    from kubomsd_amd.config import RunConfig
    from kubomsd_amd.errors import KuboMSDError
    base = ["run", "--input", "m.json", "--out", "o", "--emin", "-1", "--emax", "1", "--n-energy", "3",
            "--tmax-fs", ".1", "--spectral-min", "-4", "--spectral-max", "4"]
    assert not config_from_args(parser().parse_args(base)).block_width_autotune
    assert config_from_args(parser().parse_args(base + ["--block-width-autotune"])).block_width_autotune
    with pytest.raises(KuboMSDError, match="block_width_autotune"):
        RunConfig(energy_eV=[0.0], time_fs=[0.0], block_width_autotune="yes")


@pytest.mark.parametrize("extra, match", [(["--spectral-min", "-4"], "spectral"), ([], "emin")])
def test_parser_rejects_incomplete(extra, match, tmp_path):
    a = ["run", "--input", "m.json", "--out", "o"] + extra
    if match != "emin":
        a += ["--emin", "-1", "--emax", "1"]
    with pytest.raises(ValueError, match=match):
        config_from_args(parser().parse_args(a))


def test_cli_refuses_exact_solver_without_output(tmp_path):
    out = tmp_path / "out"
    rc = main(["run", "--input", str(EXAMPLES / "two_level" / "model.json"), "--out", str(out), "--emin", "-1",
               "--emax", "1", "--n-energy", "3", "--tmax-fs", ".01"])
    assert rc == 2 and not out.exists()


def test_writer_roundtrip_and_aliases(tmp_path):
    cfg = _cfg()
    rng = np.random.default_rng(3)
    b0, b1 = _block(cfg, 0, 2, rng), _block(cfg, 2, 4, rng, first=False)
    with ResultsWriter(tmp_path / "r", cfg) as w:
        w.append(b0)
        assert json.loads((tmp_path / "r" / "metadata.json").read_text())["run_status"] == "partial"
        w.append(b1)
        w.finalize({"ok": True})
    out = read_results(tmp_path / "r")
    assert out["metadata"]["run_status"] == "complete"
    a = out["arrays"]
    full = np.concatenate([b0.arrays["msd_A2/commutator"], b1.arrays["msd_A2/commutator"]], axis=1)
    np.testing.assert_array_equal(a["msd_A2/commutator"], full)
    np.testing.assert_array_equal(a["msd_x_A2"], full[0])
    np.testing.assert_array_equal(a["msd_direction_average_A2"], full.mean(0))
    sums = np.concatenate([b0.arrays["numerator_block_sums/msd_A2/commutator"],
                           b1.arrays["numerator_block_sums/msd_A2/commutator"]], axis=2)
    np.testing.assert_array_equal(a["numerator_block_sums/msd_A2/commutator"], sums)
    assert out["metadata"]["legacy_aliases"]["msd_y_A2"] == "msd_A2/commutator"


def test_writer_rejects_gaps_and_marks_failure(tmp_path):
    cfg = _cfg()
    rng = np.random.default_rng(4)
    with pytest.raises(RuntimeError):
        with ResultsWriter(tmp_path / "r", cfg) as w:
            with pytest.raises(KuboMSDError, match="contiguous"):
                w.append(_block(cfg, 2, 4, rng))
            w.append(_block(cfg, 0, 2, rng))
            with pytest.raises(KuboMSDError, match="incomplete"):
                w.finalize({})
            raise RuntimeError("injected device failure")
    meta = json.loads((tmp_path / "r" / "metadata.json").read_text())
    assert meta["run_status"] == "failed" and meta["written_time_points"] == 2
    assert "injected" in meta["failure"]["message"]
    with pytest.raises(KuboMSDError, match="partial"):
        read_results(tmp_path / "r")
    part = read_results(tmp_path / "r", allow_partial=True)
    assert np.isnan(part["arrays"]["msd_A2/commutator"][:, 2:]).all()


def test_writer_refuses_existing_directory(tmp_path):
    (tmp_path / "r").mkdir()
    with pytest.raises(FileExistsError):
        ResultsWriter(tmp_path / "r", _cfg())
