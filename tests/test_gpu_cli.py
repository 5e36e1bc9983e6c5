"""`python -m kubomsd_amd run` end to end on the reference's example models
(examples/two_level, examples/nonorthogonal; arguments as in the reference's
kubo_msd_gpu/tests/integration/test_gpu_cli_pipeline.py:6-15).  The written
arrays must equal a direct library run with the same configuration, bit for
bit (every device reduction is deterministic)."""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXAMPLES = Path(__file__).parent / "golden" / "examples"
ARGS = ["--solver", "stochastic-kpm", "--kernel", "jackson", "--spectral-min", "-4", "--spectral-max", "4",
        "--emin", "-.2", "--emax", ".2", "--n-energy", "5", "--tmax-fs", ".05", "--dt-fs", ".025",
        "--n-moments", "12", "--nsamples", "3", "--statistics-block", "2", "--no-vac", "--backend", "gpu"]


@pytest.mark.parametrize("example, precision", [("nonorthogonal", "fp64"), ("two_level", "fp64"),
                                                ("nonorthogonal", "mixed")])
def test_cli_run_matches_library(hip_ctx, tmp_path, example, precision):
    from kubomsd_amd.cli import config_from_args, main, parser
    from kubomsd_amd.io.model import load_model
    from kubomsd_amd.io.results import read_results
    from kubomsd_amd.operators.velocity import build_operators
    from kubomsd_amd.solvers.stochastic import iter_stochastic_blocks
    inp = EXAMPLES / example / "model.json"
    out = tmp_path / "run"
    argv = ["run", "--input", str(inp), "--out", str(out), "--precision", precision] + ARGS
    print("cli", example, precision, flush=True)
    assert main(argv) == 0
    meta = json.loads((out / "metadata.json").read_text())
    assert meta["run_status"] == "complete" and meta["primary_estimator"] == "commutator"
    res = read_results(out)["arrays"]
    cfg = config_from_args(parser().parse_args(argv))
    ops = build_operators(load_model(inp), cfg, hip_ctx)
    blocks = list(iter_stochastic_blocks(ops, cfg, "stepwise"))
    np.testing.assert_array_equal(res["dos_states_per_eV"], blocks[0].static["dos_states_per_eV"])
    want = np.concatenate([b.arrays["numerator_msd_A2/commutator"] for b in blocks], axis=1)
    np.testing.assert_array_equal(res["numerator_msd_A2/commutator"], want)
    assert np.isfinite(res["dos_states_per_eV"]).all()
    assert res["msd_x_A2"].shape == (len(cfg.time_fs), 5)
    assert np.all(res["numerator_msd_A2/commutator"][:, 0] == 0)  # MSD(t=0) = 0
