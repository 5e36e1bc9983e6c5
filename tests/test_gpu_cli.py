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


@pytest.mark.parametrize("example, precision, tol", [("nonorthogonal", "fp64", 1e-9), ("two_level", "fp64", 1e-9),
                                                     ("nonorthogonal", "mixed", 3e-7)])
def test_cli_run_matches_oracle(hip_ctx, tmp_path, example, precision, tol):
    """The CLI's block sums equal the FP64 oracle's on the same samples: the
    reference's phases (seed, ids 0..2), the same metric sampling polynomial
    applied by the oracle's Clenshaw sampler (bit-identical to the reference
    kernel), the oracle's stepwise estimator, and sums per statistics block."""
    import scipy.sparse as sp
    from kubomsd_amd.cli import config_from_args, main, parser
    from kubomsd_amd.io.model import load_model
    from kubomsd_amd.io.results import read_results
    from kubomsd_amd.metric.bounds import spd_bounds
    from kubomsd_amd.metric.sampling import inverse_sqrt_series, random_phases
    from oracle import kpm
    inp = EXAMPLES / example / "model.json"
    out = tmp_path / "run"
    argv = ["run", "--input", str(inp), "--out", str(out), "--precision", precision] + ARGS
    assert main(argv) == 0
    res = read_results(out)["arrays"]
    cfg = config_from_args(parser().parse_args(argv))
    model = load_model(inp)
    S = sp.csr_matrix(model.S)
    ops = kpm.Ops(model.H_cov, S, model.velocity, rtol=1e-13, position=model.position)
    z = 1 / np.sqrt(S.diagonal().real)
    lo, hi, _ = spd_bounds(sp.diags(z) @ S @ sp.diags(z))
    center, half = (hi + lo) / 2, (hi - lo) / 2
    coeff, _ = inverse_sqrt_series(center, half, cfg.sampling_tol, cfg.sampling_max_order)
    B = ops.metric.sample(random_phases(S.shape[0], cfg.seed, range(cfg.nsamples)), coeff, center, half)
    kcfg = kpm.config(energy_eV=cfg.energy_eV, time_fs=cfg.time_fs, spectral_bounds=cfg.spectral_bounds,
                      n_moments=cfg.n_moments, segment_fs=cfg.segment_fs, axes_vectors=cfg.axes_vectors)
    dos, nums = kpm.stepwise_samples(ops, kcfg, B)
    blocks = [slice(0, 2), slice(2, 3)]
    want_dos = np.array([dos[s].sum(0) for s in blocks])
    want = np.array([nums["msd_A2/commutator"][s].sum(0) for s in blocks])
    got = res["numerator_block_sums/msd_A2/commutator"]
    assert got.shape == want.shape
    assert np.max(np.abs(res["dos_block_sums"] - want_dos)) <= tol * np.max(np.abs(want_dos))
    assert np.max(np.abs(got - want)) <= tol * np.max(np.abs(want))
