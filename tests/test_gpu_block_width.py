"""Vector-block width tuning (RunConfig.block_width_autotune): the stochastic
driver times one fused Chebyshev step of H at the admitted sample batch and
its doublings (kmsd_time_step) and runs the width with the least time per
column.  The width changes only the launch shape, so every per-sample result
matches the fixed-width run within the rounding of the moment reductions
(rtol 1e-10, as the batch-invariance test), and the block sums equal the FP64
oracle's."""
import numpy as np
import pytest

from helpers import phases
from kubomsd_amd.synthetic import square_lattice
from oracle import kpm

pytestmark = pytest.mark.gpu


def run_config(**kw):
    from kubomsd_amd.config import RunConfig
    base = dict(solver="stochastic-kpm", kernel="jackson", backend="hip", metric_rtol=1e-13,
                include_vac=False, segment_fs=1.0)
    base.update(kw)
    return RunConfig(**base)


def test_time_step_and_width_choice(hip_ctx):
    import torch
    from kubomsd_amd.backends.hip import HipMatrix
    m = HipMatrix(hip_ctx, square_lattice(64, seed=3).H)
    t8 = hip_ctx.time_step(m, 8, torch.complex64)
    assert t8 > 0
    best, us = hip_ctx.tune_block_width(m, [8, 16, 32], torch.complex64)
    assert best in (8, 16, 32) and sorted(us) == [8, 16, 32] and all(v > 0 for v in us.values())
    # the rule: a wider block wins only when more than 3 % faster per column
    for w in (16, 32):
        if w > best:
            assert us[w] / w >= 0.97 * us[best] / best
    # remembered for the same signature
    assert hip_ctx.tune_block_width(m, [32, 8, 16], torch.complex64)[0] == best


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tuned_width_keeps_results(hip_ctx, precision):
    from kubomsd_amd.contracts import ModelBundle
    from kubomsd_amd.operators.velocity import build_operators
    from kubomsd_amd.solvers.stochastic import iter_stochastic_blocks
    m = square_lattice(16, seed=9)
    E = np.linspace(-4, 4, 17)
    B = phases(m.n, 16, 8)
    out = []
    for tune in (False, True):
        cfg = run_config(energy_eV=E, time_fs=[0, .5, 1.0], spectral_bounds=(-5.1, 5.1), n_moments=64,
                         axes=("x",), nsamples=16, statistics_block=4, sample_batch=4, precision=precision,
                         block_width_autotune=tune)
        ops = build_operators(ModelBundle(m.H, m.S, velocity=m.velocity), cfg, hip_ctx)
        out.append(list(iter_stochastic_blocks(ops, cfg, "stepwise", samples=B))[0])
    fixed, tuned = out
    assert fixed.metadata["admitted_sample_batch"] == 4 and "block_width_tuning" not in fixed.metadata
    info = tuned.metadata["block_width_tuning"]
    assert info["chosen"] == tuned.metadata["admitted_sample_batch"] in (4, 8, 16)
    assert sorted(int(k) for k in info["us_per_step"]) == [4, 8, 16]
    for key in ("msd_A2/commutator", "numerator_block_sums/msd_A2/commutator"):
        np.testing.assert_allclose(tuned.arrays[key], fixed.arrays[key], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(tuned.static["dos_block_sums"], fixed.static["dos_block_sums"], rtol=1e-10)
    if precision == "fp64":
        kcfg = kpm.config(energy_eV=E, time_fs=[0, .5, 1.0], spectral_bounds=(-5.1, 5.1), n_moments=64,
                          segment_fs=1.0, axes_vectors=np.eye(3)[:1])
        edos, enums = kpm.stepwise_samples(kpm.Ops(m.H, m.S, m.velocity, rtol=1e-13), kcfg, B)
        want = np.array([enums["msd_A2/commutator"][i:i + 4].sum(0) for i in range(0, 16, 4)])
        got = tuned.arrays["numerator_block_sums/msd_A2/commutator"]
        assert np.max(np.abs(got - want)) <= 1e-9 * np.abs(want).max()
