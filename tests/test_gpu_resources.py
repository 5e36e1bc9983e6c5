"""Resources and residency of the device estimator, mirroring the reference
GPU edition's regressions:
  - a sample batch that does not fit is halved, the samples and results stay
    the same (kubo_msd_gpu/tests/gpu/test_review_regressions.py:35-54);
  - one sample that does not fit is RESOURCE_BLOCKED, and the spectral path
    refuses its segment workspace before building any chain (:9-20);
  - no array data leaves the device while the per-batch estimators run
    (tests/gpu/test_acceptance.py:102-124: download bytes unchanged, outputs
    are device arrays), counted by the library's transfer ledger
    (kmsd_ctx_transfers) and the package's own copies;
  - shared-axes, batch and segment invariance (tests/gpu/test_acceptance.py:
    127-139);
  - the library's thread rule: a second host thread cannot create a context
    while the first thread's contexts live.
"""
from dataclasses import replace

import numpy as np
import pytest

from helpers import dense_model, phases, random_block
from kubomsd_amd.synthetic import square_lattice
from oracle import kpm
from test_gpu_solvers import bundle, run_config, small_model

pytestmark = pytest.mark.gpu


def small_cfg(**kw):
    base = dict(energy_eV=np.linspace(-4, 4, 9), time_fs=[0, .5, 1.0], spectral_bounds=(-5.1, 5.1), n_moments=32,
                axes=("x", "y"), nsamples=5, statistics_block=3, sample_batch=4, segment_fs=1.0)
    base.update(kw)
    return run_config(**base)


def test_sample_batch_halved_without_changing_results(hip_ctx, monkeypatch):
    import kubomsd_amd.solvers.stochastic as st
    m = square_lattice(10, seed=3)
    B = phases(m.n, 5, 2)
    cfg = small_cfg()
    ops = bundle(hip_ctx, m.H, m.S, m.velocity, cfg)
    full = list(st.iter_stochastic_blocks(ops, cfg, "stepwise", samples=B))[0]
    assert full.metadata["resource_batch"] == {"requested": 4, "effective": 4,
                                               "workspace_bytes": st.batch_workspace_bytes(ops, cfg, "stepwise", 4)}
    fits_two = st.batch_workspace_bytes(ops, cfg, "stepwise", 2)
    monkeypatch.setattr(st, "device_headroom", lambda ctx: fits_two)
    got = list(st.iter_stochastic_blocks(ops, cfg, "stepwise", samples=B))[0]
    assert got.metadata["resource_batch"]["effective"] == 2 and got.metadata["admitted_sample_batch"] == 2
    assert [d["sample_ids"] for d in got.metadata["sample_batch_diagnostics"]] == [[0, 2], [2, 4], [4, 5]]
    np.testing.assert_array_equal(got.static["sample_ids"], np.arange(5))
    for key in full.arrays:
        np.testing.assert_allclose(got.arrays[key], full.arrays[key], rtol=1e-9, atol=1e-11, equal_nan=True)
    np.testing.assert_allclose(got.static["dos_block_sums"], full.static["dos_block_sums"], rtol=1e-12)


def test_one_sample_over_budget_is_resource_blocked(hip_ctx, monkeypatch):
    import kubomsd_amd.solvers.stochastic as st
    from kubomsd_amd.errors import ResourceBlocked
    m = square_lattice(8, seed=3)
    cfg = small_cfg()
    ops = bundle(hip_ctx, m.H, m.S, m.velocity, cfg)
    monkeypatch.setattr(st, "device_headroom", lambda ctx: st.batch_workspace_bytes(ops, cfg, "stepwise", 1) - 1)
    with pytest.raises(ResourceBlocked, match="RESOURCE_BLOCKED"):
        list(st.iter_stochastic_blocks(ops, cfg, "stepwise", samples=phases(m.n, 5, 2)))


def test_spectral_segment_gate_before_any_chain(hip_ctx, monkeypatch):
    """The segment workspace is checked against the device headroom before
    the segment's chains are built."""
    import kubomsd_amd.solvers.spectral as spectral
    import kubomsd_amd.solvers.stochastic as st
    from kubomsd_amd.errors import NumericalError
    m = square_lattice(8, seed=3)
    cfg = small_cfg(stochastic_method="spectral", include_vac=True, axes=("x",), integration_dt_fs=.25)
    ops = bundle(hip_ctx, m.H, m.S, m.velocity, cfg)
    samples = ops.asarray(phases(m.n, 1, 2))
    monkeypatch.setattr(st, "device_headroom", lambda ctx: 1024)
    monkeypatch.setattr(spectral, "chain", lambda *a, **k: pytest.fail("chain built before the resource gate"))
    with pytest.raises(NumericalError, match="spectral workspace exceeds memory budget"):
        spectral.spectral_samples(ops, cfg, samples)


def _nonorthogonal():
    H, S, X, bounds = small_model()
    return H, S, {"x": X, "y": X.T.copy()}, bounds


@pytest.mark.parametrize("method", ["stepwise", "spectral"])
@pytest.mark.parametrize("metric", ["identity", "nonorthogonal"])
def test_estimators_keep_array_data_on_device(hip_ctx, method, metric):
    import torch
    from kubomsd_amd.solvers.spectral import spectral_samples
    from kubomsd_amd.solvers.stepwise import stepwise_samples
    if metric == "identity":
        m = square_lattice(12, seed=5)
        H, S, V, bounds = m.H, m.S, m.velocity, (-5.1, 5.1)
        B = phases(m.n, 3, 7)
    else:
        H, S, V, bounds = _nonorthogonal()
        V = {"x": V["x"]}
        B = random_block(9, 3, 7)
    axes = tuple(V)
    E = np.linspace(bounds[0] + .3, bounds[1] - .3, 9)
    cfg = run_config(energy_eV=E, time_fs=[0, .3, .6], spectral_bounds=bounds, n_moments=48, axes=axes,
                     include_vac=True, integration_dt_fs=.05, segment_fs=.5, metric_rtol=1e-13)
    ops = bundle(hip_ctx, H, S, V, cfg)
    Bd = ops.asarray(B, torch.complex128)
    fn = stepwise_samples if method == "stepwise" else spectral_samples
    fn(ops, cfg, Bd)  # first call builds per-engine tables (velocity slots, reconstruction matrix)
    before = hip_ctx.transfers()
    dos, arrays, _ = fn(ops, cfg, Bd)
    after = hip_ctx.transfers()
    assert after["library_d2h_data"] == before["library_d2h_data"]
    assert after["host_d2h_data"] == before["host_d2h_data"]
    assert isinstance(dos, torch.Tensor) and dos.device == hip_ctx.torch_device
    for value in arrays.values():
        assert isinstance(value, torch.Tensor) and value.device == hip_ctx.torch_device
    # and the results are the oracle's
    kcfg = kpm.config(energy_eV=E, time_fs=cfg.time_fs, spectral_bounds=bounds, n_moments=48, segment_fs=.5,
                      integration_dt_fs=.05, axes_vectors=np.eye(3)[["xyz".index(a) for a in axes]])
    kfn = kpm.stepwise_samples if method == "stepwise" else kpm.spectral_samples
    edos, enums = kfn(kpm.Ops(H, S, V, rtol=1e-13), kcfg, B)
    scale = np.abs(edos).max()
    assert np.max(np.abs(dos.cpu().numpy() - edos)) <= 1e-9 * scale
    for key, want in enums.items():
        got = arrays[key].cpu().numpy()
        assert np.max(np.abs(got - want)) <= 1e-8 * np.abs(want).max(), key


@pytest.mark.parametrize("method", ["stepwise", "spectral"])
def test_shared_axes_batch_and_segment_invariance(hip_ctx, method):
    from kubomsd_amd.solvers.stochastic import iter_stochastic_blocks
    m = square_lattice(10, seed=13)
    B = phases(m.n, 9, 4)
    cfg = small_cfg(nsamples=9, statistics_block=2, sample_batch=4, include_vac=False, time_fs=[0, .3, .6],
                    integration_dt_fs=.05, stochastic_method=method)

    def calc(c):
        return list(iter_stochastic_blocks(bundle(hip_ctx, m.H, m.S, m.velocity, c), c, method, samples=B))[0]
    whole = calc(cfg)
    split = calc(replace(cfg, sample_batch=1, segment_fs=.13))
    one = calc(replace(cfg, axes=("x",), axes_vectors=None))
    key = "msd_A2/" + ("commutator" if method == "stepwise" else "vac_integral")
    np.testing.assert_allclose(whole.arrays[key], split.arrays[key], rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(whole.arrays[key][0], one.arrays[key][0], rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(whole.static["block_counts"], [2, 2, 2, 2, 1], rtol=0, atol=0)


def test_second_thread_cannot_create_a_context(hip_ctx):
    import threading
    from kubomsd_amd.backends.hip import HipContext
    from kubomsd_amd.errors import KuboMSDError
    seen = []

    def other():
        try:
            HipContext(0)
            seen.append("created")
        except KuboMSDError as exc:
            seen.append(str(exc))
    t = threading.Thread(target=other)
    t.start()
    t.join()
    assert len(seen) == 1 and "one host thread" in seen[0], seen
    assert hip_ctx.memory()["budget_bytes"] > 0  # the owner's context is unaffected
