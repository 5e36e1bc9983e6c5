"""Device statistics (kmsd_block_accumulate, kmsd_time_gradient,
kmsd_bootstrap_ratio, kmsd_dos_validity) against the oracle's restatement of
the reference's statistics (oracle/statistics.py): block sums, ratios,
gradients and masks bit for bit; bootstrap standard errors within 1e-12
(the resampled sums are a BLAS product in the reference)."""
import numpy as np
import pytest

from oracle import statistics as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stats(hip_ctx):
    from kubomsd_amd.observables.statistics_device import DeviceStatistics
    return DeviceStatistics(hip_ctx)


def dev(ctx, a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, float)).to(ctx.torch_device)


def host(t):
    return t.detach().cpu().numpy()


def test_block_accumulate_in_sample_order(hip_ctx, stats):
    import torch
    from kubomsd_amd.errors import NumericalError
    rng = np.random.default_rng(1)
    x = rng.normal(size=(11, 2, 3, 5)) * 10.0 ** rng.integers(-8, 8, size=(11, 2, 3, 5))
    acc = torch.zeros((3, 2, 3, 5), dtype=torch.float64, device=hip_ctx.torch_device)
    stats.clear_flags()
    for first in range(0, 11, 3):  # batches of 3 inside blocks of 4
        ids = [i for i in range(first, min(first + 3, 11))]
        for blk in sorted({i // 4 for i in ids}):
            sel = [i for i in ids if i // 4 == blk]
            stats.accumulate(acc[blk], dev(hip_ctx, x[sel]), 2)
    stats.check_flags()
    np.testing.assert_array_equal(host(acc), O.block_sums(x, range(11), 4, 3))
    bad = x[:2].copy()
    bad[1, 0, 0, 0] = np.nan
    stats.clear_flags()
    stats.accumulate(acc[0], dev(hip_ctx, bad), 2)
    with pytest.raises(NumericalError, match="nonfinite stochastic numerator"):
        stats.check_flags()
    stats.clear_flags()


@pytest.mark.parametrize("t", [[0.0, 0.5, 1.0], list(np.linspace(0, 50, 51)), [0, .1, .31, 1.], [0., 2.5],
                               list(np.cumsum(np.random.default_rng(2).uniform(.1, 1, 12)))])
def test_time_gradient_equals_numpy(hip_ctx, stats, t):
    rng = np.random.default_rng(3)
    msd = rng.normal(size=(4, 2, len(t), 7))
    got = host(stats.msd_gradient(t, dev(hip_ctx, msd), axis=2))
    np.testing.assert_array_equal(got, O.msd_gradient(t, msd, axis=2))
    assert stats.msd_gradient([0.0], dev(hip_ctx, msd[:, :, :1]), axis=2) is None


@pytest.mark.parametrize("nb", [1, 2, 5, 9])
@pytest.mark.parametrize("dcols", ["energy", "counts"])
def test_bootstrap_matches_oracle(hip_ctx, stats, nb, dcols):
    rng = np.random.default_rng(10 + nb)
    ne = 7
    A = rng.normal(size=(nb, 2, 3, ne))
    counts = rng.integers(1, 4, size=nb)
    if dcols == "energy":
        D = rng.uniform(.5, 2, size=(nb, ne))
        D[:, 3] = 0.0  # a zero denominator column: NaN value and NaN error
    else:
        D = counts[:, None].astype(float)
    value, se, status, total = stats.estimate_ratio(dev(hip_ctx, A), dev(hip_ctx, D), counts, want_total=True)
    want_v, want_se = O.estimate_ratio(A, D, counts)
    np.testing.assert_array_equal(host(value), want_v)
    np.testing.assert_array_equal(host(total), A.sum(0))
    assert status["n_blocks"] == nb and status["reliable_interval"] == (nb >= 8)
    if nb < 2:
        assert se is None and want_se is None
    else:
        np.testing.assert_allclose(host(se), want_se, rtol=1e-12, atol=0, equal_nan=True)
        assert np.isnan(host(se)).any() == np.isnan(want_se).any()


def test_bootstrap_validity_mask_and_dos_masks(hip_ctx, stats):
    import torch
    rng = np.random.default_rng(4)
    nb, ne = 6, 9
    dos = rng.uniform(-.1, 1, size=(nb, ne))
    dos[:, 2] = 1e-12
    counts = np.full(nb, 2)
    rho, se, _ = stats.estimate_ratio(dev(hip_ctx, dos), dev(hip_ctx, counts[:, None].astype(float)), counts)
    want_rho, want_se = O.estimate_ratio(dos, counts[:, None].astype(float), counts)
    np.testing.assert_array_equal(host(rho), want_rho)
    valid, reason = stats.dos_validity(rho, se, 1e-6, nb)
    wv, wr = O.dos_validity(want_rho, want_se, 1e-6, nb)
    np.testing.assert_array_equal(host(valid), wv)
    np.testing.assert_array_equal(host(reason).astype("uint32"), wr)
    A = rng.normal(size=(nb, 1, 2, ne))
    v, s, _ = stats.estimate_ratio(dev(hip_ctx, A), dev(hip_ctx, dos), counts, valid=valid)
    wv_, ws_ = O.estimate_ratio(A, dos, counts)
    wv_[..., ~wv] = np.nan
    ws_[..., ~wv] = np.nan
    np.testing.assert_array_equal(host(v), wv_)
    np.testing.assert_allclose(host(s), ws_, rtol=1e-12, equal_nan=True)
    v1, r1 = stats.dos_validity(rho, None, 1e-6, 1)
    w1, q1 = O.dos_validity(want_rho, None, 1e-6, 1)
    np.testing.assert_array_equal(host(v1), w1)
    np.testing.assert_array_equal(host(r1).astype("uint32"), q1)
    assert torch.is_tensor(v1)


def test_stochastic_blocks_statistics_equal_oracle(hip_ctx):
    """End to end: the arrays iter_stochastic_blocks returns are the oracle's
    statistics of the block sums it returns (ratios and masks bit for bit,
    errors within 1e-12), with the gradient taken per batch."""
    from helpers import phases
    from test_gpu_solvers import bundle, run_config
    from kubomsd_amd.kernels.energy import dos_floor
    from kubomsd_amd.solvers.stochastic import iter_stochastic_blocks
    from kubomsd_amd.synthetic import square_lattice
    m = square_lattice(12, seed=9)
    E = np.linspace(-4, 4, 17)
    cfg = run_config(energy_eV=E, time_fs=[0, .5, 1.0, 2.0], spectral_bounds=(-5.1, 5.1), n_moments=64,
                     axes=("x", "y"), nsamples=9, statistics_block=2, sample_batch=3, segment_fs=1.0)
    ops = bundle(hip_ctx, m.H, m.S, m.velocity, cfg)
    blk = list(iter_stochastic_blocks(ops, cfg, "stepwise", samples=phases(m.n, 9, 8)))[0]
    counts = blk.static["block_counts"]
    rho, se = O.estimate_ratio(blk.static["dos_block_sums"], counts[:, None], counts)
    np.testing.assert_array_equal(blk.static["dos_states_per_eV"], rho)
    np.testing.assert_allclose(blk.static["stderr_dos_states_per_eV"], se, rtol=1e-12)
    thr = dos_floor(ops.n, 10.2, cfg)
    valid, reason = O.dos_validity(rho, se, thr, len(counts))
    for key in ("msd_A2/commutator", "diffusion_A2_per_fs/msd_gradient"):
        raw = blk.arrays["numerator_block_sums/" + key]
        v, s = O.estimate_ratio(raw, blk.static["dos_block_sums"], counts)
        v[..., ~valid] = np.nan
        s[..., ~valid] = np.nan
        np.testing.assert_array_equal(blk.arrays[key], v)
        np.testing.assert_allclose(blk.arrays["stderr_" + key], s, rtol=1e-12, equal_nan=True)
        np.testing.assert_array_equal(blk.arrays["numerator_" + key], raw.sum(0) / 9)
    np.testing.assert_array_equal(blk.arrays["valid/commutator"][0, 0], valid)
    np.testing.assert_array_equal(blk.arrays["reason_code/commutator"][0, 0], reason)
