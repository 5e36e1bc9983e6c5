"""Device stepwise / spectral / stochastic orchestration against the FP64
oracle orchestration (oracle/kpm.py) on identical models and samples.

Tolerances (max entry error over max entry): fp64 within 1e-9 (the
reference's own cpp-vs-reference stepwise bound is 2e-9,
tests/unit/test_native.py:139-143; measured about 1e-15); mixed (complex64
chains, complex128 propagated states, FP64 reductions) within 2e-7 for DOS and
3e-7 for the MSD and VAC numerators (measured 2-3e-8 on these models).
"""
import numpy as np
import pytest

from helpers import dense_model, phases, random_block
from kubomsd_amd.synthetic import honeycomb_lattice, square_lattice
from oracle import kpm

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-9, "mixed": 2e-7}
NTOL = {"fp64": 1e-9, "mixed": 3e-7}


def run_config(**kw):
    from kubomsd_amd.config import RunConfig
    base = dict(solver="stochastic-kpm", kernel="jackson", backend="hip", metric_rtol=1e-13,
                include_vac=False, segment_fs=.3)
    base.update(kw)
    return RunConfig(**base)


def bundle(ctx, H, S, vel, cfg, atom_slices=()):
    from kubomsd_amd.contracts import ModelBundle
    from kubomsd_amd.operators.velocity import build_operators
    return build_operators(ModelBundle(H, S, velocity=vel, atom_slices=list(atom_slices)), cfg, ctx)


def close(got, want, tol):
    got = got.detach().cpu().numpy() if hasattr(got, "detach") else np.asarray(got)
    scale = max(np.abs(want).max(), 1e-300)
    assert np.max(np.abs(got - want)) <= tol * scale, np.max(np.abs(got - want)) / scale


def small_model():
    m = dense_model(9, 903)
    H, S, X = m.H.toarray(), m.S.toarray(), m.velocity["x"].toarray()
    eps = np.linalg.eigvalsh(np.linalg.solve(S, H))
    return H, S, X, (float(eps.min()) - .3, float(eps.max()) + .3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_stepwise_nonorthogonal_matches_oracle(hip_ctx, precision):
    from kubomsd_amd.solvers.stepwise import stepwise_samples
    H, S, X, bounds = small_model()
    E = np.linspace(bounds[0] + .2, bounds[1] - .2, 7)
    cfg = run_config(energy_eV=E, time_fs=[0, .2, .5], spectral_bounds=bounds, n_moments=40, axes=("x",),
                     precision=precision)
    B = random_block(9, 3, 5)
    dos, nums, diag = stepwise_samples(bundle(hip_ctx, H, S, {"x": X}, cfg), cfg, B)
    kcfg = kpm.config(energy_eV=E, time_fs=cfg.time_fs, spectral_bounds=bounds, n_moments=40, segment_fs=.3)
    edos, enums = kpm.stepwise_samples(kpm.Ops(H, S, {"x": X}, rtol=1e-13), kcfg, B)
    close(dos, edos, TOL[precision])
    close(nums["msd_A2/commutator"], enums["msd_A2/commutator"], NTOL[precision])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_stepwise_lattice_two_axes_matches_oracle(hip_ctx, precision):
    from kubomsd_amd.solvers.stepwise import stepwise_samples
    m = square_lattice(20, seed=7)
    bounds = (-5.1, 5.1)
    E = np.linspace(-4, 4, 33)
    cfg = run_config(energy_eV=E, time_fs=[0, .5, 1.0, 1.5], spectral_bounds=bounds, n_moments=128,
                     axes=("x", "y"), precision=precision, segment_fs=1.0)
    B = phases(m.n, 4, 3)
    dos, nums, _ = stepwise_samples(bundle(hip_ctx, m.H, m.S, m.velocity, cfg), cfg, B)
    kcfg = kpm.config(energy_eV=E, time_fs=cfg.time_fs, spectral_bounds=bounds, n_moments=128, segment_fs=1.0,
                      axes_vectors=np.eye(3)[:2])
    edos, enums = kpm.stepwise_samples(kpm.Ops(m.H, m.S, m.velocity, rtol=1e-13), kcfg, B)
    close(dos, edos, TOL[precision])
    close(nums["msd_A2/commutator"], enums["msd_A2/commutator"], NTOL[precision])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_spectral_matches_oracle(hip_ctx, precision):
    from kubomsd_amd.solvers.spectral import spectral_samples
    m = honeycomb_lattice(8, 6, seed=5)
    bounds = (-8.8, 8.8)
    E = np.linspace(-8, 8, 21)
    cfg = run_config(energy_eV=E, time_fs=[0, .3, .8], spectral_bounds=bounds, n_moments=64, axes=("x",),
                     precision=precision, include_vac=True, integration_dt_fs=.02, segment_fs=.5)
    B = phases(m.n, 3, 4)
    dos, nums, diag = spectral_samples(bundle(hip_ctx, m.H, m.S, m.velocity, cfg), cfg, B)
    kcfg = kpm.config(energy_eV=E, time_fs=cfg.time_fs, spectral_bounds=bounds, n_moments=64, segment_fs=.5,
                      integration_dt_fs=.02)
    edos, enums = kpm.spectral_samples(kpm.Ops(m.H, m.S, m.velocity, rtol=1e-13), kcfg, B)
    close(dos, edos, TOL[precision])
    for key in enums:
        close(nums[key], enums[key], NTOL[precision])


def test_stochastic_blocks_batch_invariance_and_oracle_sums(hip_ctx):
    """Global sample order makes results independent of the batch width
    (tests/unit/test_shared_axes.py:5-16); block sums equal the oracle's."""
    from kubomsd_amd.solvers.stochastic import iter_stochastic_blocks
    m = square_lattice(12, seed=9)
    bounds = (-5.1, 5.1)
    E = np.linspace(-4, 4, 17)
    B = phases(m.n, 9, 8)
    out = []
    for batch in (4, 1):
        cfg = run_config(energy_eV=E, time_fs=[0, .5, 1.0], spectral_bounds=bounds, n_moments=64, axes=("x", "y"),
                         nsamples=9, statistics_block=2, sample_batch=batch, segment_fs=1.0)
        ops = bundle(hip_ctx, m.H, m.S, m.velocity, cfg)
        out.append(list(iter_stochastic_blocks(ops, cfg, "stepwise", samples=B))[0])
    a, b = out
    np.testing.assert_allclose(a.arrays["msd_A2/commutator"], b.arrays["msd_A2/commutator"], rtol=1e-10,
                               atol=1e-12)
    np.testing.assert_allclose(a.static["block_counts"], [2, 2, 2, 2, 1])
    kcfg = kpm.config(energy_eV=E, time_fs=[0, .5, 1.0], spectral_bounds=bounds, n_moments=64, segment_fs=1.0,
                      axes_vectors=np.eye(3)[:2])
    edos, enums = kpm.stepwise_samples(kpm.Ops(m.H, m.S, m.velocity, rtol=1e-13), kcfg, B)
    blocks = [slice(i, min(i + 2, 9)) for i in range(0, 9, 2)]
    want = np.array([enums["msd_A2/commutator"][s].sum(0) for s in blocks])
    close(a.arrays["numerator_block_sums/msd_A2/commutator"], want, 1e-9)
    close(a.static["dos_block_sums"], np.array([edos[s].sum(0) for s in blocks]), 1e-10)
    assert np.all(np.isfinite(a.arrays["diffusion_A2_per_fs/msd_gradient"][a.arrays["valid/commutator"]]))


def test_sampler_covariance_on_device(hip_ctx):
    """Device Clenshaw sampler: C C^H = S^-1 (tests/unit/test_sampling.py:4-14)."""
    import scipy.sparse as sp
    from kubomsd_amd.metric.sampling import MetricSampler
    S = np.array([[1, .2j], [-.2j, 2]])
    cfg = run_config(energy_eV=[0.], time_fs=[0.], spectral_bounds=(-1, 1), sampling_tol=1e-12,
                     axes=("x",))
    ops = bundle(hip_ctx, np.zeros((2, 2)), S, {"x": np.array([[0, 1j], [-1j, 0]])}, cfg)
    sampler = MetricSampler(ops.metric, cfg)
    C = sampler.transform(ops.asarray(np.eye(2, dtype=complex))).cpu().numpy()
    np.testing.assert_allclose(C @ C.conj().T, np.linalg.inv(S), rtol=1e-11, atol=1e-12)


def test_bootstrap_workspace_is_included_in_budget(hip_ctx):
    """tests/unit/test_stochastic_budget.py:9-19: 1000 energies, two blocks and
    a 5 MB budget are rejected before any sampling, naming the bootstrap."""
    from kubomsd_amd.solvers.stochastic import iter_stochastic_blocks
    cfg = run_config(energy_eV=np.linspace(-1, 1, 1000), time_fs=[0., .01], axes=("x",), nsamples=2,
                     statistics_block=1, sample_batch=1, n_moments=4, spectral_bounds=(-2., 2.),
                     memory_limit_gb=.005)
    from kubomsd_amd.contracts import ModelBundle
    from kubomsd_amd.operators.velocity import build_operators
    ops = build_operators(ModelBundle(np.array([[0., .5], [.5, 0.]]), np.eye(2), position={"x": np.diag([0., 1.])}),
                          cfg, hip_ctx)
    with pytest.raises(ValueError, match="bootstrap"):
        list(iter_stochastic_blocks(ops, cfg, "stepwise"))


def test_propagation_against_expm_inverse_segments_and_norm(hip_ctx):
    """tests/unit/test_propagation.py:5-21: propagate = expm(-i S^-1 H t / hbar),
    backward propagation inverts it, segments compose, the S-norm is kept."""
    import scipy.linalg as la
    from kubomsd_amd.contracts import ModelBundle
    from kubomsd_amd.operators.velocity import build_operators
    from kubomsd_amd.solvers.stepwise import propagate
    from kubomsd_amd.units import HBAR_EV_FS
    from helpers import host
    A = np.array([[1, .1j], [.2, 1.2]])
    S = A.conj().T @ A
    H = A.conj().T @ np.array([[.3, .4j], [-.4j, -.2]]) @ A
    V = A.conj().T @ np.array([[.5, 1j], [-1j, 0.]]) @ A
    cfg = run_config(energy_eV=[-.1, .1], time_fs=[0, .2, .7], axes=("x",), spectral_bounds=(-1, 1), n_moments=24,
                     metric_rtol=1e-12, integration_dt_fs=.002, segment_fs=.3)
    ops = build_operators(ModelBundle(H, S, velocity={"x": V}), cfg, hip_ctx)
    B = np.array([[1 + 1j, .2], [.4j, 1]])
    out = host(propagate(B, .7, ops, cfg)[0])
    want = la.expm(-1j * la.solve(S, H) * .7 / HBAR_EV_FS) @ B
    np.testing.assert_allclose(out, want, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(host(propagate(out, -.7, ops, cfg)[0]), B, rtol=1e-10, atol=1e-12)
    mid = propagate(B, .3, ops, cfg)[0]
    np.testing.assert_allclose(host(propagate(mid, .4, ops, cfg)[0]), out, atol=1e-12)
    np.testing.assert_allclose(np.sum(B.conj() * (S @ B), axis=0), np.sum(out.conj() * (S @ out), axis=0), atol=1e-12)


def test_moment_doubling_option_matches_direct_recursion(hip_ctx):
    """RunConfig(moment_doubling=True) on an identity metric: the same DOS and
    MSD as the three-term recursion to rounding (the product identities are
    exact in arithmetic); a nonorthogonal metric keeps the direct recursion."""
    from kubomsd_amd.solvers.stochastic import iter_stochastic_blocks
    m = square_lattice(12, seed=9)
    E = np.linspace(-4, 4, 17)
    B = phases(m.n, 6, 8)
    out = {}
    for doubling in (False, True):
        cfg = run_config(energy_eV=E, time_fs=[0, .5, 1.0], spectral_bounds=(-5.1, 5.1), n_moments=64,
                         axes=("x", "y"), nsamples=6, statistics_block=2, sample_batch=6, segment_fs=1.0,
                         moment_doubling=doubling)
        ops = bundle(hip_ctx, m.H, m.S, m.velocity, cfg)
        blk = list(iter_stochastic_blocks(ops, cfg, "stepwise", samples=B))[0]
        out[doubling] = blk
        assert blk.metadata["backend"]["moment_recursion"] == ("product identities (doubling)" if doubling
                                                                else "three-term")
    a, b = out[False], out[True]
    close(b.static["dos_block_sums"], a.static["dos_block_sums"], 1e-12)
    close(b.arrays["numerator_block_sums/msd_A2/commutator"], a.arrays["numerator_block_sums/msd_A2/commutator"],
          1e-10)


@pytest.mark.parametrize("mode", [dict(preconditioner="ic0"), dict(preconditioner="ic0", metric_ic_ordering="natural"),
                                  dict(metric_solver="cholesky", metric_rtol=1e-12),
                                  dict(preconditioner="block-jacobi")], ids=["ic0-rcm", "ic0-natural", "direct", "block"])
def test_stepwise_nonorthogonal_under_each_metric_mode_matches_oracle(hip_ctx, mode):
    """The stepwise estimator on a nonorthogonal model gives the oracle's
    DOS and MSD whichever metric mode solves S (the estimator only sees S^-1)."""
    from kubomsd_amd.solvers.stepwise import stepwise_samples
    H, S, X, bounds = small_model()
    E = np.linspace(bounds[0] + .2, bounds[1] - .2, 7)
    kw = dict(energy_eV=E, time_fs=[0, .2, .5], spectral_bounds=bounds, n_moments=40, axes=("x",), precision="fp64")
    kw.update(mode)
    cfg = run_config(**kw)
    B = random_block(9, 3, 5)
    dos, nums, _ = stepwise_samples(bundle(hip_ctx, H, S, {"x": X}, cfg, atom_slices=[(0, 3), (3, 9)]), cfg, B)
    kcfg = kpm.config(energy_eV=E, time_fs=cfg.time_fs, spectral_bounds=bounds, n_moments=40, segment_fs=.3)
    edos, enums = kpm.stepwise_samples(kpm.Ops(H, S, {"x": X}, rtol=1e-13), kcfg, B)
    close(dos, edos, 1e-9)
    close(nums["msd_A2/commutator"], enums["msd_A2/commutator"], 1e-8)
