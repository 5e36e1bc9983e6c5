"""The oracle's stepwise and spectral orchestration against the independent
eigen-basis closed forms (oracle/eigen.py), per sample, on small
nonorthogonal models; plus the integration-convergence and segment-partition
properties of the reference's spectral tests (tests/unit/test_spectral.py)."""
import numpy as np
import pytest

from helpers import dense_model, random_block
from oracle import eigen, kpm


def setup(n=9, seed=903):
    m = dense_model(n, seed)
    H, S, X = m.H.toarray(), m.S.toarray(), m.velocity["x"].toarray()
    eps = np.linalg.eigvalsh(np.linalg.solve(S, H))
    bounds = (float(eps.min()) - 0.3, float(eps.max()) + 0.3)
    return H, S, X, bounds


def test_stepwise_equals_eigen_closed_form_per_sample():
    H, S, X, bounds = setup()
    ops = kpm.Ops(H, S, {"x": X}, rtol=1e-14)
    E = np.linspace(bounds[0] + .2, bounds[1] - .2, 7)
    cfg = kpm.config(energy_eV=E, time_fs=[0, .2, .5], spectral_bounds=bounds, n_moments=40, segment_fs=.3,
                     axes_vectors=np.eye(3)[:1])
    B = random_block(9, 3, 5)
    dos, nums = kpm.stepwise_samples(ops, cfg, B)
    edos, enums = eigen.observables(H, S, [X], B, E, bounds, 40, cfg.time_fs, .01)
    np.testing.assert_allclose(dos, edos, rtol=1e-10, atol=1e-12 * np.abs(edos).max())
    np.testing.assert_allclose(nums["msd_A2/commutator"], enums["msd_A2/commutator"], rtol=1e-9,
                               atol=1e-11 * np.abs(enums["msd_A2/commutator"]).max())


def test_spectral_equals_eigen_closed_form_and_converges():
    H, S, X, bounds = setup()
    ops = kpm.Ops(H, S, {"x": X}, rtol=1e-14)
    E = np.linspace(bounds[0] + .2, bounds[1] - .2, 5)
    B = random_block(9, 2, 6)
    cfg = kpm.config(energy_eV=E, time_fs=[0, .2, .5], spectral_bounds=bounds, n_moments=32, segment_fs=.3,
                     integration_dt_fs=.01)
    _, nums = kpm.spectral_samples(ops, cfg, B)
    _, enums = eigen.observables(H, S, [X], B, E, bounds, 32, cfg.time_fs, .01)
    for key in ("vac_A2_per_fs2/vac_integral", "msd_A2/vac_integral", "diffusion_A2_per_fs/vac_integral"):
        np.testing.assert_allclose(nums[key], enums[key], rtol=1e-9, atol=1e-11 * np.abs(enums[key]).max())
    # With a complete metric basis B = sqrt(n) L^-H (S = L L^H) both estimators
    # are traces, so the VAC-integral MSD converges to the exact commutator
    # MSD; the error falls by > 3x per halving of the integration step
    # (tests/unit/test_spectral.py:14-18).
    import scipy.linalg as la
    Bc = np.sqrt(9) * la.solve_triangular(la.cholesky(S, lower=True).conj().T, np.eye(9), lower=False)
    _, exact = eigen.observables(H, S, [X], Bc, E, bounds, 32, [0, .2, .5], .01)
    errors = []
    for dt in (.02, .01, .005):
        _, z = kpm.spectral_samples(ops, kpm.config(energy_eV=E, time_fs=[0, .2, .5], spectral_bounds=bounds,
                                                     n_moments=32, segment_fs=.3, integration_dt_fs=dt), Bc)
        errors.append(np.max(np.abs(z["msd_A2/vac_integral"].mean(0) - exact["msd_A2/commutator"].mean(0))))
    assert errors[1] < errors[0] * .3 and errors[2] < errors[1] * .3


def test_spectral_segment_partition_is_equivalent():
    H, S, X, bounds = setup()
    ops = kpm.Ops(H, S, {"x": X}, rtol=1e-14)
    E = np.linspace(bounds[0] + .2, bounds[1] - .2, 4)
    B = random_block(9, 2, 7)
    a = kpm.spectral_samples(ops, kpm.config(energy_eV=E, time_fs=[0, .2, .7], spectral_bounds=bounds,
                                             n_moments=24, segment_fs=.3), B)[1]
    b = kpm.spectral_samples(ops, kpm.config(energy_eV=E, time_fs=[0, .2, .7], spectral_bounds=bounds,
                                             n_moments=24, segment_fs=.23), B)[1]
    for key in a:
        np.testing.assert_allclose(a[key], b[key], rtol=1e-10, atol=1e-12 * np.abs(a[key]).max())


def test_planted_dropped_term_in_eigen_form_is_detected():
    """Dropping the velocity from the commutator amplitude (a plausible
    mistake) changes the closed form well beyond the tolerances used above."""
    H, S, X, bounds = setup()
    E = np.linspace(bounds[0] + .2, bounds[1] - .2, 7)
    B = random_block(9, 3, 5)
    _, good = eigen.observables(H, S, [X], B, E, bounds, 40, [0, .2, .5], .01)
    _, bad = eigen.observables(H, S, [np.eye(9)], B, E, bounds, 40, [0, .2, .5], .01)
    assert not np.allclose(good["msd_A2/commutator"], bad["msd_A2/commutator"], rtol=1e-3)
