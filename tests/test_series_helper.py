"""evaluate_series (reference solvers/chebyshev.py:32-39) with the Bessel
coefficients of bessel_coefficients reproduces the matrix exponential
exp(-i H dt / hbar) (scipy.linalg.expm), on CPU torch tensors."""
import numpy as np
import pytest
import scipy.linalg as la


def test_bessel_series_equals_expm():
    import torch
    from kubomsd_amd.config import RunConfig
    from kubomsd_amd.errors import NumericalError
    from kubomsd_amd.solvers.chebyshev import bessel_coefficients, evaluate_series, scaling
    from kubomsd_amd.units import HBAR_EV_FS
    rng = np.random.default_rng(4)
    X = rng.normal(size=(12, 12)) + 1j * rng.normal(size=(12, 12))
    H = (X + X.conj().T) / 8
    ev = np.linalg.eigvalsh(H)
    cfg = RunConfig([0], [0, 1], solver="stochastic-kpm", kernel="jackson",
                    spectral_bounds=(ev[0] - .1, ev[-1] + .1), propagation_tol=1e-13)
    c, a = scaling(cfg)
    dt = 0.7
    coeff, report = bessel_coefficients(dt, cfg)
    Ht = torch.from_numpy((H - c * np.eye(12)) / a)
    B = torch.from_numpy(rng.normal(size=(12, 3)) + 1j * rng.normal(size=(12, 3)))
    got = evaluate_series(B, lambda Y: Ht @ Y, coeff).numpy()
    want = la.expm(-1j * H * dt / HBAR_EV_FS) @ B.numpy()
    assert np.abs(got - want).max() < 1e-11 * np.abs(want).max()
    with pytest.raises(NumericalError, match="nonfinite"):
        evaluate_series(B, lambda Y: Y * float("inf"), coeff)
