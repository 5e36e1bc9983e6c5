"""Random-phase samples and the S^-1/2 metric correction (reference
src/kubo_msd/metric/sampling.py:8-46).

random_phases: column id -> exp(2 pi i u), u = NumPy PCG64(SeedSequence([seed, id])).random(n),
the reference's generator (a NumPy library routine), so sample ids select the
same vectors on every backend.  MetricSampler builds the Chebyshev series of
x^-1/2 on the spectrum of W = Z S Z (Z = diag(S)^-1/2) from a type-II DCT and
applies B = Z p(W) xi on the device by Clenshaw recursion (kmsd_metric_sample).
"""
import numpy as np
from scipy.fft import dct

from ..errors import NumericalError
from .bounds import spd_bounds


def random_phases(n, seed, sample_ids):
    cols = [np.exp(2j * np.pi * np.random.Generator(np.random.PCG64(np.random.SeedSequence([int(seed), int(i)])))
                   .random(n)) for i in sample_ids]
    return np.column_stack(cols)


def inverse_sqrt_series(center, half, tol, max_order):
    """Chebyshev coefficients of 1/sqrt(center + half x) on [-1, 1], order
    doubling from 16 until the max relative grid error < tol/4."""
    if half < 1e-14 * center:
        return np.array([1 / np.sqrt(center)]), dict(order=0, max_relative_grid_error=0.)
    for k in range(4, 20):
        order = min(2 ** k, max_order)
        n = max(2 * order, 64)
        theta = np.pi * (np.arange(n) + .5) / n
        coef = dct(1 / np.sqrt(center + half * np.cos(theta)), type=2) / n
        coef[0] *= .5
        coef = coef[:order + 1]
        xx = np.cos(np.linspace(0, np.pi, max(4097, 8 * order + 1)))
        err = np.max(np.abs(np.polynomial.chebyshev.chebval(xx, coef) * np.sqrt(center + half * xx) - 1))
        if err < tol * .25:
            return coef, dict(order=order, max_relative_grid_error=float(err))
        if order >= max_order:
            break
    raise NumericalError("metric inverse sqrt polynomial failed sampling tolerance")


class MetricSampler:
    def __init__(self, metric, config):
        self.metric = metric
        S = metric.host_S
        z = 1 / np.sqrt(metric.diagonal)
        import scipy.sparse as sp
        W = sp.diags(z) @ S @ sp.diags(z)
        lo, hi, report = spd_bounds(W)
        self.center, self.half = (hi + lo) / 2, (hi - lo) / 2
        self.coeff, rep = inverse_sqrt_series(self.center, self.half, config.sampling_tol, config.sampling_max_order)
        self.report = {"bounds": [lo, hi], "bounds_report": report, "preconditioner": "diagonal congruence",
                       "target": config.sampling_tol, **rep}

    def transform(self, xi):
        return self.metric.device.sample(xi, self.coeff, self.center, self.half)


def sample_metric(seed, sample_ids, ops, config):
    """Complex128 samples in both precisions: the propagated states start
    from them unrounded; each moment chain rounds its own input to the
    precision's storage type."""
    import torch
    if getattr(ops, "sampler", None) is None:
        ops.sampler = MetricSampler(ops.metric, config)
    xi = ops.asarray(random_phases(ops.n, seed, sample_ids), torch.complex128)
    return ops.sampler.transform(xi)
