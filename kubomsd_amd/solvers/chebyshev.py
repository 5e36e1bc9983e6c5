"""Chebyshev scaling, Bessel propagation coefficients, chains and moment
projections on the device (reference src/kubo_msd/solvers/chebyshev.py:8-61).

Every chain runs inside the C ABI (kmsd_chain / kmsd_moments): one fused
launch per Chebyshev order for S = I, SpMM + batched PCG + fused update
otherwise; only the (ne x M) reconstruction G @ mu happens here, on device.
"""
import numpy as np
from scipy.special import gammaln, jv

from ..errors import NumericalError
from ..kernels.energy import reconstruction_matrix
from ..units import HBAR_EV_FS


def scaling(cfg):
    if cfg.spectral_bounds is None:
        raise NumericalError("propagation needs resolved full spectral bounds")
    lo, hi = cfg.spectral_bounds
    return (lo + hi) / 2, (hi - lo) / 2


def scaled_action(ops, cfg, B):
    c, a = scaling(cfg)
    return (ops.apply_h_action(B) - c * B) / a


def evaluate_series(B, action, coeff, xp=None):
    """sum_k coeff[k] T_k(A) B for a caller-supplied action A (reference
    chebyshev.py:32-39), by the three-term recursion on device tensors; the
    solvers here use the fused kmsd_series instead.  xp is accepted for the
    reference's signature and unused."""
    import torch
    prev = B
    out = coeff[0] * B
    if len(coeff) > 1:
        curr = action(B)
        out = out + coeff[1] * curr
        for v in coeff[2:]:
            nxt = 2 * action(curr) - prev
            out = out + v * nxt
            prev, curr = curr, nxt
    if not bool(torch.isfinite(torch.view_as_real(out) if out.is_complex() else out).all()):
        raise NumericalError("nonfinite Chebyshev propagation; check bounds and solves")
    return out


def bessel_coefficients(dt, cfg, weighted_tail=False):
    """c_k = (-i)^k J_k(z) (2 for k >= 1) exp(-i c dt/hbar), z = a dt/hbar, with
    the order chosen by the closed-form Bessel tail bound (chebyshev.py:17-30)."""
    c, a = scaling(cfg)
    z = a * dt / HBAR_EV_FS
    az = abs(z)
    if az == 0:
        return np.array([1 + 0j]), {"order": 0, "tail_bound": 0.}
    chosen = None
    bound = np.inf
    for k in range(max(1, int(np.ceil(az))), cfg.propagation_max_order + 1):
        n = k + 1
        ratio = az / (2 * (n + 1))
        if ratio >= 1:
            continue
        logbound = n * np.log(az / 2) - gammaln(n + 1) + az * az / (4 * (n + 1))
        bound = 2 * np.exp(min(logbound, 700)) / (1 - ratio)
        if weighted_tail:
            bound *= 4 * (n + 1) ** 2 / (1 - ratio) ** 2
        if bound <= cfg.propagation_tol:
            chosen = k
            break
    if chosen is None:
        raise NumericalError("propagation Bessel tail not converged before max order")
    m = np.arange(chosen + 1)
    b = jv(m, z) * (-1j) ** m
    b[1:] *= 2
    b *= np.exp(-1j * c * dt / HBAR_EV_FS)
    return b, {"order": chosen, "tail_bound": float(bound),
               "tail_policy": "absolute Bessel-series bound; weighted for commutator if requested"}


def chain(B, ops, cfg, length):
    c, a = scaling(cfg)
    return ops.native.chain(B, length, c, a)


def reconstruction(ops, cfg):
    """Device float64 G (ne, M), cached on the bundle per (energies, bounds, M)."""
    import torch
    key = (cfg.energy_eV.tobytes(), tuple(cfg.spectral_bounds), cfg.n_moments)
    cache = getattr(ops, "_reconstruction", None)
    if cache is None or cache[0] != key:
        G = torch.from_numpy(reconstruction_matrix(cfg.energy_eV, cfg.spectral_bounds, cfg.n_moments))
        ops._reconstruction = (key, G.to(ops.ctx.torch_device))
    return ops._reconstruction[1]


def projected_quadratic(B, ops, cfg):
    """(nb, ne) device float64: DOS_b(E) = sum_l G[E, l] Re<S B_b, T_l B_b>.
    With cfg.moment_doubling and S = I the moments come from the product
    identities mu_2k = 2<T_k, T_k> - mu_0, mu_2k+1 = 2<T_k+1, T_k> - mu_1
    (exact in arithmetic, about half the orders; the rounding differs from
    the direct recursion, which stays the default)."""
    c, a = scaling(cfg)
    doubling = bool(getattr(cfg, "moment_doubling", False)) and ops.metric.identity
    mu = ops.native.moments(B, cfg.n_moments, c, a, doubling=doubling)
    return (reconstruction(ops, cfg) @ mu).T
