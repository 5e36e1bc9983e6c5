"""Segmented VAC from Chebyshev product tables on the device (reference
src/kubo_msd/solvers/spectral.py:15-53).

Per segment: left = chain(c, M+K) and, per axis, cov = V chain(S^-1 V c, K+1)
stay in HBM; table[b, j, k] = sum_i conj(left[j, i, b]) cov[k, i, b] is the
batched tall-skinny GEMM of kmsd_mu_table.  Per integration node the Bessel
contraction q = table beta, the product identity
mu_l = Re sum_p 1/2 (q_{l+p} + q_{|l-p|}) conj(beta_p) and vac = mu G^T are
small device tensor ops; M and D follow the reference's recurrence.
"""
import time

import numpy as np

from ..config import resolve_time_grid
from ..errors import NumericalError
from .chebyshev import bessel_coefficients, chain, projected_quadratic, reconstruction, scaling
from .stepwise import propagate


def spectral_workspace_bytes(ops, cfg, nb):
    """Device bytes of one segment: chains (M+K and K+1 orders) + velocity
    copies + tables (the reference's estimate, spectral.py:31, per element)."""
    from .chebyshev import scaling  # noqa: F401
    es = 16 if cfg.precision == "fp64" else 8
    na = len(cfg.axes)
    K = _max_order(cfg)
    L = cfg.n_moments
    return es * nb * (ops.n * (L + 3 * K + 3 + 2 * na)) + 16 * nb * na * (L + K) * (K + 1)


def _max_order(cfg):
    dt = min(cfg.segment_fs, float(cfg.time_fs[-1])) if len(cfg.time_fs) > 1 else cfg.segment_fs
    try:
        return len(bessel_coefficients(max(dt, 1e-12), cfg)[0]) - 1
    except NumericalError:
        return cfg.propagation_max_order


def _table(ops, left, cov):
    """(nb, L+K, K+1) complex128 tables on the FP64 matrix cores (both
    precisions: complex64 chains are widened exactly)."""
    from ..backends.hip import KMSD_MU_FP64, mu_table
    return mu_table(ops.ctx, left, cov, KMSD_MU_FP64)


def spectral_samples(ops, cfg, samples):
    import torch
    from .stochastic import device_headroom
    # Anchor states are propagated in complex128 in both precisions, from the
    # samples as given; the chains of each segment (and the DOS chain) run in
    # the precision's storage type.
    c_anchor = ops.asarray(samples, torch.complex128)
    samples = c_anchor if ops.dtype == torch.complex128 else c_anchor.to(ops.dtype)
    nb = samples.shape[1]
    na, ne, nt, L = len(cfg.axes), len(cfg.energy_eV), len(cfg.time_fs), cfg.n_moments
    dev = samples.device
    t0 = time.perf_counter()
    dos = projected_quadratic(samples, ops, cfg)
    keys = ("msd_A2/vac_integral", "diffusion_A2_per_fs/vac_integral", "vac_A2_per_fs2/vac_integral")
    output = {key: torch.zeros((nb, na, nt, ne), dtype=torch.float64, device=dev) for key in keys}
    times = np.unique(np.r_[resolve_time_grid(cfg.time_fs[-1], cfg.integration_dt_fs), cfg.time_fs])
    output_indices = {float(t): j for j, t in enumerate(cfg.time_fs)}
    G = reconstruction(ops, cfg)  # (ne, L)
    v_anchor = [ops.apply_v_action(a, c_anchor) for a in cfg.axes]
    anchor, next_anchor = 0., -1.
    tables = None
    maxorder = segments = 0
    maxnorm = 0.
    M = torch.zeros((nb, na, ne), dtype=torch.float64, device=dev)
    D = torch.zeros_like(M)
    previous = None
    prev_t = 0.
    n = ops.n
    for t in times:
        if tables is None or t > next_anchor + 1e-14:
            new_anchor = 0. if tables is None else float(prev_t)
            if new_anchor > anchor:
                both = torch.cat([c_anchor] + v_anchor, dim=1).contiguous()
                both, report = propagate(both, new_anchor - anchor, ops, cfg)
                maxnorm = max(maxnorm, report["relative_S_norm_drift"])
                c_anchor = both[:, :nb].contiguous()
                v_anchor = [both[:, (i + 1) * nb:(i + 2) * nb].contiguous() for i in range(na)]
            anchor = new_anchor
            next_anchor = min(max(anchor + cfg.segment_fs, float(t)), float(times[-1]))
            coeff, _ = bessel_coefficients(next_anchor - anchor, cfg)
            K = len(coeff) - 1
            maxorder = max(maxorder, K)
            need = 16 * nb * (n * (L + 2 * K + 1) + na * (L + K) * (K + 1))
            if need > device_headroom(ops.ctx):
                raise NumericalError("spectral workspace exceeds memory budget; reduce sample_batch/segment_fs or "
                                     "increase explicit budget")
            left = chain(c_anchor.to(samples.dtype), ops, cfg, L + K)  # (L+K, n, nb)
            tables = []
            for ai, axis in enumerate(cfg.axes):
                # cov = V chain(S^-1 V c, K+1): chain and one velocity product over all orders
                cov = ops.apply_v_chain(axis, v_anchor[ai].to(samples.dtype), K + 1, scaling(cfg))  # (K+1, n, nb)
                tables.append(_table(ops, left, cov))  # (nb, L+K, K+1)
                del cov
            del left
            segments += 1
        b, _ = bessel_coefficients(float(t - anchor), cfg)
        Kt = len(b) - 1
        if Kt > tables[0].shape[-1] - 1:
            raise NumericalError("internal spectral propagation order exceeded segment allocation")
        bd = torch.from_numpy(b).to(dev)
        l = torch.arange(L, device=dev)[:, None]
        p = torch.arange(Kt + 1, device=dev)[None, :]
        vac = torch.empty((nb, na, ne), dtype=torch.float64, device=dev)
        for ai, table in enumerate(tables):
            q = table[:, :, :Kt + 1] @ bd  # (nb, L+K)
            moments = (0.5 * (q[:, l + p] + q[:, (l - p).abs()]) * bd.conj()[None, None, :]).sum(2).real
            vac[:, ai, :] = moments @ G.T
        if previous is not None:
            h = float(t - prev_t)
            M = M + 2 * h * D + h * h / 3 * (2 * previous + vac)
            D = D + h / 2 * (previous + vac)
        if float(t) in output_indices:
            j = output_indices[float(t)]
            output["msd_A2/vac_integral"][:, :, j, :] = M
            output["diffusion_A2_per_fs/vac_integral"][:, :, j, :] = D
            output["vac_A2_per_fs2/vac_integral"][:, :, j, :] = vac
        previous = vac
        prev_t = t
    if not cfg.include_vac:
        output.pop("vac_A2_per_fs2/vac_integral")
    return dos, output, {
        "algorithm": "segmented trace-first Chebyshev product VAC", "segments": segments,
        "max_propagation_order": maxorder, "max_relative_S_norm_drift": maxnorm,
        "integration_scheme": "piecewise linear VAC, simultaneous M/D recurrence", "internal_nodes": len(times),
        "integration_dt_fs": cfg.integration_dt_fs, "seconds": time.perf_counter() - t0}


def iter_spectral_blocks(ops, config, samples=None):
    from .stochastic import iter_stochastic_blocks
    yield from iter_stochastic_blocks(ops, config, "spectral", samples)
