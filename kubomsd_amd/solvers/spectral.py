"""Segmented VAC from Chebyshev product tables on the device (reference
src/kubo_msd/solvers/spectral.py:15-53).

Per segment: left = chain(c, M+K) and, per axis, cov = V chain(S^-1 V c, K+1)
stay in HBM; table[b, j, k] = sum_i conj(left[j, i, b]) cov[k, i, b] is the
batched tall-skinny GEMM of kmsd_mu_table.  The integration nodes a
segment's tables serve are batched: the Bessel contractions q = table beta,
the product identity mu_l = Re sum_p 1/2 (q_{l+p} + q_{|l-p|}) conj(beta_p)
and vac = mu G^T for all of them are a few device products, and M and D
follow the reference's recurrence as running sums over the nodes.
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


def _segments(times, segment_fs):
    """[(anchor, end, node indices)]: the integration nodes each segment's
    tables serve.  A segment starts at the first node past the previous
    segment's end, anchored at the node before it, and ends segment_fs
    later, at the latest at the last node (reference spectral.py:33-36)."""
    segments = []
    end, prev_t = -1., 0.
    for i, t in enumerate(times):
        t = float(t)
        if not segments or t > end + 1e-14:
            anchor = 0. if not segments else prev_t
            end = min(max(anchor + segment_fs, t), float(times[-1]))
            segments.append((anchor, end, []))
        segments[-1][2].append(i)
        prev_t = t
    return segments


def _node_vac(table, beta, G, L, chunk_bytes):
    """vac (nodes, nb, ne) of one segment's nodes.  beta (K+1, nodes):
    each node's Bessel coefficients, zero past its own order.  q = table beta
    for all nodes at once; mu_l = 1/2 Re sum_p (q_{l+p} + q_{|l-p|})
    conj(beta_p) as two windowed products, the second over the reflected
    sequence R_j = q_{|j-K|} (q_{|l-p|} = R_{l+K-p}); vac = mu G^T."""
    import torch
    nb, LK, K1 = table.shape
    K = K1 - 1
    nodes = beta.shape[1]
    Q = torch.matmul(table, beta).permute(2, 0, 1)  # (nodes, nb, L+K)
    w = beta.conj().T.contiguous()  # (nodes, K+1)
    wf = w.flip(-1)
    per_node = 16 * nb * L * K1 * 3
    step = max(1, min(nodes, chunk_bytes // per_node))
    out = []
    for a in range(0, nodes, step):
        q = Q[a:a + step]
        r = torch.cat([q[..., 1:K + 1].flip(-1), q[..., :L]], dim=-1)  # (n, nb, L+K)
        m = torch.matmul(q.unfold(-1, K1, 1), w[a:a + step, None, :, None]) + \
            torch.matmul(r.unfold(-1, K1, 1), wf[a:a + step, None, :, None])
        out.append((0.5 * m.squeeze(-1).real) @ G.T)  # (n, nb, ne)
    return torch.cat(out, dim=0)


def spectral_samples(ops, cfg, samples):
    """The reference's segmented VAC (spectral.py:15-53) with the integration
    nodes of each segment batched: their Bessel contractions, product
    identities and energy projections are a few device products per
    segment, and M and D follow the reference's recurrence
    M += 2hD + h^2/3 (2C_prev + C), D += h/2 (C_prev + C) as running sums
    over the nodes."""
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
    anchor = 0.
    maxorder = 0
    maxnorm = 0.
    M = torch.zeros((nb, na, ne), dtype=torch.float64, device=dev)
    D = torch.zeros_like(M)
    previous = None  # vac of the node before the current segment's first
    prev_t = 0.
    n = ops.n
    segments = _segments(times, cfg.segment_fs)
    for new_anchor, end, nodes in segments:
        if new_anchor > anchor:
            both = torch.cat([c_anchor] + v_anchor, dim=1).contiguous()
            both, report = propagate(both, new_anchor - anchor, ops, cfg)
            maxnorm = max(maxnorm, report["relative_S_norm_drift"])
            c_anchor = both[:, :nb].contiguous()
            v_anchor = [both[:, (i + 1) * nb:(i + 2) * nb].contiguous() for i in range(na)]
        anchor = new_anchor
        coeff, _ = bessel_coefficients(end - anchor, cfg)
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
        ts = [float(times[i]) for i in nodes]
        beta = np.zeros((K + 1, len(nodes)), dtype=np.complex128)
        for j, t in enumerate(ts):
            b, _ = bessel_coefficients(t - anchor, cfg)
            if len(b) - 1 > K:
                raise NumericalError("internal spectral propagation order exceeded segment allocation")
            beta[:len(b), j] = b
        beta = torch.from_numpy(beta).to(dev)
        chunk = max(16 * nb * L * (K + 1) * 3, min(1 << 30, device_headroom(ops.ctx) // 4))
        vac = torch.stack([_node_vac(table, beta, G, L, chunk) for table in tables], dim=2)  # (nodes, nb, na, ne)
        del tables
        # M, D at every node of the segment: running sums of the recurrence's
        # increments (none at the first node of the run, t = 0)
        h = torch.tensor(np.diff(np.r_[prev_t, ts]), dtype=torch.float64, device=dev)[:, None, None, None]
        C_prev = torch.cat([vac[:1] if previous is None else previous[None], vac[:-1]], dim=0)
        dD = h / 2 * (C_prev + vac)
        if previous is None:
            dD[0] = 0
        Dn = torch.cumsum(torch.cat([D[None], dD], dim=0), dim=0)  # Dn[i] = D before node i ... after node i-1
        dM = 2 * h * Dn[:-1] + h * h / 3 * (2 * C_prev + vac)
        if previous is None:
            dM[0] = 0
        Mn = torch.cumsum(torch.cat([M[None], dM], dim=0), dim=0)
        for j, t in enumerate(ts):
            if t in output_indices:
                k = output_indices[t]
                output["msd_A2/vac_integral"][:, :, k, :] = Mn[j + 1]
                output["diffusion_A2_per_fs/vac_integral"][:, :, k, :] = Dn[j + 1]
                output["vac_A2_per_fs2/vac_integral"][:, :, k, :] = vac[j]
        M, D, previous, prev_t = Mn[-1], Dn[-1], vac[-1], ts[-1]
    if not cfg.include_vac:
        output.pop("vac_A2_per_fs2/vac_integral")
    return dos, output, {
        "algorithm": "segmented trace-first Chebyshev product VAC", "segments": len(segments),
        "max_propagation_order": maxorder, "max_relative_S_norm_drift": maxnorm,
        "integration_scheme": "piecewise linear VAC, simultaneous M/D recurrence", "internal_nodes": len(times),
        "integration_dt_fs": cfg.integration_dt_fs, "seconds": time.perf_counter() - t0}


def iter_spectral_blocks(ops, config, samples=None):
    from .stochastic import iter_stochastic_blocks
    yield from iter_stochastic_blocks(ops, config, "spectral", samples)
