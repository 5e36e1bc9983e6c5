"""Spectral velocity-correlation double moments mu_nm on the device.

The reference's spectral estimator contracts Chebyshev product tables
table[b, j, k] = <T_j c_b, V T_k S^-1 V c_b> (src/kubo_msd/solvers/
spectral.py:33-37).  The square form of the same contraction is

    mu[b, n, m] = sum_i conj(A[n, i, b]) B[m, i, b],
    A[n] = T_n(H~) S^-1 V r_b,   B[m] = V T_m(H~) r_b,   n, m < M,

the velocity-correlation double moment from which the energy-resolved
Kubo-Greenwood correlation follows by a double Chebyshev expansion.  Both
chains stay in HBM ((M, N, nb) blocks of the precision's storage type): A by
kmsd_chain from the velocity action, B by kmsd_chain_v (V T_m formed inside
the chain's own steps where the velocity fits H's pattern).  The table runs
on the matrix cores:
  "fp32"   FP32 MFMA products and accumulation over runs of at most 2^16
           entries, FP64 sums of the runs (complex64 chains, one table per
           sample);
  "bf16x3" operands split x = hi + lo in bf16, three bf16 MFMA products per
           real product (complex64 chains, one table per sample);
  "fp64"   FP64 MFMA, products and sums exact to FP64 rounding for either
           chain precision, all samples in one batched launch.
"""
import time

from ..errors import KuboMSDError, ResourceBlocked
from .chebyshev import chain, scaling

MODES = ("fp32", "bf16x3", "fp64")


def velocity_chains(ops, cfg, r, axis, M):
    """(A, B) as above for the (N, nb) device block r, each (M, N, nb)."""
    c, a = scaling(cfg)
    A = chain(ops.apply_v_action(axis, r), ops, cfg, M)
    B = ops.apply_v_chain(axis, r, M, (c, a))
    return A, B


def velocity_correlation_moments(ops, cfg, samples, axis="x", n_moments=None, mode="fp32", reference=False):
    """mu (nb, M, M) complex128 on the device for the samples (N, nb), with
    M = n_moments or cfg.n_moments.  mode: one of MODES, or a sequence of
    them (each table from the same chains; mu and the errors are then dicts
    by mode).  reference=True also forms the FP64 table of the same chains
    (FP64 matrix cores) and reports, per sample, the largest entry error over
    the largest entry.  Returns dict(mu, error_vs_fp64, ms, ...) with HIP-event
    times per stage summed over the samples."""
    import torch
    from ..backends.hip import KMSD_MU_BF16X3, KMSD_MU_FP32, KMSD_MU_FP64, mu_table
    from .stochastic import device_headroom
    modes = (mode,) if isinstance(mode, str) else tuple(mode)
    if not modes or any(md not in MODES for md in modes):
        raise KuboMSDError(f"mode must be one of {MODES} or a sequence of them")
    if any(md != "fp64" for md in modes) and ops.dtype != torch.complex64:
        raise KuboMSDError("fp32 and bf16x3 tables take complex64 chains (precision='mixed')")
    M = int(n_moments or cfg.n_moments)
    R = ops.asarray(samples, torch.complex128).to(ops.dtype)
    n, nb = R.shape
    # MFMA fp32/bf16x3 tables take one sample per launch; FP64 tables all at once
    width = nb if modes == ("fp64",) else 1
    need = 2 * M * n * width * R.element_size() + 16 * M * M * nb * (len(modes) + int(reference))
    if need > device_headroom(ops.ctx):
        raise ResourceBlocked(f"RESOURCE_BLOCKED: velocity-correlation chains need {need} bytes of device memory")
    stream = torch.cuda.current_stream(ops.ctx.torch_device)
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    ms = {"chain": 0.0, "velocity_chain": 0.0, **{"table_" + md: 0.0 for md in modes}}
    if reference:
        ms["table_fp64_reference"] = 0.0
    mu = {md: torch.empty((nb, M, M), dtype=torch.complex128, device=R.device) for md in modes}
    errors = {md: [] for md in modes}
    codes = {"fp32": KMSD_MU_FP32, "bf16x3": KMSD_MU_BF16X3, "fp64": KMSD_MU_FP64}
    c, a = scaling(cfg)
    t0 = time.perf_counter()
    for j0 in range(0, nb, width):
        r = R[:, j0:j0 + width].contiguous()
        marks = [ev() for _ in range(4 + len(modes))]
        marks[0].record(stream)
        A = chain(ops.apply_v_action(axis, r), ops, cfg, M)
        marks[1].record(stream)
        B = ops.apply_v_chain(axis, r, M, (c, a))
        marks[2].record(stream)
        for q, md in enumerate(modes):
            mu[md][j0:j0 + width] = mu_table(ops.ctx, A, B, codes[md])
            marks[3 + q].record(stream)
        if reference:
            ref = mu_table(ops.ctx, A, B, KMSD_MU_FP64)
            marks[3 + len(modes)].record(stream)
            scale = ref.abs().amax(dim=(1, 2)).clamp(min=1e-300)
            for md in modes:
                err = (mu[md][j0:j0 + width] - ref).abs().amax(dim=(1, 2)) / scale
                errors[md].extend(ops.ctx.to_host(err, status=True).tolist())
            del ref
        torch.cuda.synchronize(R.device)
        ms["chain"] += marks[0].elapsed_time(marks[1])
        ms["velocity_chain"] += marks[1].elapsed_time(marks[2])
        for q, md in enumerate(modes):
            ms["table_" + md] += marks[2 + q].elapsed_time(marks[3 + q])
        if reference:
            ms["table_fp64_reference"] += marks[2 + len(modes)].elapsed_time(marks[3 + len(modes)])
        del A, B
    single = isinstance(mode, str)
    return {"mu": mu[modes[0]] if single else mu, "mode": mode, "moments": M,
            "error_vs_fp64": (errors[modes[0]] if single else errors) if reference else None, "ms": ms,
            "table_flops_per_sample": 8.0 * n * M * M, "samples": nb, "seconds": time.perf_counter() - t0}
