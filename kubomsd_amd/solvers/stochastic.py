"""Sample-batch orchestration and statistics for the device solvers
(reference src/kubo_msd/solvers/stochastic.py:21-228; device residency and
batch admission as in the GPU edition's solvers/stochastic_gpu.py:18-145).

Samples are taken in global order, `sample_batch` at a time (a batch may
span statistics blocks), so results do not depend on the batch width beyond
rounding.  DOS and numerators stay on the
device: each batch's per-sample arrays are added into device-resident block
sums in sample order, the MSD time gradient, the paired block bootstrap and
the validity masks run as device kernels (observables/statistics_device.py),
and only the finished arrays of each output time slab are copied to the host.
A batch whose device workspace does not fit the HBM budget is halved until it
does (never below one sample, else RESOURCE_BLOCKED).
"""
from dataclasses import replace
import time

import numpy as np

from ..contracts import Reason, TransportBlock
from ..errors import NumericalError, ResourceBlocked
from ..kernels.energy import dos_floor, kernel_resolution
from ..metric.device_bounds import estimate_spectral_bounds_device
from ..metric.sampling import sample_metric
from ..observables.statistics_device import FLAG_DOS, FLAG_NUMERATOR, DeviceStatistics
from .spectral import spectral_samples, spectral_workspace_bytes
from .stepwise import stepwise_samples


def _quantity_keys(cfg, method):
    keys = ["msd_A2/commutator"] if method == "stepwise" else []
    if method == "spectral" or cfg.include_vac:
        keys += ["msd_A2/vac_integral", "diffusion_A2_per_fs/vac_integral"]
        if cfg.include_vac:
            keys += ["vac_A2_per_fs2/vac_integral"]
    if len(cfg.time_fs) > 1:
        keys += ["diffusion_A2_per_fs/msd_gradient"]
    return keys


def batch_workspace_bytes(ops, cfg, method, nb):
    """Device bytes a batch of nb samples needs beyond the resident operators."""
    es = 16 if cfg.precision == "fp64" else 8
    na = len(cfg.axes)
    w = nb * (1 + na)
    # The augmented state is complex128 in both precisions: total + series
    # out + 2 recursion buffers (+ PCG work, 7 blocks, for a nonorthogonal
    # metric); the moment chains of the projections run in the storage type.
    vec = ops.n * (16 * (4 * w + (0 if ops.metric.identity else 7 * w)) + es * 3 * na * nb)
    if method == "spectral" or cfg.include_vac:
        vec = max(vec, spectral_workspace_bytes(ops, cfg, nb))
    return int(vec)


def check_workspace_budget(cfg, keys, nbk):
    """The reference's host workspace budget (solvers/stochastic.py:139-157):
    a batch's per-sample arrays, the merged slabs and, with two or more
    blocks, the bootstrap's resampled numerators, ratios and scratch must fit
    memory_limit_gb; the same estimate gives the same rejection here."""
    na, nt, ne = len(cfg.axes), len(cfg.time_fs), len(cfg.energy_eV)
    tb = min(cfg.time_block, nt)
    working = min(cfg.sample_batch, cfg.statistics_block) * na * nt * ne * 8 * len(keys) * 2
    root = nbk * na * tb * ne * 8 * (len(keys) + 6) + nbk * ne * 16
    if nbk > 1:
        root += 4 * 256 * na * tb * ne * 8 + 2 * 256 * nbk * 8 + 2 * 256 * ne * 8
    if max(working, root) > cfg.memory_limit_gb * 1024 ** 3:
        raise NumericalError("sample/merge/bootstrap workspace exceeds memory budget; reduce "
                             "sample_batch/time_block/energy count or raise explicit budget")


def device_headroom(ctx):
    """Bytes a new workspace may take on the context's device.  Both the
    library's own allocations and torch's (vector blocks, accumulators,
    series outputs) are active memory, as the reference GPU edition counts its
    whole memory pool (backends/gpu_context.py:128-136): the headroom is the
    smaller of the budget minus all active bytes and 85 % of what the device
    can still give (free memory plus torch's cached, unused blocks)."""
    import torch
    mem = ctx.memory()
    dev = ctx.torch_device
    allocated = torch.cuda.memory_allocated(dev)
    reusable = torch.cuda.memory_reserved(dev) - allocated
    free, _ = torch.cuda.mem_get_info(dev)
    return int(min(mem["budget_bytes"] - mem["used_bytes"] - allocated, 0.85 * (free + reusable)))


def admitted_batch(ops, cfg, method):
    """Largest batch of at most cfg.sample_batch samples whose workspace fits
    the headroom, halving from cfg.sample_batch (reference GPU edition,
    solvers/stochastic_gpu.py:31-45); RESOURCE_BLOCKED when one sample does
    not fit."""
    free = device_headroom(ops.ctx)
    nb = cfg.sample_batch
    while nb >= 1:
        if batch_workspace_bytes(ops, cfg, method, nb) <= free:
            return nb
        nb //= 2
    raise ResourceBlocked("RESOURCE_BLOCKED: one sample's device workspace exceeds the HBM budget")


def tuned_batch(ops, cfg, method):
    """Sample batch width from measured step times (cfg.block_width_autotune):
    the admitted cfg.sample_batch and its doublings up to the sample count
    (at most 256) whose workspace fits the headroom, timed as one fused
    Chebyshev step of H on a block of that width (kmsd_time_step); the width
    with the least time per column wins (a wider one only when more than 3 %
    faster).  Per-sample results do not depend on the width beyond the
    rounding of the moment reductions.  Returns (width, {width: us})."""
    import torch
    base = admitted_batch(ops, cfg, method)
    free = device_headroom(ops.ctx)
    widths = [base]
    w = 2 * base
    while w <= min(cfg.nsamples, 256) and batch_workspace_bytes(ops, cfg, method, w) <= free:
        widths.append(w)
        w *= 2
    if len(widths) == 1:
        return base, {}
    dtype = torch.complex64 if cfg.precision == "mixed" else torch.complex128
    return ops.ctx.tune_block_width(ops.H, widths, dtype)


def _sample_batches(ops, cfg, method, samples, accum, dos_blocks, batch, stats, batch_results=None):
    """Runs the samples in global order, `batch` at a time, and adds each
    sample's DOS and numerators to its statistics block in sample order
    (reference _sample_batches, solvers/stochastic.py:32-71).  The reference
    cuts batches at statistics-block boundaries; here a batch may span
    blocks, because a sample's results do not depend on the batch it runs in
    (tests/unit/test_shared_axes.py:5-16) and the device runs wide blocks of
    vectors far faster than narrow ones.  Block sums keep the reference's
    per-sample order."""
    import torch
    diagnostics = []
    primary = "commutator" if method == "stepwise" else "vac_integral"
    ns = cfg.nsamples if samples is None else samples.shape[1]
    sb = cfg.statistics_block
    for first in range(0, ns, batch):
        last = min(first + batch, ns)
        ids = np.arange(first, last)
        if batch_results and (first, last) in batch_results:
            dos, nums, diag = batch_results[(first, last)]  # computed with other jobs' batches
        elif method == "stepwise":
            B = sample_metric(cfg.seed, ids, ops, cfg) if samples is None else \
                ops.asarray(samples[:, first:last], torch.complex128)
            dos, nums, diag = stepwise_samples(ops, cfg, B)
            if cfg.include_vac:
                _, extra, extra_diag = spectral_samples(ops, cfg, B)
                nums.update(extra)
                diag["additional_vac_path"] = extra_diag
                diag["additional_cost_recorded"] = True
        else:
            B = sample_metric(cfg.seed, ids, ops, cfg) if samples is None else \
                ops.asarray(samples[:, first:last], torch.complex128)
            dos, nums, diag = spectral_samples(ops, cfg, B)
        if len(cfg.time_fs) > 1:
            nums["diffusion_A2_per_fs/msd_gradient"] = stats.msd_gradient(cfg.time_fs, nums["msd_A2/" + primary],
                                                                          axis=2)
        diag["sample_ids"] = [int(first), int(last)]
        diagnostics.append(diag)
        # Per-sample sums into each block the batch touches, in sample order
        # (the reference's loop of += per sample), with the nonfinite checks
        # as device flags.
        stats.clear_flags()
        for block in range(first // sb, (last - 1) // sb + 1):
            lo, hi = max(first, block * sb) - first, min(last, (block + 1) * sb) - first
            stats.accumulate(dos_blocks[block], dos[lo:hi], FLAG_DOS)
            for key, value in nums.items():
                stats.accumulate(accum[key][block], value[lo:hi], FLAG_NUMERATOR)
        stats.check_flags()
    return diagnostics


def _normalize_dos(ops, cfg, ns, counts, dos_blocks, report, method, diagnostics, batch, stats, resolution):
    nbk = len(counts)
    primary = "commutator" if method == "stepwise" else "vac_integral"
    import torch
    _host = ops.ctx.to_host
    dcounts = ops.ctx.upload(counts[:, None].astype(float))
    rho, stderr_dos, stat = stats.estimate_ratio(dos_blocks, dcounts, counts)
    threshold = dos_floor(ops.n, cfg.spectral_bounds[1] - cfg.spectral_bounds[0], cfg)
    valid, reason = stats.dos_validity(rho, stderr_dos, threshold, nbk)
    static = {"dos_states_per_eV": _host(rho), "dos_block_sums": _host(dos_blocks), "block_counts": counts,
              "sample_count": np.array(ns), "sample_ids": np.arange(ns)}
    if stderr_dos is not None:
        static["stderr_dos_states_per_eV"] = _host(stderr_dos)
    metadata = {
        "primary_estimator": primary,
        "primary_diffusion_estimator": "msd_gradient" if method == "stepwise" else "vac_integral",
        "resolved_config": cfg.as_dict(), "spectral_report": report,
        "kernel_resolution": resolution, "dos_floor": threshold,
        "statistical_method": stat,
        "sample_generator": "NumPy PCG64 SeedSequence([seed,sample_id]); unit modulus complex phases",
        "sampling": getattr(getattr(ops, "sampler", None), "report", {"source": "caller supplied samples"}),
        "sample_batch_diagnostics": diagnostics, "admitted_sample_batch": batch,
        "resource_batch": {"requested": int(cfg.sample_batch), "effective": int(batch),
                           "workspace_bytes": batch_workspace_bytes(ops, cfg, method, batch)},
        "raw_accumulation": "numerator_* are means over nsamples; *_block_sums are sums",
        "scratch_bytes": 0,
        "parallel": {"distribution": "serial", "world_size": 1, "operator_storage": "one device",
                     "reduction": "device-resident block sums, normalized once"},
        "science_status": {
            "NUMERICS_VERIFIED": "UNVERIFIED for this run",
            "STATISTICS_VERIFIED": "NOT RUN (no multi-seed convergence study for this run)",
            "TRUNCATION_VERIFIED": "not applicable"},
        "statistics_location": "device-resident block sums and bootstrap",
        "backend": {"name": "hip", "precision": cfg.precision, "memory": ops.ctx.memory(),
                    "moment_recursion": "product identities (doubling)"
                    if getattr(cfg, "moment_doubling", False) and ops.metric.identity else "three-term",
                    "metric": ops.metric.diagnostics(), "h_action_calls": ops.h_calls},
    }
    return valid, _host(reason).astype("uint32"), static, metadata


def _normalize_slab(raw, dos_blocks, counts, ns, valid, key, stats):
    _host = stats.ctx.to_host
    raw = raw.contiguous()
    value, se, _, total = stats.estimate_ratio(raw, dos_blocks, counts, valid=valid, want_total=True)
    arrays = {key: _host(value), "numerator_" + key: _host(total) / ns,
              "numerator_block_sums/" + key: _host(raw)}
    if se is not None:
        arrays["stderr_" + key] = _host(se)
    return arrays


def iter_stochastic_blocks(ops, config, method, samples=None, bounds_report=None, batch_results=None):
    """Spectral bounds come from the device two-start generalized Ritz check
    (metric/device_bounds.py), which also validates user bounds as the
    reference does; bounds_report: precomputed bounds (e.g.
    metric.bounds.certified_bounds) used instead.  batch_results:
    {(first, last): (dos, numerators, diagnostics)} of sample batches already
    computed (the serving group runner's); other batches are computed here."""
    cfg = config
    report = bounds_report or estimate_spectral_bounds_device(ops, cfg.spectral_bounds,
                                                              ops.metric.bounds_report["certified"])
    cfg = replace(cfg, spectral_bounds=report["bounds"])
    ns = cfg.nsamples if samples is None else samples.shape[1]
    nbk = (ns + cfg.statistics_block - 1) // cfg.statistics_block
    counts = np.array([min(cfg.statistics_block, ns - i * cfg.statistics_block) for i in range(nbk)])
    keys = _quantity_keys(cfg, method)
    check_workspace_budget(cfg, keys, nbk)
    shape = (nbk, len(cfg.axes), len(cfg.time_fs), len(cfg.energy_eV))
    import torch
    dev = ops.ctx.torch_device
    accum = {key: torch.zeros(shape, dtype=torch.float64, device=dev) for key in keys}
    dos_blocks = torch.zeros((nbk, len(cfg.energy_eV)), dtype=torch.float64, device=dev)
    stats = DeviceStatistics(ops.ctx)
    # precomputed batches were admitted by their runner
    width_times = {}
    if batch_results:
        batch = max(b - a for a, b in batch_results)
    elif getattr(cfg, "block_width_autotune", False):
        batch, width_times = tuned_batch(ops, cfg, method)
    else:
        batch = admitted_batch(ops, cfg, method)
    # The reported kernel resolution is host root-finding only; it runs on a
    # helper thread while the samples' device work (inside library calls
    # that release the GIL) proceeds.
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=1) as pool:
        resolution = pool.submit(kernel_resolution, cfg)
        t0 = time.perf_counter()
        diagnostics = _sample_batches(ops, cfg, method, samples, accum, dos_blocks, batch, stats, batch_results)
        seconds = time.perf_counter() - t0
        valid, reason, static, metadata = _normalize_dos(ops, cfg, ns, counts, dos_blocks, report, method,
                                                         diagnostics, batch, stats, resolution.result())
    metadata["sample_seconds"] = seconds
    if width_times:
        metadata["block_width_tuning"] = {"chosen": int(batch),
                                          "us_per_step": {str(k): float(v) for k, v in width_times.items()}}
    valid_host = ops.ctx.to_host(valid)
    for start in range(0, len(cfg.time_fs), cfg.time_block):
        stop = min(start + cfg.time_block, len(cfg.time_fs))
        arrays = {}
        for key in keys:
            arrays.update(_normalize_slab(accum[key][:, :, start:stop, :], dos_blocks, counts, ns, valid, key,
                                          stats))
        for estimator in {k.split("/")[-1] for k in keys}:
            r = reason.copy()
            if estimator == "vac_integral":
                r |= int(Reason.INTEGRATION_UNVERIFIED)
            dims = (len(cfg.axes), stop - start, len(cfg.energy_eV))
            arrays["valid/" + estimator] = np.broadcast_to(valid_host, dims).copy()
            arrays["reason_code/" + estimator] = np.broadcast_to(r, dims).copy()
        yield TransportBlock(start, cfg.time_fs[start:stop], arrays, static if start == 0 else {},
                             metadata if start == 0 else {})
