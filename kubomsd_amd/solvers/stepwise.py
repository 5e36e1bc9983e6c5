"""Stepwise commutator MSD on the device (reference
src/kubo_msd/solvers/stepwise.py:8-38).

State X = [B | 0 ... 0] (one zero block per axis) evolves by the augmented
Bessel series; the MSD numerator at each output time is the DOS projection of
each axis block.  All axis blocks are projected in ONE moment chain of width
na*nb (columns are independent, so this equals one chain per axis) and the
projection at t = 0 of the zero blocks is exactly zero.
"""
import logging
import time

import numpy as np

from ..errors import NumericalError
from .chebyshev import bessel_coefficients, projected_quadratic, scaling

log = logging.getLogger(__name__)


def drift_limit(cfg):
    """Largest accepted relative S-norm drift of a propagated block: the
    reference's 1e-7 (solvers/stepwise.py:16-17).  Propagated states are
    complex128 in both precisions (see propagate), so the limit is the same."""
    return 1e-7


def _drift(ops, B, norm0):
    norm = ops.inner(B, B).real
    return ops.ctx.status_value(((norm - norm0).abs() / norm0.clamp(min=np.finfo(float).tiny)).max())


def propagate(states, dt_fs, ops, cfg):
    """exp(-i S^-1 H dt / hbar) states by the Bessel series in segments of at
    most segment_fs.  The propagated state is complex128 whatever the
    precision: it carries rounding from one output time to the next over the
    whole time grid, while the moment chains started from it (the bulk of the
    work) run in the precision's storage type."""
    import torch
    B = ops.asarray(states, torch.complex128)
    initial = ops.inner(B, B).real
    nsegments = max(1, int(np.ceil(abs(dt_fs) / cfg.segment_fs)))
    dt = dt_fs / nsegments
    coeff, diag = bessel_coefficients(dt, cfg)
    c, a = scaling(cfg)
    for _ in range(nsegments):
        B = ops.native.series(B, coeff, c, a)
    drift = _drift(ops, B, initial)
    if drift > drift_limit(cfg):
        raise NumericalError(f"S-norm propagation drift {drift:g}; bounds/solver/propagation need correction")
    return B, {**diag, "segments": nsegments, "relative_S_norm_drift": drift}


def _stepwise_program(ops, cfg, samples):
    """The stepwise estimator as a generator: it yields each moment-chain
    input (the DOS chain's, then one per output time) as (ops, block),
    receives that chain's DOS projection (width, ne), and returns
    (dos, numerators, diagnostics).  stepwise_samples runs it with one chain
    per request; stepwise_samples_group runs several programs in lockstep and
    serves their requests with shared launches."""
    import torch
    # The augmented state is propagated in complex128 in both precisions
    # (about 3 % of the column-steps), starting from the samples as given
    # (complex64 samples are widened exactly); each moment chain, the DOS
    # chain included, runs on its projection to the chain storage type.
    state0 = ops.asarray(samples, torch.complex128)
    initial = state0 if ops.dtype == torch.complex128 else state0.to(ops.dtype)
    nb = initial.shape[1]
    na, nt, ne = len(cfg.axes), len(cfg.time_fs), len(cfg.energy_eV)
    dos = yield ops, initial
    total = torch.cat([state0] + [torch.zeros_like(state0) for _ in cfg.axes], dim=1).contiguous()
    output = torch.zeros((nb, na, nt, ne), dtype=torch.float64, device=initial.device)
    c, a = scaling(cfg)
    orders = []
    norm0 = ops.inner(state0, state0).real
    drift = 0.
    segments = 0
    t0 = time.perf_counter()
    for j, t in enumerate(cfg.time_fs):
        if j:
            span = t - cfg.time_fs[j - 1]
            count = max(1, int(np.ceil(span / cfg.segment_fs)))
            coeff, report = bessel_coefficients(span / count, cfg, weighted_tail=True)
            orders.append(report["order"])
            for _ in range(count):
                total = ops.native.series(total, coeff, c, a, cfg.axes_vectors)
                segments += 1
            drift = max(drift, _drift(ops, total[:, :nb].contiguous(), norm0))
            if drift > drift_limit(cfg):
                raise NumericalError(f"stepwise base propagation S-norm drift {drift:g}")
            proj = yield ops, total[:, nb:].to(initial.dtype).contiguous()  # (na*nb, ne)
            output[:, :, j, :] = proj.reshape(na, nb, ne).permute(1, 0, 2)
            log.info("stepwise: output time %g fs done (%d of %d), %.1f s", t, j, nt - 1, time.perf_counter() - t0)
    return dos, {"msd_A2/commutator": output}, {
        "algorithm": "velocity-driven augmented Chebyshev commutator", "segments": segments,
        "max_propagation_order": max(orders, default=0), "max_relative_S_norm_drift": drift,
        "seconds": time.perf_counter() - t0}


def stepwise_samples(ops, cfg, samples):
    program = _stepwise_program(ops, cfg, samples)
    request = next(program)
    while True:
        try:
            request = program.send(projected_quadratic(request[1], ops, cfg))
        except StopIteration as done:
            return done.value


def stepwise_samples_group(items):
    """stepwise_samples for several (ops, cfg, samples) at once, on one device
    context: the programs advance in lockstep, and the moment chains they
    request together (identity metrics, one moment count, chain width and
    storage type, no moment doubling) run with their orders in shared
    launches (kmsd_moments_group), bit for bit what each would compute alone.
    Returns the list of (dos, numerators, diagnostics)."""
    from ..backends.hip import moments_group
    from .chebyshev import reconstruction
    programs = [_stepwise_program(ops, cfg, B) for ops, cfg, B in items]
    cfgs = [cfg for _, cfg, _ in items]
    requests = [next(p) for p in programs]
    results = [None] * len(programs)
    while any(r is not None for r in requests):
        live = [q for q, r in enumerate(requests) if r is not None]
        buckets = {}
        for q in live:
            ops, B = requests[q]
            cfg = cfgs[q]
            key = (cfg.n_moments, B.shape[1], B.dtype) \
                if ops.metric.identity and not getattr(cfg, "moment_doubling", False) else ("alone", q)
            buckets.setdefault(key, []).append(q)
        answers = {}
        for key, qs in buckets.items():
            if len(qs) == 1:
                ops, B = requests[qs[0]]
                answers[qs[0]] = projected_quadratic(B, ops, cfgs[qs[0]])
                continue
            sc = [scaling(cfgs[q]) for q in qs]
            mus = moments_group([requests[q][0].native for q in qs], [requests[q][1] for q in qs],
                                cfgs[qs[0]].n_moments, [c for c, _ in sc], [a for _, a in sc])
            for q, mu in zip(qs, mus):
                answers[q] = (reconstruction(requests[q][0], cfgs[q]) @ mu).T
        for q in live:
            try:
                requests[q] = programs[q].send(answers[q])
            except StopIteration as done:
                results[q] = done.value
                requests[q] = None
    return results


def iter_stepwise_blocks(ops, config, samples=None):
    from .stochastic import iter_stochastic_blocks
    yield from iter_stochastic_blocks(ops, config, "stepwise", samples)
