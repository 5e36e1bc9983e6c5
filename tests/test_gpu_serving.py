"""Continuous batching: jobs admitted by slots and HBM budget overlap on the
device (one worker process per slot), and every job's result is bitwise
identical alone and in the mix (deterministic reductions); mixed-precision
D(E) stays near the FP64 run."""
from dataclasses import replace

import numpy as np
import pytest

from kubomsd_amd.serving import TransportServer, synthetic_trace

pytestmark = pytest.mark.gpu


def trace(precision="mixed"):
    return synthetic_trace(6, rate_per_s=50.0, n_range=(800, 6000), moments=(64, 128), vectors=(4, 8),
                           nonorthogonal_fraction=0.34, precision=precision, times=(0.0, 0.5, 1.0), seed=11)


def test_alone_equals_mix_and_overlap():
    """Each job's DOS and D(E) are bitwise the same run alone (one slot, jobs
    one after another) and in a 4-slot mix where jobs overlap on the device;
    slots are worker processes with their own device contexts."""
    jobs = trace()
    for j in jobs:
        j.arrival_s = 0.0
    with TransportServer(slots=1, hbm_budget_bytes=8 << 30, group_max=1) as srv:
        alone = srv.run(jobs, realtime=False)
        assert srv.peak_running == 1
    with TransportServer(slots=4, hbm_budget_bytes=8 << 30) as srv:
        mix = srv.run(trace(), realtime=False)
        assert srv.peak_running >= 2, "jobs never overlapped"
    assert any(j.meta["nonorthogonal"] for j in jobs)
    assert len(mix) == len(alone) == len(jobs)
    for jid, r in mix.items():
        assert r.error is None and alone[jid].error is None, (r.error, alone[jid].error)
        np.testing.assert_array_equal(r.dos, alone[jid].dos)
        np.testing.assert_array_equal(r.diffusion, alone[jid].diffusion)


def test_grouped_jobs_equal_alone():
    """Small orthogonal jobs of one shape share launches when a slot takes
    several at once (group_max): each job's DOS and D(E) are bitwise those it
    gets alone; the nonorthogonal ones in the trace run alone in the mix."""
    def grouped_trace():
        jobs = synthetic_trace(8, rate_per_s=50.0, n_range=(800, 6000), moments=(64,), vectors=(4,),
                               nonorthogonal_fraction=0.25, precision="mixed", times=(0.0, 0.5, 1.0), seed=12)
        for j in jobs:
            j.arrival_s = 0.0
        return jobs
    jobs = grouped_trace()
    assert sum(j.group_key is not None for j in jobs) >= 4
    with TransportServer(slots=1, hbm_budget_bytes=8 << 30, group_max=1) as srv:
        alone = srv.run(jobs, realtime=False)
        assert not srv.groups_formed
    with TransportServer(slots=2, hbm_budget_bytes=8 << 30, group_max=8) as srv:
        mix = srv.run(grouped_trace(), realtime=False)
        assert srv.groups_formed and max(len(g) for g in srv.groups_formed) >= 2
    assert max(r.shared for r in mix.values()) >= 2, "no job ran in shared launches"
    assert all(r.shared == 1 for r in alone.values())
    for jid, r in mix.items():
        assert r.error is None and alone[jid].error is None, (r.error, alone[jid].error)
        np.testing.assert_array_equal(r.dos, alone[jid].dos)
        np.testing.assert_array_equal(r.diffusion, alone[jid].diffusion)


def test_mixed_precision_diffusion_vs_fp64():
    with TransportServer(slots=3, hbm_budget_bytes=8 << 30) as srv:
        mixed = srv.run(trace("mixed"), realtime=False)
    with TransportServer(slots=3, hbm_budget_bytes=8 << 30) as srv:
        fp64 = srv.run(trace("fp64"), realtime=False)
    for jid in mixed:
        a, b = mixed[jid].diffusion, fp64[jid].diffusion
        assert mixed[jid].error is None and fp64[jid].error is None
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        assert err < 2e-7, (jid, err)  # measured 5.5e-9 .. 1.6e-8 on this trace


def test_budget_admission_rejects_oversized_job():
    from kubomsd_amd.errors import ResourceBlocked
    jobs = trace()
    srv = TransportServer(slots=1, hbm_budget_bytes=1 << 20)
    with pytest.raises(ResourceBlocked, match="RESOURCE_BLOCKED"):
        srv.submit(jobs[0])


def trace_job(n, nonorthogonal, seed, moments=128, vectors=64):
    """One job of the serving trace's construction, with two explicit sample
    columns so the oracle can run it."""
    from helpers import phases
    job = synthetic_trace(1, n_range=(n, n), moments=(moments,), vectors=(vectors,),
                          nonorthogonal_fraction=1.0 if nonorthogonal else 0.0, seed=seed)[0]
    job.samples = phases(job.meta["n"], 2, seed)
    return job


def oracle_block_sums(job, bounds):
    """The oracle's stepwise per-sample DOS and MSD numerators for a job,
    summed over its statistics blocks."""
    from oracle import kpm
    m = job.model()
    kcfg = kpm.config(energy_eV=job.cfg.energy_eV, time_fs=job.cfg.time_fs, spectral_bounds=bounds,
                      n_moments=job.cfg.n_moments, segment_fs=job.cfg.segment_fs)
    ops = kpm.Ops(m.H_cov, m.S, m.velocity, rtol=1e-13, threads=8)
    dos, nums = kpm.stepwise_samples(ops, kcfg, job.samples)
    sb = job.cfg.statistics_block
    ns = job.samples.shape[1]
    blocks = [slice(i, min(i + sb, ns)) for i in range(0, ns, sb)]
    return (np.array([dos[s].sum(0) for s in blocks]),
            np.array([nums["msd_A2/commutator"][s].sum(0) for s in blocks]))


def check_job_vs_oracle(job, out, tol=3e-7):
    first = out["blocks"][0]
    edos, enum = oracle_block_sums(job, tuple(first.metadata["spectral_report"]["bounds"]))
    got_dos = first.static["dos_block_sums"]
    got_num = np.concatenate([b.arrays["numerator_block_sums/msd_A2/commutator"] for b in out["blocks"]], axis=2)
    assert np.max(np.abs(got_dos - edos)) <= tol * np.max(np.abs(edos))
    assert np.max(np.abs(got_num - enum)) <= tol * np.max(np.abs(enum))


def test_jobs_of_each_size_class_match_oracle(hip_ctx):
    """One serving job per size class, run as the slots run them (a grouped
    pair of small orthogonal flakes, a small nonorthogonal flake, a 2e5-orbital
    flake in the large class): per-block DOS and MSD numerator sums match the
    oracle's stepwise orchestration on the same model and samples (mixed
    precision, within 3e-7)."""
    from kubomsd_amd.serving import run_job, run_job_group
    srv = TransportServer.__new__(TransportServer)
    srv.class_edges = (1 << 26, 1 << 30)
    small = [trace_job(1e4, False, 31, vectors=4), trace_job(1.2e4, False, 32, vectors=4)]
    assert small[0].group_key == small[1].group_key is not None
    outs = run_job_group(small, hip_ctx, keep_blocks=True)
    for job, (out, err) in zip(small, outs):
        assert err is None and out["shared"] == 2
        check_job_vs_oracle(job, out)
    nonorth = trace_job(2e4, True, 33, vectors=16)
    large = trace_job(2e5, False, 34, moments=64, vectors=64)
    assert srv.size_class(nonorth.est_bytes) == 1 and srv.size_class(large.est_bytes) == 2
    for job in (nonorth, large):
        check_job_vs_oracle(job, run_job(job, hip_ctx, keep_blocks=True))
