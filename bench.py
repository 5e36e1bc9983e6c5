#!/usr/bin/env python3
"""Chebyshev KPM moment-chain throughput on one MI355X.

Workload (BASELINE configuration "Batched FP32/FP64-accum"): honeycomb lattice
1000 x 1000 cells (N = 2,000,000 orbitals), t = -2.7 eV, Anderson disorder
U[-0.5, 0.5] eV, a = 1.42 A, periodic, CSR with 3 off-diagonal nnz per row
plus the diagonal; 64 unit-modulus random-phase vectors (NumPy PCG64
SeedSequence([20261019, id]) phases, the reference's generator) in one block;
M = 2048 Jackson moments; spectral bounds (-8.8, 8.8) eV.

One step = one full moment chain mu_l = Re<B, T_l(H~) B>, l < 2048, over the
64-vector block: 2047 fused launches (CSR product + three-term update + moment
partials) plus partial reductions.  Vectors are complex64 in HBM with FP64
arithmetic in registers and FP64 reductions ("mixed"); an FP64 (complex128)
chain on the same input gives the error vs FP64.

The JSON line also carries the roofline of the dominant kernel (the fused
step) from HIP events on the library's stream, and the CPU baseline: the
reference's own CPU kernel (oracle/_ref, built from cpp/native.cpp) or, if that
build is absent, the FP64 C++ restatement (oracle/libkpm_oracle.so), timed on a
bounded slice of the same workload on the host cores.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np

HBM_PEAK_GBS = 8000.0
SEED = 20261019


def parse():
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5, help="timed moment chains")
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--cells", type=int, default=1000, help="honeycomb cells per side")
    p.add_argument("--vectors", type=int, default=None, help="default 64 (honeycomb), 32 (nonorthogonal)")
    p.add_argument("--moments", type=int, default=None, help="default 2048 (honeycomb), 1024 (nonorthogonal)")
    p.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    p.add_argument("--no-fp64-check", action="store_true")
    p.add_argument("--no-doubling", action="store_true",
                   help="skip the moment-doubling chains (profiles of the direct recursion only)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-plain-baseline", action="store_true",
                   help="skip the FP64 GPU comparison paths (fused complex128 chain, unfused torch path)")
    p.add_argument("--cpu-seconds", type=float, default=5.0, help="target seconds per CPU baseline run")
    p.add_argument("--groups", type=int, default=0, help="override workgroups per launch")
    p.add_argument("--cpl", type=int, default=0)
    p.add_argument("--pcg-poll", type=int, default=0, help="override PCG iterations between host polls")
    p.add_argument("--min-blocks", type=int, default=0,
                   help="override the least row blocks per workgroup of row-walking launches")
    p.add_argument("--tile-bytes", type=int, default=None,
                   help="nonorthogonal: gathered-block bytes per tile of the plane-tiled 4x4-block walk "
                        "(0: natural row order; default: the library's)")
    p.add_argument("--sample-batch", type=int, default=0,
                   help="square: samples per batch (0: all vectors in one batch)")
    p.add_argument("--block-width-autotune", action="store_true",
                   help="square: let the estimator pick the batch width from measured step times, starting "
                        "from --sample-batch")
    p.add_argument("--no-bsr-mfma", action="store_true",
                   help="nonorthogonal: FP32-valued 4x4-block products on the VALU kernel instead of the "
                        "matrix cores")
    p.add_argument("--moment-doubling", action="store_true",
                   help="serving/stepwise: orthogonal jobs take moments from the product identities "
                        "(errors are still reported against the FP64 three-term recursion)")
    p.add_argument("--serving", action="store_true",
                   help="run the continuous-batching serving trace instead of the chain benchmark")
    p.add_argument("--jobs", type=int, default=64)
    p.add_argument("--rate", type=float, default=8.0,
                   help="Poisson arrival rate (jobs/s); 0: every job waiting at the start (service capacity)")
    p.add_argument("--n-max", type=float, default=4e5, help="largest job size (orbitals); reference: 4e6")
    p.add_argument("--slots", type=int, default=6)
    p.add_argument("--group-max", type=int, default=8,
                   help="serving: most jobs of one shape a slot runs with shared launches (1: none)")
    p.add_argument("--check-jobs", type=int, default=6, help="jobs re-run alone in FP64 for the D(E) error")
    p.add_argument("--tmax", type=float, default=5.0, help="stepwise: last output time (fs); reference: 50")
    p.add_argument("--dt", type=float, default=1.0, help="stepwise: output time step (fs)")
    p.add_argument("--cpu-vectors", type=int, default=2,
                   help="square: vectors of the CPU reference run (the full estimator on the first ones)")
    p.add_argument("--method", choices=["stepwise", "spectral"], default="stepwise",
                   help="square: the stochastic estimator (spectral: VAC by segmented Chebyshev products, "
                        "the mu tables on the FP64 matrix cores)")
    p.add_argument("--workload", choices=["honeycomb", "nonorthogonal", "mu", "stepwise", "square", "flake"],
                   default="honeycomb",
                   help="honeycomb: orthogonal moment chain (default); nonorthogonal: simple-cubic 64^3 x 4 "
                        "orbitals, S^-1 H chain with batched block-Jacobi PCG; mu: square 2048^2 double-moment "
                        "tables mu_mn on the matrix cores; flake: one nonorthogonal serving job alone")
    p.add_argument("--flake-n", type=float, default=17672, help="flake: orbitals (about)")
    p.add_argument("--pcg-min-blocks", type=int, default=0,
                   help="override the least row blocks per workgroup of PCG vector passes")
    p.add_argument("--bcheb-lo-pct", type=int, default=0,
                   help="nonorthogonal: block-Jacobi Chebyshev iteration interval lower end in percent of the smallest Ritz value")
    p.add_argument("--side", type=int, default=2048, help="mu: square lattice side")
    p.add_argument("--samples", type=int, default=32, help="mu: random vectors (one table each)")
    p.add_argument("--atoms", type=int, default=64, help="nonorthogonal: atoms per cube side")
    p.add_argument("--preconditioner", choices=["block-jacobi", "diagonal", "ic0"], default="block-jacobi",
                   help="nonorthogonal: metric preconditioner (ic0: RCM order, shifts up to 1e-2)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only run a plain 1 GiB device copy (torch copy kernel), to scale HBM counters")
    a = p.parse_args()
    nonorth = a.workload == "nonorthogonal"
    a.vectors = a.vectors or (16 if a.workload == "square" else 32 if nonorth else 64)
    a.moments = a.moments or {"honeycomb": 2048, "stepwise": 2048, "square": 512, "flake": 256}.get(a.workload, 1024)
    return a


def calibrate_copy(reps=5):
    """A plain copy of known size: reads and writes 2^30 bytes per launch."""
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30, "ms_per_copy": ms,
                      "copy_GBps_read_plus_write": 2 * (1 << 30) / (ms / 1e3) / 1e9}))


def serving(args):
    """Serving workload: Poisson arrivals of honeycomb-flake transport jobs
    (N log-uniform in [1e4, n_max], W ~ U[0.5,3] eV, M in {256..4096},
    vectors in {4,16,64}, 20% nonorthogonal), admitted by slots and HBM budget."""
    import torch
    from kubomsd_amd.serving import TransportServer, synthetic_trace
    jobs = synthetic_trace(args.jobs, rate_per_s=args.rate if args.rate > 0 else 1.0, n_range=(1e4, args.n_max),
                           moment_doubling=args.moment_doubling)
    if args.rate == 0:
        for j in jobs:
            j.arrival_s = 0.0
    def report(r):
        print(f"# job {r.job_id} done on slot {r.slot}: latency {r.latency_s:.2f} s"
              + (f", shared launches with {r.shared - 1} others" if r.shared > 1 else "")
              + (f", error {r.error}" if r.error else ""), file=sys.stderr, flush=True)
    srv = TransportServer(slots=args.slots, on_result=report, group_max=args.group_max)
    srv.start()  # worker processes up before the clock starts
    t0 = time.perf_counter()
    res = srv.run(jobs, realtime=True)
    wall = time.perf_counter() - t0
    srv.close()
    stats = TransportServer.summary(res, wall)
    # D(E) error vs FP64: the first check_jobs re-run alone in complex128
    from dataclasses import replace
    errs = []
    check = [j for j in synthetic_trace(args.jobs, rate_per_s=args.rate if args.rate > 0 else 1.0, n_range=(1e4, args.n_max),
                                        precision="fp64")][:args.check_jobs]
    with TransportServer(slots=1) as one:
        ref = one.run([replace(j, arrival_s=0.0) for j in check], realtime=False)
    for j in check:
        a, b = res[j.job_id].diffusion, ref[j.job_id].diffusion
        if a is not None and b is not None and np.max(np.abs(b)) > 0:
            errs.append(float(np.max(np.abs(a - b)) / np.max(np.abs(b))))
    work = sum(3 * j.meta["M"] * j.meta["vectors"] for j in jobs)
    arrivals = "all waiting at t = 0" if args.rate == 0 else f"Poisson {args.rate}/s"
    line = {"metric": "Chebyshev vector-steps/s & HBM GB/s vs peak; p50/p99 job latency; err vs FP64",
            "value": stats["p99_latency_s"], "unit": "s (p99 job latency)", "n_gpus": 1, "steps": len(jobs),
            "warmup": 0, "ms_per_step": 1e3 * wall / len(jobs), "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "c64+f64", "data": "synthetic (seeded honeycomb flakes, Poisson arrivals)",
            "config": {"workload": f"serving: {len(jobs)} jobs, arrivals {arrivals}, N logU[1e4,{args.n_max:.0e}], "
                                   f"M in 256..4096, vectors in 4/16/64, 20% nonorthogonal, {args.slots} slots, "
                                   f"up to {args.group_max} jobs per shared launch"
                                   + (", moments by doubling (S = I jobs)" if args.moment_doubling else ""),
                       "parallelism": "single GPU, one worker process per slot"},
            "serving": {**stats, "peak_concurrent_jobs": srv.peak_running, "peak_budget_bytes": srv.peak_used,
                        "groups": len(srv.groups_formed),
                        "jobs_in_shared_launches": sum(len(g) for g in srv.groups_formed),
                        "vector_orders_total": work, "vector_orders_per_s": work / wall,
                        "diffusion_rel_err_vs_fp64": errs}}
    print(json.dumps(line))


def flake(args):
    """One serving job alone: a nonorthogonal honeycomb flake of about
    --flake-n orbitals (S = I + 0.08 U[-1,1] on the bonds, diagonal-
    preconditioned PCG to 1e-10 per order), --vectors vectors, --moments
    moments, stepwise DOS + MSD + D(E) at t = 0, 0.5, 1 fs in mixed precision,
    the job shape of the serving trace.  Reports seconds per job (a first run
    warms up; the median of --steps timed runs)."""
    import torch
    from kubomsd_amd.backends.hip import HipContext
    from kubomsd_amd.config import RunConfig
    from kubomsd_amd.serving import FlakeModel, TransportJob, estimate_job_bytes, run_job
    assert torch.cuda.is_available(), "bench needs a HIP device"
    ctx = HipContext(0)
    if args.min_blocks:
        ctx.set_tuning("min_blocks", args.min_blocks)
    if args.pcg_min_blocks:
        ctx.set_tuning("pcg_min_blocks", args.pcg_min_blocks)
    if args.pcg_poll:
        ctx.set_tuning("pcg_poll", args.pcg_poll)
    L = max(2, int(round(np.sqrt(args.flake_n / 2))))
    n, nb, M = 2 * L * L, args.vectors, args.moments
    cfg = RunConfig(energy_eV=np.linspace(-7.5, 7.5, 61), time_fs=[0.0, 0.5, 1.0], solver="stochastic-kpm",
                    kernel="jackson", backend="hip", n_moments=M, axes=("x",), nsamples=nb, sample_batch=nb,
                    statistics_block=max(1, nb // 4), include_vac=False, precision="mixed", metric_rtol=1e-10,
                    seed=SEED, autotune=False)
    job = TransportJob(0, FlakeModel(n, 1.0, True, 77), cfg, "stepwise", None, 0.0,
                       estimate_job_bytes(n, 4 * n, nb, 1, "mixed", True))
    run_job(job, ctx)
    times = []
    for _ in range(max(1, args.steps)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_job(job, ctx)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    sec = float(np.median(times))
    print(json.dumps({"metric": "seconds per serving job (nonorthogonal flake, alone)", "value": sec, "unit": "s",
                      "n_gpus": 1, "steps": len(times), "warmup": 1, "ms_per_step": 1e3 * sec,
                      "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "c64+f64",
                      "data": "synthetic (seeded honeycomb flake, nonorthogonal overlap)",
                      "config": {"workload": f"nonorthogonal honeycomb flake N={n}, {nb} vectors, M={M}, "
                                             "stepwise t = 0, 0.5, 1 fs, diagonal-preconditioned PCG to 1e-10",
                                 "n": n, "vectors": nb, "moments": M},
                      "runs_s": times}))


# |lambda(S^-1 H)| <= |H|_2 / lambda_min(S) for the seeded simple-cubic model
# (64^3 atoms, 4 orbitals): |H|_2 and lambda_min(S) from host Lanczos (scipy
# eigsh, tol 1e-6 / 1e-4: |H|_2 = 9.6389, lambda_min(S) = 0.31957), widened
# by 2%.  A chain run with bounds that miss the spectrum grows exponentially;
# the bench checks |mu_l| <= mu_0.
CUBIC64_BOUND = 1.02 * 9.6389 / 0.31957


def atom_block_factors(S, norb):
    """Lower Cholesky factors of the norb x norb on-site blocks of S, as the
    reference's block-Jacobi input [(a, b, L)] (metric/setup.py:73-76)."""
    coo = S.tocoo()
    nat = S.shape[0] // norb
    keep = coo.row // norb == coo.col // norb
    blk = np.zeros((nat, norb, norb), np.complex128)
    blk[coo.row[keep] // norb, coo.row[keep] % norb, coo.col[keep] % norb] = coo.data[keep]
    L = np.linalg.cholesky(blk)
    return [(q * norb, (q + 1) * norb, L[q]) for q in range(nat)]


def inner_step_bytes(nnz_s, n, nb, s_val, s_vec):
    """Lower bound of HBM bytes per inner step of a metric solve on an
    nb-column block: the S product (values, indices, row pointers) and four
    block streams.  A PCG iteration moves at least six (read P; read and
    write R; read and write Y; write P), a step of the block-Jacobi Chebyshev
    semi-iteration four (read d_k; read d_(k-1); read M^-1 r; write
    d_(k+1)), so four bounds both."""
    return nnz_s * (s_val + 4) + 4 * (n + 1) + 4 * n * nb * s_vec


def nonorthogonal(args):
    """Simple-cubic 64^3 atoms x 4 orbitals (N = 1,048,576), 4x4 blocks to the
    second shell, nonorthogonal S; 32 vectors; M = 1024 moments of S^-1 H, each
    order one H product plus a block-Jacobi solve of S (FP32 inner solves to
    1e-6 inside FP64 refinement to 1e-10: PCG for the first inner solve, whose
    CG coefficients give the Ritz bounds of the block-Jacobi Chebyshev
    semi-iteration that the later inner solves use).  Reports vector-steps/s
    and inner steps/s (S products of PCG iterations or semi-iteration steps,
    max over columns per solve)."""
    import torch
    from kubomsd_amd.backends.hip import HipContext, HipEngine, HipMetric
    from kubomsd_amd.synthetic import cubic_orbital_model
    assert torch.cuda.is_available(), "bench needs a HIP device"
    ctx = HipContext(0)
    if args.min_blocks:
        ctx.set_tuning("min_blocks", args.min_blocks)
    if args.pcg_poll:
        ctx.set_tuning("pcg_poll", args.pcg_poll)
    if args.tile_bytes is not None:
        ctx.set_tuning("tile_bytes", args.tile_bytes)
    if args.no_bsr_mfma:
        ctx.set_tuning("bsr_mfma", 0)
    if args.bcheb_lo_pct:
        ctx.set_tuning("bcheb_lo_pct", args.bcheb_lo_pct)
    t_build = time.perf_counter()
    model = cubic_orbital_model(args.atoms)
    n, nb, M = model.n, args.vectors, args.moments
    blocks = atom_block_factors(model.S, 4) if args.preconditioner == "block-jacobi" else ()
    metric = HipMetric(ctx, model.S, model.S.diagonal().real, blocks)
    if args.preconditioner == "ic0":
        from kubomsd_amd.metric.preconditioners import prepare_ic0_validated
        import scipy.sparse as sp
        ic = prepare_ic0_validated(ctx, sp.csr_matrix(model.S, dtype=np.complex128), ordering="rcm", shift_max=1e-2)
        metric.set_ic0(ic.factor, ic.permutation, ic.scaling)
        print(f"# IC(0) factor: shift {ic.report['shift']}, {ic.report['seconds']:.1f}s", file=sys.stderr, flush=True)
    eng = HipEngine(ctx, model.H, metric, [(-1, []), (-1, []), (-1, [])], 1e-10, 0.0, 1000)
    # the chain's dominant products: S in the FP32 inner solves
    tuning = apply_tuning(ctx, args, metric.S, nb, torch.complex64, f32_values=True, tile=True)
    bound = CUBIC64_BOUND if args.atoms == 64 and CUBIC64_BOUND else None
    if bound is None:
        import scipy.sparse.linalg as sla
        lmin = sla.eigsh(model.S.real.tocsr(), k=1, which="SA", tol=1e-6, return_eigenvectors=False,
                         v0=np.random.default_rng(1).normal(size=n))[0]
        hmax = abs(sla.eigsh(model.H.real.tocsr(), k=1, which="LM", tol=1e-4, return_eigenvectors=False,
                             v0=np.random.default_rng(2).normal(size=n))[0])
        bound = 1.02 * hmax / lmin
    c, a = 0.0, bound
    Bh = random_phases(n, SEED, range(nb))
    B = torch.from_numpy(Bh).to(torch.complex64).to(ctx.torch_device)
    print(f"# setup {time.perf_counter() - t_build:.1f}s: N={n} nnz(H)={model.H.nnz} nnz(S)={model.S.nnz} nb={nb} "
          f"M={M} bounds=+-{bound:.3f}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        eng.moments(B, 8, c, a)
    torch.cuda.synchronize()
    d0 = metric.diagnostics()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for step in range(args.steps):
        mu = eng.moments(B, M, c, a)  # returns after the chain's final reduction
        print(f"# chain {step + 1}/{args.steps} done", file=sys.stderr, flush=True)
    e1.record(stream)
    torch.cuda.synchronize()
    d1 = metric.diagnostics()
    gpu_ms = e0.elapsed_time(e1)
    ms_per_step = gpu_ms / args.steps
    its = d1["max_iterations_sum"] - d0["max_iterations_sum"]
    solves = d1["calls"] - d0["calls"]
    mu_h = mu.cpu().numpy()
    growth = float(np.max(np.abs(mu_h)) / np.max(np.abs(mu_h[0])))
    print(f"# timed {args.steps} chains: {gpu_ms:.1f} ms, {solves} solves, {its} inner steps, "
          f"max|mu_l|/mu_0 = {growth:.3f}", file=sys.stderr, flush=True)
    if growth > 1.01:
        raise SystemExit("moments grow: spectral bounds do not enclose the spectrum")
    value = nb * (M - 1) * args.steps / (gpu_ms / 1e3)
    s_vec = 8
    step_b = step_bytes(model.H.nnz, n, nb, 4, s_vec)
    pcg_b = inner_step_bytes(model.S.nnz, n, nb, 4, s_vec)
    alg = (M - 1) * args.steps * step_b + its * pcg_b
    achieved = alg / (gpu_ms / 1e3) / 1e9
    err = None
    if not args.no_fp64_check:
        L = min(64, M)
        B64 = torch.from_numpy(Bh).to(ctx.torch_device)
        mu64 = eng.moments(B64, L, c, a).cpu().numpy()
        err = dict(moments_compared=L, mu_max_rel=float(np.max(np.abs(mu_h[:L] - mu64)) / np.max(np.abs(mu64))))
        del B64
    cpu = None
    if not args.no_cpu_baseline:
        # 64 orders of this chain take minutes on the host (each order is a
        # PCG solve): a slice of about --cpu-seconds, 2 vectors
        cpu = cpu_baseline(model, Bh[:, :2], c, a, args.cpu_seconds, blocks=blocks)
    line = {
        "metric": "Chebyshev vector-steps/s & HBM GB/s vs peak; p50/p99 job latency; err vs FP64",
        "value": value, "unit": "vector-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64+f64", "data": "synthetic (seeded simple-cubic orbital model, NumPy PCG64 random phases)",
        "config": {"workload": f"simple cubic {args.atoms}^3 atoms x 4 orbitals (N={n}), nonorthogonal S, "
                               f"{nb} vectors, M={M} moments of S^-1 H, {args.preconditioner} preconditioned FP32 inner "
                               f"solves to 1e-6 (PCG, then the Chebyshev semi-iteration on its Ritz bounds) inside FP64 refinement "
                               f"to 1e-10, one moment chain per step",
                   "n": n, "nnz_H": int(model.H.nnz), "nnz_S": int(model.S.nnz), "vectors": nb, "moments": M,
                   "precision": "mixed", "parallelism": "single GPU", "storage": eng.H.info()["format"],
                   "launch_shape": tuning},
        "pcg": {"solves": solves, "inner_steps": its, "inner_steps_per_solve": its / max(solves, 1),
                "inner_steps_per_s": its / (gpu_ms / 1e3),
                "algorithmic_bytes_per_inner_step": pcg_b},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_chain": alg / args.steps,
                     "kernel": "whole chain: H products, recursion and inner solve steps (lower-bound bytes "
                               "over event time; FP64 refinement passes not counted)"},
        "err_vs_fp64": err,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line))


MFMA_PEAK_TFLOPS = {"fp32": 157.3, "bf16": 2500.0}


def double_moments(args):
    """Square 2048 x 2048 periodic lattice (N = 4,194,304), t = -1 eV, Anderson
    W = 1.5 eV; per random vector r: A = [T_n(H~) V r] and B = [V T_m(H~) r]
    for n, m < M = 1024 (complex64 chains, (M, N) each), then the table
    mu_nm = sum_i conj(A_ni) B_mi on the matrix cores, FP32 MFMA and the bf16x3
    split variant, through the package API
    (kubomsd_amd.solvers.velocity_correlation).  Algorithmic flops per table
    8 N M^2 (complex MAC = 8 real flops).  Errors against FP64: the FP64
    table of the same complex64 chains (all entries, every vector), and the
    full FP64 pipeline (complex128 chains and table) for n, m < 64 of the
    first vector."""
    import torch
    from kubomsd_amd.backends.hip import HipContext
    from kubomsd_amd.config import RunConfig
    from kubomsd_amd.contracts import ModelBundle
    from kubomsd_amd.operators.velocity import build_operators
    from kubomsd_amd.solvers.velocity_correlation import velocity_correlation_moments
    from kubomsd_amd.synthetic import square_lattice
    assert torch.cuda.is_available(), "bench needs a HIP device"
    ctx = HipContext(0)
    if args.min_blocks:
        ctx.set_tuning("min_blocks", args.min_blocks)
    t_build = time.perf_counter()
    model = square_lattice(args.side, t=-1.0, W=1.5)
    n, M = model.n, args.moments
    bound = 4.0 + 0.75 + 0.05  # Gershgorin |E| <= 4|t| + W/2, plus margin
    bundle = ModelBundle(model.H, model.S, velocity={"x": model.velocity["x"]})

    def config(precision, moments):
        return RunConfig(solver="stochastic-kpm", kernel="jackson", backend="hip", energy_eV=[0.0], time_fs=[0.0],
                         axes=("x",), spectral_bounds=(-bound, bound), n_moments=moments, precision=precision,
                         autotune=False)
    cfg = config("mixed", M)
    ops = build_operators(bundle, cfg, ctx)
    flops = 8.0 * n * M * M
    print(f"# setup {time.perf_counter() - t_build:.1f}s: N={n} nnz={model.H.nnz} M={M} samples={args.samples}",
          file=sys.stderr, flush=True)
    totals = {}
    errors = {"fp32": [], "bf16x3": []}
    t_all = time.perf_counter()
    for sid in range(args.samples):
        r = random_phases(n, SEED, [sid])
        out = velocity_correlation_moments(ops, cfg, r, "x", mode=("fp32", "bf16x3"),
                                           reference=not args.no_fp64_check)
        for key, v in out["ms"].items():
            totals[key] = totals.get(key, 0.0) + v
        if out["error_vs_fp64"]:
            for md in errors:
                errors[md].extend(out["error_vs_fp64"][md])
        if sid == 0 and not args.no_fp64_check:
            # full FP64 pipeline for n, m < 64 (chains are prefix-consistent)
            k = 64
            cfg64 = config("fp64", k)
            ops64 = build_operators(bundle, cfg64, ctx)
            full = velocity_correlation_moments(ops64, cfg64, r, "x", mode="fp64")["mu"][0]
            del ops64
            s2 = full.abs().max().item()
            pipe = {md: (out["mu"][md][0, :k, :k] - full).abs().max().item() / s2 for md in errors}
        del out
        print(f"# sample {sid}: fp32 {totals['table_fp32']:.1f} ms, bf16x3 {totals['table_bf16x3']:.1f} ms "
              f"(cumulative)", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t_all
    ns = args.samples
    tf32 = flops * ns / (totals["table_fp32"] / 1e3) / 1e12
    tfb = flops * ns / (totals["table_bf16x3"] / 1e3) / 1e12
    err = None
    if not args.no_fp64_check:
        err = {"gemm_fp32_max_rel": max(errors["fp32"]), "gemm_bf16x3_max_rel": max(errors["bf16x3"]),
               "pipeline_fp32_max_rel_64x64": pipe["fp32"], "pipeline_bf16x3_max_rel_64x64": pipe["bf16x3"],
               "reference": "FP64 table (FP64 matrix cores) of the same complex64 chains, every vector; complex128 "
                            "chains and table for n, m < 64 of the first vector"}
    line = {
        "metric": "Chebyshev vector-steps/s & HBM GB/s vs peak; p50/p99 job latency; err vs FP64",
        "value": tf32, "unit": "TFLOP/s (mu_nm table, FP32 MFMA)", "n_gpus": 1, "steps": ns, "warmup": 0,
        "ms_per_step": totals["table_fp32"] / ns, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 (FP32 MFMA, FP64 split-sum)",
        "data": "synthetic (seeded square lattice, NumPy PCG64 phases)",
        "config": {"workload": f"square {args.side}^2 (N={n}), t=-1 eV, W=1.5 eV, {ns} random vectors, "
                               f"mu_nm {M}x{M} complex per vector (package API velocity_correlation_moments)",
                   "n": n, "moments": M, "samples": ns, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": tf32, "peak": MFMA_PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                     "frac": tf32 / MFMA_PEAK_TFLOPS["fp32"], "traffic": None,
                     "algorithmic_flops_per_launch": flops, "kernel": "mu_f32_kernel (+ split sum)"},
        "bf16x3": {"achieved_TFLOPs": tfb, "mfma_executed_TFLOPs": 3 * tfb,
                   "frac_of_bf16_peak": 3 * tfb / MFMA_PEAK_TFLOPS["bf16"],
                   "ms_per_table": totals["table_bf16x3"] / ns},
        "pipeline_ms_per_vector": {k: v / ns for k, v in totals.items()},
        "wall_s": wall,
        "err_vs_fp64": err,
        "cpu_baseline": None,
    }
    if not args.no_fp64_check:
        line["fp64_table_TFLOPs"] = flops * ns / (totals["table_fp64_reference"] / 1e3) / 1e12
    print(json.dumps(line))


def stepwise(args):
    """The 'Batched FP32/FP64-accum' configuration through the whole stepwise
    estimator (iter_stochastic_blocks, method 'stepwise'): honeycomb 1000 x
    1000 cells, 64 random-phase vectors in one batch (8 statistics blocks of
    8), M = 2048 Jackson moments, axes x and y, output times 0 .. tmax fs
    every dt fs, 161 energies on [-8, 8] eV, bounds (-8.8, 8.8) eV.  Runs
    once with complex64 vectors (mixed) and once in complex128 (fp64) on the
    same samples, and reports the estimator's vector-steps/s and the relative
    error of D(E, t) (diffusion_A2_per_fs/msd_gradient), MSD(E, t) and
    DOS(E) of the mixed run against the FP64 run."""
    import torch
    from kubomsd_amd.backends.hip import HipContext
    from kubomsd_amd.config import RunConfig
    from kubomsd_amd.contracts import ModelBundle
    from kubomsd_amd.metric.bounds import certified_bounds
    from kubomsd_amd.operators.velocity import build_operators
    from kubomsd_amd.solvers.stochastic import iter_stochastic_blocks
    from kubomsd_amd.synthetic import honeycomb_lattice
    assert torch.cuda.is_available(), "bench needs a HIP device"
    import logging
    logging.basicConfig(level=logging.INFO, stream=sys.stderr, format="# %(message)s")
    ctx = HipContext(0)
    if args.min_blocks:
        ctx.set_tuning("min_blocks", args.min_blocks)
    t_build = time.perf_counter()
    model = honeycomb_lattice(args.cells, t=-2.7, W=1.0, a=1.42)
    n, nb, M = model.n, args.vectors, args.moments
    E = np.linspace(-8.0, 8.0, 161)
    T = np.round(np.arange(0.0, args.tmax + 0.5 * args.dt, args.dt), 12)
    samples = random_phases(n, SEED, range(nb))
    bundle = ModelBundle(model.H, model.S, velocity=model.velocity)
    print(f"# setup {time.perf_counter() - t_build:.1f}s: N={n} nb={nb} M={M} times={len(T)} axes=x,y",
          file=sys.stderr, flush=True)
    runs = {}
    for prec in ("mixed", "fp64"):
        cfg = RunConfig(solver="stochastic-kpm", kernel="jackson", backend="hip", stochastic_method="stepwise",
                        energy_eV=E, time_fs=T, axes=("x", "y"), spectral_bounds=(-8.8, 8.8), n_moments=M,
                        nsamples=nb, sample_batch=nb, statistics_block=max(1, nb // 8), include_vac=False,
                        precision=prec, segment_fs=args.dt, autotune=prec == "mixed",
                        moment_doubling=bool(args.moment_doubling and prec == "mixed"))
        ops = build_operators(bundle, cfg, ctx)
        # The explicit bounds are checked against Gershgorin discs (a proof of
        # containment); a Lanczos Ritz check to 1e-10 residuals does not
        # converge in 512 steps on the dense band edges of a 2M-site lattice.
        report = certified_bounds(ops.host_H, ops.host_S, ops.metric.s_lower, spectral_bounds=cfg.spectral_bounds)
        stream = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        blocks = list(iter_stochastic_blocks(ops, cfg, "stepwise", samples=samples, bounds_report=report))
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        cat = lambda key: np.concatenate([b.arrays[key] for b in blocks], axis=1)
        diag = blocks[0].metadata["sample_batch_diagnostics"][0]
        runs[prec] = dict(ms=e0.elapsed_time(e1), wall=wall, D=cat("diffusion_A2_per_fs/msd_gradient"),
                          msd=cat("msd_A2/commutator"), dos=blocks[0].static["dos_states_per_eV"],
                          valid=blocks[0].arrays["valid/commutator"][0, 0], order=diag["max_propagation_order"],
                          segments=diag["segments"], tuning=ops.tuning)
        print(f"# {prec}: {runs[prec]['ms'] / 1e3:.1f} s on HIP events ({wall:.1f} s wall), "
              f"propagation order {diag['max_propagation_order']}", file=sys.stderr, flush=True)
        del ops, blocks
    mx, f64 = runs["mixed"], runs["fp64"]

    def rel(a, b):
        ok = np.isfinite(a) & np.isfinite(b)
        return float(np.max(np.abs(a[ok] - b[ok])) / np.max(np.abs(b[ok])))

    na, nt = 2, len(T)
    # Column-steps: the DOS chain, then per output interval the augmented
    # series (order K, one segment per interval, width nb(1+na)) and the
    # moment chain of both axis blocks (width na*nb).
    chain_steps = (M - 1) * nb + (nt - 1) * (M - 1) * na * nb
    series_steps = mx["segments"] * mx["order"] * nb * (1 + na)
    vsteps = chain_steps + series_steps
    value = vsteps / (mx["ms"] / 1e3)
    chain_bytes = (M - 1) * step_bytes(model.H.nnz, n, nb, 8, 8) + \
        (nt - 1) * (M - 1) * step_bytes(model.H.nnz, n, na * nb, 8, 8)
    achieved = chain_bytes / (mx["ms"] / 1e3) / 1e9
    line = {
        "metric": "Chebyshev vector-steps/s & HBM GB/s vs peak; p50/p99 job latency; err vs FP64",
        "value": value, "unit": "vector-steps/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": mx["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64+f64", "data": "synthetic (seeded honeycomb lattice, NumPy PCG64 random phases)",
        "config": {"workload": f"stepwise MSD: honeycomb {args.cells}x{args.cells} cells (N={n}), {nb} vectors, "
                               f"M={M}, axes x,y, t=0..{args.tmax:g} fs every {args.dt:g} fs, 161 energies"
                               + (" (mixed run: moments by doubling, counted as the M-1 orders they replace)"
                                  if args.moment_doubling else ""),
                   "n": n, "vectors": nb, "moments": M, "times": nt, "precision": "mixed",
                   "parallelism": "single GPU", "launch_shape": mx["tuning"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "whole estimator: algorithmic bytes of its moment-chain steps over its event time "
                               "(series steps and statistics counted in the time only)"},
        "vector_steps": {"moment_chains": chain_steps, "series": series_steps, "propagation_order": mx["order"]},
        "err_vs_fp64": {"D_E_t_max_rel": rel(mx["D"], f64["D"]), "msd_max_rel": rel(mx["msd"], f64["msd"]),
                        "dos_max_rel": rel(mx["dos"], f64["dos"]),
                        "valid_masks_equal": bool(np.array_equal(mx["valid"], f64["valid"]))},
        "fp64_run": {"ms": f64["ms"], "vector_steps_per_s": vsteps / (f64["ms"] / 1e3)},
        "cpu_baseline": None,
    }
    print(json.dumps(line))


class _RefStepwiseOps:
    """The reference's own CPU kernel (oracle/_ref Engine) behind the
    interface the oracle's stepwise orchestration (oracle/kpm.py) drives."""

    def __init__(self, ref, model, threads):
        import scipy.sparse as sp

        def csr(m):
            m = sp.csr_matrix(m, dtype=np.complex128)
            return ref.CSR(m.indptr.astype(np.int64), m.indices.astype(np.int64), m.data, m.shape[0], threads)
        self.metric = ref.Metric(csr(model.S), model.S.diagonal().real.copy(), [], threads)
        defs = [(0, [csr(model.velocity["x"])]), (-1, []), (-1, [])]
        self.eng = ref.Engine(csr(model.H), self.metric, defs, 1e-10, 0.0, 1000, 0.6582119569509067)
        self.native = self
        self.host_S = sp.csr_matrix(model.S, dtype=np.complex128)
        self.n = model.n

    def moments(self, B, length, center, half):
        return self.eng.moments(np.ascontiguousarray(B), int(length), center, half)

    def chain(self, B, length, center, half):
        return self.eng.chain(np.ascontiguousarray(B), int(length), center, half)

    def inner(self, A, B):
        return np.sum(A.conj() * (self.host_S @ B), axis=0)

    def apply_v_cov(self, direction, B):
        return self.eng.v_cov(np.ascontiguousarray(B), [float(v) for v in direction])

    def solve(self, B):
        return self.metric.solve(np.ascontiguousarray(B), 1e-10, 0.0, 1000)[0]

    def series(self, B, coeff, center, half, directions=()):
        return self.eng.series(np.ascontiguousarray(B), np.ascontiguousarray(coeff, np.complex128), center, half,
                               [[float(v) for v in d] for d in directions])


def square(args):
    """The reference's CPU-runnable configuration in full: square 100 x 100
    periodic lattice (N = 10,000), t = -1 eV, Anderson U[-1, 1] eV, bond
    velocity along x, 16 random-phase vectors, M = 512 Jackson moments,
    161 energies on [-4, 4] eV, stepwise MSD at t = 0 .. 10 fs every 0.5 fs,
    bounds (-5.1, 5.1) eV, FP64.  The GPU runs the whole estimator
    (iter_stochastic_blocks) on all vectors; the reference's own CPU kernel
    (oracle/_ref, all host cores) runs the same stepwise estimator on the first
    --cpu-vectors of them, and the per-sample MSD numerators of both are
    compared."""
    import torch
    from kubomsd_amd.backends.hip import HipContext
    from kubomsd_amd.config import RunConfig
    from kubomsd_amd.contracts import ModelBundle
    from kubomsd_amd.metric.bounds import certified_bounds
    from kubomsd_amd.operators.velocity import build_operators
    from kubomsd_amd.solvers.spectral import spectral_samples
    from kubomsd_amd.solvers.stepwise import stepwise_samples
    from kubomsd_amd.solvers.stochastic import iter_stochastic_blocks
    from kubomsd_amd.synthetic import square_lattice
    assert torch.cuda.is_available(), "bench needs a HIP device"
    ctx = HipContext(0)
    if args.min_blocks:
        ctx.set_tuning("min_blocks", args.min_blocks)
    side = args.cells if args.cells != 1000 else 100
    model = square_lattice(side, t=-1.0, W=2.0)
    n, nb, M = model.n, args.vectors, args.moments
    E = np.linspace(-4.0, 4.0, 161)
    T = np.round(np.arange(0.0, 10.0 + 0.25, 0.5), 12)
    samples = random_phases(n, SEED, range(nb))
    method = args.method
    cfg = RunConfig(solver="stochastic-kpm", kernel="jackson", backend="hip", stochastic_method=method,
                    energy_eV=E, time_fs=T, axes=("x",), spectral_bounds=(-5.1, 5.1), n_moments=M, nsamples=nb,
                    sample_batch=args.sample_batch or nb, statistics_block=8, include_vac=False, precision="fp64",
                    autotune=False, block_width_autotune=args.block_width_autotune)
    ops = build_operators(ModelBundle(model.H, model.S, velocity={"x": model.velocity["x"]}), cfg, ctx)
    report = certified_bounds(ops.host_H, ops.host_S, ops.metric.s_lower, spectral_bounds=cfg.spectral_bounds)
    list(iter_stochastic_blocks(ops, cfg, method, samples=samples, bounds_report=report))  # warm-up
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(max(1, args.steps)):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        blocks = list(iter_stochastic_blocks(ops, cfg, method, samples=samples, bounds_report=report))
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    ms = float(np.median(times))
    diag = blocks[0].metadata["sample_batch_diagnostics"][0]
    batch_info = {"sample_batch": cfg.sample_batch, "admitted": blocks[0].metadata["admitted_sample_batch"],
                  "width_tuning": blocks[0].metadata.get("block_width_tuning")}
    K, nt = diag["max_propagation_order"], len(T)
    if method == "stepwise":
        # column-steps: DOS chain, then per interval the augmented series (width 2 nb) and the axis chain
        vsteps = (M - 1) * nb + (nt - 1) * (K * 2 * nb + (M - 1) * nb)
    else:
        # column-steps: DOS chain; per segment the left chain (M + K orders), the velocity chain (K + 1)
        # and, from the second segment on, the anchors' propagation (width 2 nb, K orders)
        seg = diag["segments"]
        vsteps = (M - 1) * nb + seg * ((M + K - 1) * nb + K * nb) + (seg - 1) * K * 2 * nb
    print(f"# GPU: whole estimator {ms:.1f} ms (median of {len(times)}), order {K}", file=sys.stderr, flush=True)

    # Reference CPU kernel on the first cpu_vectors samples, and the GPU's
    # per-sample numerators of the same samples.
    from oracle import kpm
    from oracle import native as O
    cb = min(args.cpu_vectors, nb)
    threads = cpu_threads()
    ref = O.load_reference_native()
    kcfg = kpm.config(energy_eV=E, time_fs=T, spectral_bounds=(-5.1, 5.1), n_moments=M, segment_fs=1.0,
                      axes_vectors=np.eye(3)[:1])
    if ref is not None:
        cops, kind = _RefStepwiseOps(ref, model, threads), "reference"
    else:
        cops, kind = kpm.Ops(model.H, model.S, {"x": model.velocity["x"]}, rtol=1e-10, threads=threads), "port"
    run_cpu = kpm.stepwise_samples if method == "stepwise" else kpm.spectral_samples
    run_gpu = stepwise_samples if method == "stepwise" else spectral_samples
    key = "msd_A2/commutator" if method == "stepwise" else "msd_A2/vac_integral"
    t0 = time.perf_counter()
    cdos, cnums = run_cpu(cops, kcfg, samples[:, :cb])
    cpu_s = time.perf_counter() - t0
    cpu_vsteps = vsteps * cb / nb
    gdos, gnums, _ = run_gpu(ops, cfg, ops.asarray(samples[:, :cb]))
    g = gnums[key].cpu().numpy()
    c = cnums[key]
    parity = {"msd_numerators_max_rel": float(np.max(np.abs(g - c)) / np.max(np.abs(c))),
              "dos_max_rel": float(np.max(np.abs(gdos.cpu().numpy() - cdos)) / np.max(np.abs(cdos))),
              "samples_compared": cb}
    print(f"# CPU ({kind}, {threads} threads): {cpu_s:.1f} s for {cb} vectors", file=sys.stderr, flush=True)
    value = vsteps / (ms / 1e3)
    line = {
        "metric": "Chebyshev vector-steps/s & HBM GB/s vs peak; p50/p99 job latency; err vs FP64",
        "value": value, "unit": "vector-steps/s", "n_gpus": 1, "steps": len(times), "warmup": 1,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (seeded square lattice, NumPy PCG64 random phases)",
        "config": {"workload": f"square {side}x{side} periodic (N={n}), t=-1 eV, W=2 eV, {nb} vectors, M={M}, "
                               + ("stepwise MSD" if method == "stepwise" else
                                  "spectral VAC (segments of 1 fs, integration step 0.01 fs) MSD")
                               + " along x, t=0..10 fs every 0.5 fs, 161 energies, FP64, one whole estimator run "
                                 "per step", "n": n, "vectors": nb, "moments": M, "times": nt,
                   "method": method, "precision": "fp64", "parallelism": "single GPU", "batch": batch_info},
        "roofline": {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                     "traffic": None, "kernel": "launch-bound at this size (N = 10^4: about 2.5 MB per block)"},
        "parity_vs_reference_cpu": parity,
        "cpu_baseline": {"value": cpu_vsteps / cpu_s, "unit": "vector-steps/s", "cores": threads, "kind": kind,
                         "sample": f"the whole {method} estimator on the first {cb} of the {nb} vectors, one run, "
                                   f"OMP threads={threads}"},
    }
    line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line))


def plain_gpu_fp64(H, Bh, c, a, orders, device):
    """The FP64 plain GPU path for comparison: the reference GPU edition's
    unfused moment loop (kubo_msd_gpu/src/kubo_msd/solvers/chebyshev.py:60-71:
    vendor sparse product, separate recursion and dot kernels), here with
    torch's sparse CSR product and elementwise ops in complex128 and S = I (no
    solve).  Returns moments/s over `orders` orders and the moments."""
    import torch
    Hc = H.tocsr()
    Ht = torch.sparse_csr_tensor(torch.from_numpy(Hc.indptr.astype(np.int64)),
                                 torch.from_numpy(Hc.indices.astype(np.int64)),
                                 torch.from_numpy(Hc.data.astype(np.complex128)), size=Hc.shape).to(device)
    B = torch.from_numpy(Bh).to(device)
    nb = B.shape[1]
    mu = torch.empty((orders, nb), dtype=torch.float64, device=device)

    def run():
        prev, curr = B, (Ht @ B - c * B) / a
        mu[0] = (B.conj() * B).sum(0).real
        mu[1] = (B.conj() * curr).sum(0).real
        for l in range(2, orders):
            nxt = 2 * (Ht @ curr - c * curr) / a - prev
            mu[l] = (B.conj() * nxt).sum(0).real
            prev, curr = curr, nxt
    run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    return nb * orders / (ms / 1e3), mu


def apply_tuning(ctx, args, matrix, nb, dtype, f32_values=False, tile=False):
    """Explicit --cpl/--groups, else the library's autotuner (cached); tile:
    also the 4x4-block tile size unless --tile-bytes gives it."""
    if args.cpl or args.groups:
        if args.cpl:
            ctx.set_tuning("cpl", args.cpl)
        if args.groups:
            ctx.set_tuning("groups", args.groups)
        return {"cpl": args.cpl or None, "groups": args.groups or None, "source": "command line"}
    choice = ctx.autotune(matrix, nb, dtype, f32_values=f32_values,
                          tile=tile and getattr(args, "tile_bytes", None) is None)
    choice["source"] = "kmsd_autotune"
    return choice


def random_phases(n, seed, ids):
    """exp(2 pi i u), u = PCG64(SeedSequence([seed, id])).random(n) per column
    (reference src/kubo_msd/metric/sampling.py:8-9)."""
    out = np.empty((n, len(ids)), np.complex128)
    for j, i in enumerate(ids):
        u = np.random.Generator(np.random.PCG64(np.random.SeedSequence([int(seed), int(i)]))).random(n)
        out[:, j] = np.exp(2j * np.pi * u)
    return out


def step_bytes(nnz, n, nb, s_val, s_vec, d=1):
    """Algorithmic HBM bytes of one fused Chebyshev step on an nb-column block:
    matrix values + column indices, row pointers, and (3+d) block streams
    (read T_n, read T_{n-1}, write T_{n+1}, read the projection block)."""
    return nnz * (s_val + 4) + 4 * (n + 1) + (3 + d) * n * nb * s_vec


def measured_traffic(key, **expect):
    """HBM bytes per launch of the dominant kernel measured by rocprofv3
    (profiles/traffic.json, copy-calibrated FETCH_SIZE + WRITE_SIZE) when the
    run's workload is the profiled one, else None."""
    try:
        with open(ROOT / "profiles" / "traffic.json") as f:
            rec = json.load(f).get(key)
    except (OSError, ValueError):
        return None
    if not rec or any(rec.get(k) != v for k, v in expect.items()):
        return None
    return rec["traffic_bytes"]


def cpu_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return max(1, min(aff, int(env))) if env else aff


def cpu_baseline(model, B, c, a, target_s, blocks=(), orders=None):
    """Times the first orders of the FP64 DOS moment chain on the host cores
    (median of 3 after a warm-up, as benchmarks/benchmark_native.py:68-75):
    exactly `orders` orders when given (the stated 64-order slice), else as
    many as take about target_s seconds."""
    threads = cpu_threads()
    os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
    from oracle import native as O
    ref = O.load_reference_native()
    HB = np.ascontiguousarray(B)
    if ref is not None:
        import scipy.sparse as sp
        def csr(m):
            m = sp.csr_matrix(m, dtype=np.complex128)
            return ref.CSR(m.indptr.astype(np.int64), m.indices.astype(np.int64), m.data, m.shape[0], threads)
        S = csr(model.S)
        metric = ref.Metric(S, model.S.diagonal().real.copy(), list(blocks), threads)
        eng = ref.Engine(csr(model.H), metric, [(-1, []), (-1, []), (-1, [])], 1e-10, 0.0, 1000,
                         0.6582119569509067)
        kind = "reference"
    else:
        metric = O.Metric(O.Csr(model.S, threads), model.S.diagonal().real, blocks, threads)
        eng = O.Engine(O.Csr(model.H, threads), metric, [(-1, []), (-1, []), (-1, [])], 1e-10, 0.0, 1000)
        kind = "port"
    t0 = time.perf_counter()
    eng.moments(HB, 2, c, a)  # warm-up: one H action (SpMM + metric PCG)
    one = time.perf_counter() - t0
    if orders is None:
        orders = int(np.clip(target_s / max(one, 1e-3), 3, 256))
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        eng.moments(HB, orders, c, a)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    vs = B.shape[1] * (orders - 1) / med
    return dict(value=vs, unit="vector-steps/s", cores=threads, kind=kind,
                sample=(f"first {orders} orders of the FP64 DOS moment chain (each an H action with the "
                        f"reference's PCG metric solve), N={B.shape[0]}, {B.shape[1]} vectors, median of 3 "
                        f"after warm-up; OMP threads={threads}"))


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.serving:
        serving(args)
        return
    if args.workload == "mu":
        if args.gpus != 1:
            sys.exit("this build runs on one GPU only")
        double_moments(args)
        return
    if args.workload == "nonorthogonal":
        if args.gpus != 1:
            sys.exit("this build runs on one GPU only")
        nonorthogonal(args)
        return
    if args.workload == "stepwise":
        if args.gpus != 1:
            sys.exit("this build runs on one GPU only")
        stepwise(args)
        return
    if args.workload == "square":
        if args.gpus != 1:
            sys.exit("this build runs on one GPU only")
        square(args)
        return
    if args.workload == "flake":
        if args.gpus != 1:
            sys.exit("this build runs on one GPU only")
        flake(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sample sharding is a design "
                                   "note (independent statistics blocks per GPU, one final gather)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from kubomsd_amd.backends.hip import HipContext, HipEngine, HipMetric
    from kubomsd_amd.synthetic import honeycomb_lattice

    assert torch.cuda.is_available(), "bench needs a HIP device"
    ctx = HipContext(0)
    if args.min_blocks:
        ctx.set_tuning("min_blocks", args.min_blocks)
    t_build = time.perf_counter()
    model = honeycomb_lattice(args.cells, t=-2.7, W=1.0, a=1.42)
    n, nb, M = model.n, args.vectors, args.moments
    lo, hi = -8.8, 8.8
    c, a = (lo + hi) / 2, (hi - lo) / 2
    Bh = random_phases(n, SEED, range(nb))
    metric = HipMetric(ctx, model.S, model.S.diagonal().real)
    eng = HipEngine(ctx, model.H, metric, [(-1, []), (-1, []), (-1, [])], 1e-10, 0.0, 1000)
    dt = torch.complex64 if args.precision == "mixed" else torch.complex128
    tuning = apply_tuning(ctx, args, eng.H, nb, dt)
    B = torch.from_numpy(Bh).to(dt).to(ctx.torch_device)
    print(f"# setup {time.perf_counter() - t_build:.1f}s: N={n} nnz={model.H.nnz} nb={nb} M={M} "
          f"{args.precision}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        mu = eng.moments(B, M, c, a)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        mu = eng.moments(B, M, c, a)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    ms_per_step = gpu_ms / args.steps
    vector_steps = nb * (M - 1)
    value = vector_steps * args.steps / (gpu_ms / 1e3)
    print(f"# timed {args.steps} chains: {gpu_ms:.1f} ms on HIP events, {wall * 1e3:.1f} ms wall",
          file=sys.stderr, flush=True)

    s_vec = 8 if args.precision == "mixed" else 16
    s_val = 8  # real FP64 H values (FP64 arithmetic on complex64 storage)
    per_launch = step_bytes(model.H.nnz, n, nb, s_val, s_vec)
    # Event time over the whole chain divided by its fused steps: includes the
    # mu_0 pass and the partial reductions, so it bounds the kernel from above.
    step_s = (ms_per_step / 1e3) / (M - 1)
    achieved = per_launch / step_s / 1e9

    # Same moments from the Chebyshev doubling identities (about half the
    # orders, no projection stream): reported beside the direct recursion.
    dbl_ms, dbl_orders, mu_d = None, (M - 1) // 2 + 1, None
    if not args.no_doubling:
        e2, e3 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        mu_d = eng.moments(B, M, c, a, doubling=True)
        e2.record(stream)
        for _ in range(args.steps):
            mu_d = eng.moments(B, M, c, a, doubling=True)
        e3.record(stream)
        torch.cuda.synchronize()
        dbl_ms = e2.elapsed_time(e3) / args.steps

    err = None
    if not args.no_fp64_check and args.precision == "mixed":
        B64 = torch.from_numpy(Bh).to(ctx.torch_device)
        mu64 = eng.moments(B64, M, c, a).cpu().numpy()
        mu32 = mu.cpu().numpy()
        from kubomsd_amd.kernels.energy import reconstruction_matrix
        G = reconstruction_matrix(np.linspace(-8.0, 8.0, 161), (lo, hi), M)
        dos64, dos32 = G @ mu64.mean(1), G @ mu32.mean(1)
        err = dict(mu_max_rel=float(np.max(np.abs(mu32 - mu64)) / np.max(np.abs(mu64))),
                   dos_max_rel=float(np.max(np.abs(dos32 - dos64)) / np.max(np.abs(dos64))))
        if mu_d is not None:
            mud = mu_d.cpu().numpy()
            dosd = G @ mud.mean(1)
            err.update(doubling_mu_max_rel=float(np.max(np.abs(mud - mu64)) / np.max(np.abs(mu64))),
                       doubling_dos_max_rel=float(np.max(np.abs(dosd - dos64)) / np.max(np.abs(dos64))))
        del B64
    # FP64 paths on the same GPU: this library's fused complex128 chain, and the
    # plain unfused path (vendor sparse product + separate kernels).
    plain = None
    if not args.no_plain_baseline:
        B64 = torch.from_numpy(Bh).to(ctx.torch_device)
        e4, e5 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        mu64f = eng.moments(B64, 256, c, a)
        e4.record(stream)
        mu64f = eng.moments(B64, 256, c, a)
        e5.record(stream)
        torch.cuda.synchronize()
        fused64 = nb * 256 / (e4.elapsed_time(e5) / 1e3)
        del B64
        try:
            plain_rate, mu_plain = plain_gpu_fp64(model.H, Bh, c, a, 256, ctx.torch_device)
            agree = float((mu_plain - mu64f).abs().max() / mu64f.abs().max())
            plain = {"moments_per_s": plain_rate, "orders_timed": 256, "max_rel_diff_vs_fused_fp64": agree,
                     "fused_fp64_moments_per_s": fused64,
                     "mixed_direct_over_plain": (nb * M / (ms_per_step / 1e3)) / plain_rate,
                     "mixed_doubling_over_plain": (nb * M / (dbl_ms / 1e3)) / plain_rate if dbl_ms else None,
                     "path": "torch sparse CSR complex128 product + elementwise recursion + dot per order"}
        except RuntimeError as exc:  # the vendor sparse product may not take complex128
            plain = {"error": str(exc)[:200], "fused_fp64_moments_per_s": fused64}
    cpu = None
    if not args.no_cpu_baseline:
        # the stated slice: the first 64 orders of the DOS chain, on 2 of the
        # 64 vectors (about 10-15 s per run on 16 host cores)
        cpu = cpu_baseline(model, Bh[:, :2], c, a, args.cpu_seconds, orders=64)

    line = {
        "metric": "Chebyshev vector-steps/s & HBM GB/s vs peak; p50/p99 job latency; err vs FP64",
        "value": value,
        "unit": "vector-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+f64" if args.precision == "mixed" else "c128",
        "data": "synthetic (seeded honeycomb lattice, NumPy PCG64 random phases)",
        "config": {"workload": f"honeycomb {args.cells}x{args.cells} cells (N={n}), t=-2.7 eV, W=1 eV, "
                               f"{nb} random-phase vectors, M={M} Jackson moments, one moment chain per step",
                   "n": n, "nnz": int(model.H.nnz), "vectors": nb, "moments": M,
                   "precision": args.precision, "parallelism": "single GPU", "storage": eng.H.info()["format"],
                   "launch_shape": tuning},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("ell_step_kernel_honeycomb_n2e6_nb64_c64", n=n, vectors=nb,
                                                 precision=args.precision),
                     "algorithmic_bytes_per_launch": per_launch,
                     "kernel": "ell_step_kernel (fused ELL product + 3-term update + moment partials)"
                     if eng.H.info()["format"] == "csr" else "bsr4_step_kernel"},
        "moments_per_s": nb * M / (ms_per_step / 1e3),
        "moment_doubling": None if dbl_ms is None else
        {"moments_per_s": nb * M / (dbl_ms / 1e3), "ms_per_chain": dbl_ms, "orders": dbl_orders,
         "achieved_GBps": step_bytes(model.H.nnz, n, nb, s_val, s_vec, d=0) * dbl_orders / (dbl_ms / 1e3) / 1e9},
        "err_vs_fp64": err,
        "fp64_gpu_baselines": plain,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
