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
    p.add_argument("--vectors", type=int, default=64)
    p.add_argument("--moments", type=int, default=2048)
    p.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    p.add_argument("--no-fp64-check", action="store_true")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=5.0, help="target seconds per CPU baseline run")
    p.add_argument("--groups", type=int, default=0, help="override workgroups per launch")
    p.add_argument("--cpl", type=int, default=0)
    p.add_argument("--serving", action="store_true",
                   help="run the continuous-batching serving trace instead of the chain benchmark")
    p.add_argument("--jobs", type=int, default=64)
    p.add_argument("--rate", type=float, default=8.0, help="Poisson arrival rate (jobs/s)")
    p.add_argument("--n-max", type=float, default=4e5, help="largest job size (orbitals); reference: 4e6")
    p.add_argument("--slots", type=int, default=6)
    p.add_argument("--check-jobs", type=int, default=6, help="jobs re-run alone in FP64 for the D(E) error")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only run a plain 1 GiB device copy (torch copy kernel), to scale HBM counters")
    return p.parse_args()


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
    jobs = synthetic_trace(args.jobs, rate_per_s=args.rate, n_range=(1e4, args.n_max))
    srv = TransportServer(slots=args.slots)
    t0 = time.perf_counter()
    res = srv.run(jobs, realtime=True)
    wall = time.perf_counter() - t0
    stats = TransportServer.summary(res, wall)
    # D(E) error vs FP64: the first check_jobs re-run alone in complex128
    from dataclasses import replace
    errs = []
    check = [j for j in synthetic_trace(args.jobs, rate_per_s=args.rate, n_range=(1e4, args.n_max),
                                        precision="fp64")][:args.check_jobs]
    ref = TransportServer(slots=1).run([replace(j, arrival_s=0.0) for j in check], realtime=False)
    for j in check:
        a, b = res[j.job_id].diffusion, ref[j.job_id].diffusion
        if a is not None and b is not None and np.max(np.abs(b)) > 0:
            errs.append(float(np.max(np.abs(a - b)) / np.max(np.abs(b))))
    work = sum(3 * j.meta["M"] * j.meta["vectors"] for j in jobs)
    line = {"metric": "Chebyshev vector-steps/s & HBM GB/s vs peak; p50/p99 job latency; err vs FP64",
            "value": stats["p99_latency_s"], "unit": "s (p99 job latency)", "n_gpus": 1, "steps": len(jobs),
            "warmup": 0, "ms_per_step": 1e3 * wall / len(jobs), "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "c64+f64", "data": "synthetic (seeded honeycomb flakes, Poisson arrivals)",
            "config": {"workload": f"serving: {len(jobs)} jobs, rate {args.rate}/s, N logU[1e4,{args.n_max:.0e}], "
                                   f"M in 256..4096, vectors in 4/16/64, 20% nonorthogonal, {args.slots} slots",
                       "parallelism": "single GPU, concurrent streams"},
            "serving": {**stats, "peak_concurrent_jobs": srv.peak_running, "peak_budget_bytes": srv.peak_used,
                        "vector_orders_total": work, "diffusion_rel_err_vs_fp64": errs}}
    print(json.dumps(line))


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


def cpu_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return max(1, min(aff, int(env))) if env else aff


def cpu_baseline(model, B, c, a, target_s):
    """Times the first orders of the FP64 DOS moment chain on the host cores
    (median of 3 after a warm-up, as benchmarks/benchmark_native.py:68-75)."""
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
        metric = ref.Metric(S, model.S.diagonal().real.copy(), [], threads)
        eng = ref.Engine(csr(model.H), metric, [(-1, []), (-1, []), (-1, [])], 1e-10, 0.0, 1000,
                         0.6582119569509067)
        kind = "reference"
    else:
        metric = O.Metric(O.Csr(model.S, threads), model.S.diagonal().real, (), threads)
        eng = O.Engine(O.Csr(model.H, threads), metric, [(-1, []), (-1, []), (-1, [])], 1e-10, 0.0, 1000)
        kind = "port"
    t0 = time.perf_counter()
    eng.moments(HB, 2, c, a)  # warm-up: one H action (SpMM + metric PCG)
    one = time.perf_counter() - t0
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
    if args.groups:
        ctx.set_tuning("groups", args.groups)
    if args.cpl:
        ctx.set_tuning("cpl", args.cpl)
    t_build = time.perf_counter()
    model = honeycomb_lattice(args.cells, t=-2.7, W=1.0, a=1.42)
    n, nb, M = model.n, args.vectors, args.moments
    lo, hi = -8.8, 8.8
    c, a = (lo + hi) / 2, (hi - lo) / 2
    Bh = random_phases(n, SEED, range(nb))
    metric = HipMetric(ctx, model.S, model.S.diagonal().real)
    eng = HipEngine(ctx, model.H, metric, [(-1, []), (-1, []), (-1, [])], 1e-10, 0.0, 1000)
    dt = torch.complex64 if args.precision == "mixed" else torch.complex128
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
    e2, e3 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    mu_d = eng.moments(B, M, c, a, doubling=True)
    e2.record(stream)
    for _ in range(args.steps):
        mu_d = eng.moments(B, M, c, a, doubling=True)
    e3.record(stream)
    torch.cuda.synchronize()
    dbl_ms = e2.elapsed_time(e3) / args.steps
    dbl_orders = (M - 1) // 2 + 1

    err = None
    if not args.no_fp64_check and args.precision == "mixed":
        B64 = torch.from_numpy(Bh).to(ctx.torch_device)
        mu64 = eng.moments(B64, M, c, a).cpu().numpy()
        mu32 = mu.cpu().numpy()
        from kubomsd_amd.kernels.energy import reconstruction_matrix
        G = reconstruction_matrix(np.linspace(-8.0, 8.0, 161), (lo, hi), M)
        dos64, dos32 = G @ mu64.mean(1), G @ mu32.mean(1)
        mud = mu_d.cpu().numpy()
        dosd = G @ mud.mean(1)
        err = dict(mu_max_rel=float(np.max(np.abs(mu32 - mu64)) / np.max(np.abs(mu64))),
                   dos_max_rel=float(np.max(np.abs(dos32 - dos64)) / np.max(np.abs(dos64))),
                   doubling_mu_max_rel=float(np.max(np.abs(mud - mu64)) / np.max(np.abs(mu64))),
                   doubling_dos_max_rel=float(np.max(np.abs(dosd - dos64)) / np.max(np.abs(dos64))))
        del B64
    cpu = None
    if not args.no_cpu_baseline:
        cb_nb = 8
        cpu = cpu_baseline(model, Bh[:, :cb_nb], c, a, args.cpu_seconds)

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
                   "precision": args.precision, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": per_launch,
                     "kernel": "step_kernel (fused CSR product + 3-term update + moment partials)"},
        "moments_per_s": nb * M / (ms_per_step / 1e3),
        "moment_doubling": {"moments_per_s": nb * M / (dbl_ms / 1e3), "ms_per_chain": dbl_ms,
                            "orders": dbl_orders,
                            "achieved_GBps": step_bytes(model.H.nnz, n, nb, s_val, s_vec, d=0) * dbl_orders
                            / (dbl_ms / 1e3) / 1e9},
        "err_vs_fp64": err,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
