"""Continuous-batching transport service on one MI355X.

Requests (transport jobs of different sizes) arrive over time.  The server
keeps ``slots`` worker PROCESSES, each with its own HIP context, stream and
kmsd context on the device; a job is admitted when a slot is free AND its
estimated device footprint fits the HBM budget left by the jobs already on the
device.  Admitted jobs run concurrently: their kernels interleave on the GPU
through the processes' own hardware queues, so small jobs that cannot fill
256 CUs alone overlap with each other and with large ones, and each slot's
host-side orchestration runs in its own interpreter (no shared interpreter
lock between slots).  Waiting jobs are kept in size classes (by estimated
bytes of one Chebyshev step); the scheduler takes the oldest waiting job
whose footprint fits, scanning classes smallest-first when the oldest does not
fit, so a large job never blocks small ones that fit (and is guaranteed to run
once enough of the budget drains, because no new work is admitted past a job
that has waited longer than ``max_wait_s``).

The parent process only schedules: one dispatcher thread per slot takes an
admitted job, sends it to its worker process over a pipe and waits for the
result; it makes no device calls.  A worker that dies is reported as the
failure of the job it held and is replaced.

Every kernel on the path reduces deterministically (fixed-order partial sums,
launch shapes fixed by the job alone: jobs run with the default shapes, no
timing-based tuning), so a job's result is bitwise the same alone and in any
mix; tests/test_gpu_serving.py checks it.

The reference has no serving layer; its closest analogue is the GPU edition's
per-run admission (require_bytes, batch halving: kubo_msd_gpu/src/kubo_msd/
backends/gpu_context.py:128-136, solvers/stochastic_gpu.py:31-52), which the
per-job budget check here generalizes to many jobs.
"""
from collections import deque
from dataclasses import dataclass, field
import logging
import threading
import time

import numpy as np

from .errors import ResourceBlocked

log = logging.getLogger(__name__)


@dataclass
class TransportJob:
    """One request: a model (or a callable that builds it), a RunConfig and
    optional explicit samples.  method: 'stepwise' or 'spectral'."""
    job_id: int
    model: object
    cfg: object
    method: str = "stepwise"
    samples: object = None
    arrival_s: float = 0.0
    est_bytes: int = 0
    size_class: int = 0
    meta: dict = field(default_factory=dict)
    group_key: object = None  # jobs with equal keys may share launches (None: always alone)
    est_work: float = 0.0  # relative service-time estimate (orders jobs that arrive together)


@dataclass
class JobResult:
    job_id: int
    arrival_s: float
    start_s: float
    finish_s: float
    slot: int
    dos: np.ndarray = None
    diffusion: np.ndarray = None
    blocks: list = None
    error: str = None
    shared: int = 1  # jobs whose moment chains ran in shared launches with this one's (1: alone)

    @property
    def latency_s(self):
        return self.finish_s - self.arrival_s


def estimate_job_bytes(n, nnz, nb, na, precision, nonorthogonal, n_velocity=1):
    """Device bytes of a stepwise job: matrices (FP64 + FP32 values, int32
    indices, for H, S and the velocities) and the vector blocks of one batch."""
    es = 16 if precision == "fp64" else 8
    mats = (2 + n_velocity) * nnz * (8 + 4 + 4) + 64 * n
    w = nb * (1 + na)
    vec = n * es * (4 * w + 3 * na * nb + (7 * w + 12 * nb if nonorthogonal else 0))
    return int(1.15 * (mats + vec)) + (64 << 20)


class TransportServer:
    def __init__(self, device=0, slots=4, hbm_budget_bytes=None, class_edges=(1 << 26, 1 << 30),
                 max_wait_s=30.0, keep_blocks=False, on_result=None, group_max=8):
        """on_result: optional callable(JobResult), called as each job finishes.
        group_max: most jobs of one group key a slot takes at once (their
        moment chains share launches); 1 disables grouping."""
        self.on_result = on_result
        self.group_max = max(1, int(group_max))
        self.groups_formed = []
        self.device = device
        self.slots = slots
        self.procs = []
        if hbm_budget_bytes is None:
            import torch
            free, _ = torch.cuda.mem_get_info(device)
            hbm_budget_bytes = int(0.85 * free)
        self.budget = int(hbm_budget_bytes)
        self.used = 0
        self.class_edges = tuple(class_edges)
        self.queues = [deque() for _ in range(len(class_edges) + 1)]
        self.max_wait_s = max_wait_s
        self.keep_blocks = keep_blocks
        self.cv = threading.Condition()
        self.results = {}
        self.pending = 0
        self.closed = False
        self.t0 = None
        self.peak_used = 0
        self.peak_running = 0
        self.running = 0

    # -- admission -------------------------------------------------------------
    def size_class(self, est_bytes):
        return int(np.searchsorted(self.class_edges, est_bytes, side="right"))

    def submit(self, job):
        if job.est_bytes > self.budget:
            raise ResourceBlocked(f"RESOURCE_BLOCKED: job {job.job_id} needs {job.est_bytes} bytes, "
                                  f"budget {self.budget}")
        job.size_class = self.size_class(job.est_bytes)
        with self.cv:
            self.queues[job.size_class].append(job)
            self.pending += 1
            self.cv.notify_all()

    def _pick(self, now, longest=False):
        heads = [q[0] for q in self.queues if q]
        if not heads:
            return None
        # The oldest arrival first.  Among jobs that arrived together, slot 0
        # (longest=True) takes the one with the most estimated work and the
        # other slots the one with the least: the longest jobs start at once
        # on one slot while the short ones stream through the others, so a
        # batch arriving at once neither waits behind its longest jobs nor
        # finishes with them (first in queue order on equal estimates).
        t_old = min(j.arrival_s for j in heads)
        oldest = None
        for q in self.queues:
            for j in q:
                if j.arrival_s != t_old:
                    continue
                if oldest is None or (j.est_work > oldest.est_work if longest else j.est_work < oldest.est_work):
                    oldest = j
        free = self.budget - self.used
        if oldest.est_bytes <= free:
            return oldest
        if now - oldest.arrival_s > self.max_wait_s:
            return None  # drain: let the starving job get the budget next
        for q in self.queues:  # smallest classes first
            for job in q:
                if job.est_bytes <= free:
                    return job
        return None

    def _admit(self, job):
        self.queues[job.size_class].remove(job)
        self.pending -= 1
        self.used += job.est_bytes
        self.running += 1
        self.peak_used = max(self.peak_used, self.used)
        self.peak_running = max(self.peak_running, self.running)

    def _take(self, slot=1):
        """The next admitted work for a slot: a list of one job, or of up to
        group_max waiting jobs with the first one's group key (in arrival
        order, each within the budget left)."""
        with self.cv:
            while True:
                if self.closed and self.pending == 0:
                    return None
                job = self._pick(self.clock(), longest=slot == 0 and self.slots > 1)
                if job is not None:
                    self._admit(job)
                    group = [job]
                    if job.group_key is not None and self.group_max > 1:
                        waiting = sorted((j for q in self.queues for j in q if j.group_key == job.group_key),
                                         key=lambda j: j.arrival_s)
                        for other in waiting:
                            if len(group) == self.group_max:
                                break
                            if other.est_bytes <= self.budget - self.used:
                                self._admit(other)
                                group.append(other)
                    if len(group) > 1:
                        self.groups_formed.append([j.job_id for j in group])
                    return group
                self.cv.wait(0.05)

    def _release(self, job):
        with self.cv:
            self.used -= job.est_bytes
            self.running -= 1
            self.cv.notify_all()

    def clock(self):
        return time.perf_counter() - self.t0

    # -- execution -------------------------------------------------------------
    MAX_SLOTS = 15  # worker processes on the device, plus this one

    def _launch(self, slot):
        import multiprocessing as mp
        ctx = mp.get_context("spawn")  # a fresh interpreter: no device state is inherited
        parent, child = ctx.Pipe()
        proc = ctx.Process(target=_worker_main, args=(child, self.device, self.budget, self.keep_blocks),
                           name=f"kubomsd-slot-{slot}", daemon=True)
        proc.start()
        child.close()
        return proc, parent

    @staticmethod
    def _ready(slot, proc, conn):
        try:
            ok = conn.recv() == "ready"  # the worker's device context exists
        except EOFError:
            ok = False
        if not ok:
            proc.join(5)
            raise RuntimeError(f"serving slot {slot} failed to start (exit code {proc.exitcode})")

    def _spawn(self, slot):
        proc, conn = self._launch(slot)
        self._ready(slot, proc, conn)
        return proc, conn

    def _replace(self, slot):
        """Ends the slot's worker and starts a new one; False if that fails."""
        proc, conn = self.procs[slot]
        proc.join(5)
        if proc.is_alive():
            proc.kill()
            proc.join()
        conn.close()
        try:
            self.procs[slot] = self._spawn(slot)
        except Exception:
            return False
        return True

    def start(self):
        """Starts the worker processes, all at once (once; run() calls it)."""
        if self.procs:
            return
        if not 1 <= self.slots <= self.MAX_SLOTS:
            raise ValueError(f"slots must be in [1, {self.MAX_SLOTS}]")
        procs = [self._launch(s) for s in range(self.slots)]
        try:
            for s, (proc, conn) in enumerate(procs):
                self._ready(s, proc, conn)
        except Exception:
            self.procs = procs
            self.close()
            raise
        self.procs = procs

    def close(self):
        """Stops the worker processes."""
        for proc, conn in self.procs:
            try:
                conn.send(None)
            except (OSError, ValueError):
                pass
        for proc, conn in self.procs:
            proc.join(30)
            if proc.is_alive():
                proc.kill()
                proc.join()
            conn.close()
        self.procs = []

    def __enter__(self):
        self.start()
        return self

    def __exit__(self, *exc):
        self.close()

    def _dispatch(self, slot):
        while True:
            group = self._take(slot)
            if group is None:
                return
            start = self.clock()
            outs = [({}, None)] * len(group)
            lost = False
            for attempt in (0, 1):
                proc, conn = self.procs[slot]
                try:
                    conn.send(group)
                except (EOFError, OSError) as exc:  # the worker had ended before this work
                    outs = [({}, f"worker process unavailable ({type(exc).__name__})")] * len(group)
                    lost = not self._replace(slot)
                    if lost:
                        break
                    continue  # one retry on a fresh worker
                except Exception as exc:  # the jobs cannot be sent (e.g. not picklable)
                    outs = [({}, f"job could not be sent to a worker: {type(exc).__name__}: {exc}")] * len(group)
                    break
                try:
                    outs = conn.recv()
                except (EOFError, OSError) as exc:  # the worker died running this work
                    lost = not self._replace(slot)
                    outs = [({}, f"worker process ended ({type(exc).__name__}, exit code {proc.exitcode})")] * \
                        len(group)
                break
            finish = self.clock()
            for job, (res, err) in zip(group, outs):
                self._release(job)
                result = JobResult(job.job_id, job.arrival_s, start, finish, slot, res.get("dos"),
                                   res.get("diffusion"), res.get("blocks"), err, res.get("shared", 1))
                self.results[job.job_id] = result
                if self.on_result is not None:
                    self.on_result(result)
            if lost:  # no worker could be started for this slot: the other slots keep serving
                return

    def run(self, jobs, realtime=True):
        """Feeds jobs at their arrival times (seconds from start) and returns
        {job_id: JobResult} once all have finished.  The worker processes
        are started first (not timed) and kept for later runs until close()."""
        self.start()
        with self.cv:
            self.closed = False
        self.t0 = time.perf_counter()
        ordered = sorted(jobs, key=lambda j: j.arrival_s)
        if not realtime:  # every job is waiting before the first slot takes work
            for job in ordered:
                self.submit(job)
            ordered = []
        threads = [threading.Thread(target=self._dispatch, args=(s,), daemon=True) for s in range(self.slots)]
        for t in threads:
            t.start()
        for job in ordered:
            delay = job.arrival_s - self.clock()
            if delay > 0:
                time.sleep(delay)
            self.submit(job)
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        for t in threads:
            t.join()
        # Jobs still queued had no slot left (every worker failed to restart).
        with self.cv:
            for q in self.queues:
                while q:
                    job = q.popleft()
                    self.pending -= 1
                    now = self.clock()
                    self.results[job.job_id] = JobResult(job.job_id, job.arrival_s, now, now, -1,
                                                         error="no serving slot could take the job")
        return self.results

    @staticmethod
    def summary(results, wall_s):
        lat = np.array([r.latency_s for r in results.values() if r.error is None])
        failed = [r.job_id for r in results.values() if r.error is not None]
        return dict(jobs=len(results), failed=failed, wall_s=wall_s,
                    jobs_per_s=len(lat) / wall_s if wall_s > 0 else None,
                    p50_latency_s=float(np.percentile(lat, 50)) if len(lat) else None,
                    p99_latency_s=float(np.percentile(lat, 99)) if len(lat) else None,
                    mean_latency_s=float(lat.mean()) if len(lat) else None)


def _worker_main(conn, device, budget, keep_blocks):
    """One serving slot: a process with its own HIP context, running the work
    it receives (a list of jobs, through run_jobs; None ends it)."""
    import torch
    from .backends.hip import HipContext
    torch.cuda.set_device(device)
    ctx = HipContext(device, hbm_budget_bytes=budget)
    conn.send("ready")
    while True:
        group = conn.recv()
        if group is None:
            break
        conn.send(run_jobs(group, ctx, keep_blocks))
    conn.close()


def run_jobs(group, ctx, keep_blocks=False):
    """[(result, error)] for a list of jobs on one context: one job alone, or
    several through run_job_group; errors are reported per job."""
    if len(group) > 1:
        try:
            return run_job_group(group, ctx, keep_blocks)
        except Exception as exc:  # noqa: BLE001 -- the jobs are then run one by one, each reporting its own error
            log.warning("shared launches for jobs %s failed (%s: %s); running them one by one",
                        [j.job_id for j in group], type(exc).__name__, exc)
    outs = []
    for job in group:
        try:
            outs.append((run_job(job, ctx, keep_blocks), None))
        except Exception as exc:  # reported per job, the slot keeps serving
            outs.append(({}, f"{type(exc).__name__}: {exc}"))
    return outs


class FlakeModel:
    """Builds one synthetic serving job's model on demand (a picklable
    callable, so the job travels to a worker process as a few numbers)."""

    def __init__(self, n_sites, W, nonorthogonal, seed):
        self.n_sites, self.W, self.nonorthogonal, self.seed = n_sites, W, nonorthogonal, seed

    def __call__(self):
        from .contracts import ModelBundle
        from .synthetic import honeycomb_flake
        m = honeycomb_flake(self.n_sites, W=self.W, s_scale=0.08 if self.nonorthogonal else 0.0, seed=self.seed)
        return ModelBundle(m.H, m.S, velocity={"x": m.velocity["x"]})


def _prepare(job, ctx):
    """(ops, bounds report) of a job: its operators on the slot's context and
    its spectral bounds, resolved as iter_stochastic_blocks resolves them."""
    from .metric.bounds import certified_bounds
    from .metric.device_bounds import estimate_spectral_bounds_device
    from .operators.velocity import build_operators
    model = job.model() if callable(job.model) else job.model
    ops = build_operators(model, job.cfg, ctx)
    if job.cfg.spectral_bounds is None:
        report = certified_bounds(ops.host_H, ops.host_S, ops.metric.s_lower)
    else:
        report = estimate_spectral_bounds_device(ops, job.cfg.spectral_bounds, ops.metric.bounds_report["certified"])
    return ops, report


def _result(blocks, keep_blocks):
    first = blocks[0]
    key = "numerator_diffusion_A2_per_fs/msd_gradient"
    diff = np.concatenate([b.arrays[key] for b in blocks], axis=1) if key in first.arrays else None
    out = {"dos": first.static["dos_states_per_eV"], "diffusion": diff}
    if keep_blocks:
        out["blocks"] = blocks
    return out


def run_job(job, ctx, keep_blocks=False):
    """Builds the job's operators on the slot's context and runs the
    stochastic estimator; returns DOS and the diffusion estimate D(E, t)."""
    from .solvers.stochastic import iter_stochastic_blocks
    ops, report = _prepare(job, ctx)
    blocks = list(iter_stochastic_blocks(ops, job.cfg, job.method, samples=job.samples, bounds_report=report))
    del ops
    return _result(blocks, keep_blocks)


def run_job_group(jobs, ctx, keep_blocks=False):
    """Several stepwise jobs on one context whose moment chains share launches
    (solvers.stepwise.stepwise_samples_group): each job's samples form one
    batch, computed together with the other jobs' batches, then its
    statistics.  A job that cannot join (nonorthogonal metric, VAC path,
    batch not admitted whole) runs alone.  Every job's result is bit for bit
    what run_job gives it.  Returns [(result, error)] in the jobs' order; an
    error in the shared part propagates (run_jobs then runs the jobs one by
    one)."""
    from dataclasses import replace
    import torch
    from .metric.sampling import sample_metric
    from .solvers.stepwise import stepwise_samples_group
    from .solvers.stochastic import admitted_batch, iter_stochastic_blocks
    outs = [None] * len(jobs)
    members = []  # (position, job, ops, report, samples)
    for q, job in enumerate(jobs):
        ops, report = _prepare(job, ctx)
        cfg = replace(job.cfg, spectral_bounds=report["bounds"])
        ns = cfg.nsamples if job.samples is None else job.samples.shape[1]
        if (job.method != "stepwise" or cfg.include_vac or not ops.metric.identity
                or admitted_batch(ops, cfg, "stepwise") < ns):
            blocks = list(iter_stochastic_blocks(ops, job.cfg, job.method, samples=job.samples, bounds_report=report))
            outs[q] = (_result(blocks, keep_blocks), None)
            continue
        B = sample_metric(cfg.seed, np.arange(ns), ops, cfg) if job.samples is None else \
            ops.asarray(job.samples, torch.complex128)
        members.append((q, job, ops, report, cfg, B))
    shared = stepwise_samples_group([(ops, cfg, B) for _, _, ops, _, cfg, B in members])
    for (q, job, ops, report, cfg, B), res in zip(members, shared):
        ns = B.shape[1]
        try:
            blocks = list(iter_stochastic_blocks(ops, job.cfg, "stepwise", samples=job.samples, bounds_report=report,
                                                 batch_results={(0, ns): res}))
            outs[q] = ({**_result(blocks, keep_blocks), "shared": len(members)}, None)
        except Exception as exc:  # this job's statistics failed; the others stand
            outs[q] = ({}, f"{type(exc).__name__}: {exc}")
    return outs


def synthetic_trace(n_jobs=256, rate_per_s=4.0, n_range=(1e4, 4e6), moments=(256, 512, 1024, 2048, 4096),
                    vectors=(4, 16, 64), nonorthogonal_fraction=0.2, precision="mixed", times=(0.0, 0.5, 1.0),
                    seed=20261208, moment_doubling=False, group_elements=1 << 20):
    """The serving workload: Poisson arrivals; honeycomb flakes with N
    log-uniform in n_range, W ~ U[0.5, 3] eV, M and the vector count drawn from
    the given sets, a fraction with a nonorthogonal overlap; each job is a
    stepwise transport run (DOS + MSD + D(E) at `times`) on one batch.
    Orthogonal jobs with at most group_elements orbitals x vectors get a
    group key (M, vectors, times, precision), so a slot may run several of
    them with shared launches."""
    from .config import RunConfig
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_per_s, n_jobs))
    arrivals -= arrivals[0]
    jobs = []
    for i in range(n_jobs):
        n = float(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        W = float(rng.uniform(0.5, 3.0))
        M = int(rng.choice(moments))
        nb = int(rng.choice(vectors))
        nonorth = bool(rng.random() < nonorthogonal_fraction)
        L = max(2, int(round(np.sqrt(n / 2))))
        n_act = 2 * L * L

        build = FlakeModel(2 * L * L, W, nonorth, seed + 17 * i)
        cfg = RunConfig(energy_eV=np.linspace(-7.5, 7.5, 61), time_fs=list(times), solver="stochastic-kpm",
                        kernel="jackson", backend="hip", n_moments=M, axes=("x",), nsamples=nb, sample_batch=nb,
                        statistics_block=max(1, nb // 4), include_vac=False, precision=precision,
                        metric_rtol=1e-10, seed=20261019 + i, autotune=False, moment_doubling=moment_doubling)
        est = estimate_job_bytes(n_act, 4 * n_act, nb, 1, precision, nonorth)
        # launch-bound jobs (a Chebyshev order moves under ~32 MB) share launches
        # with others of the same shape
        key = (M, nb, tuple(times), precision) if not nonorth and n_act * nb <= group_elements else None
        # service-time estimate: orbitals x orders x vectors, a PCG solve per order costing
        # about ten products when the overlap is nonorthogonal
        work = float(n_act) * M * nb * (10.0 if nonorth else 1.0)
        jobs.append(TransportJob(i, build, cfg, "stepwise", None, float(arrivals[i]), est,
                                 meta=dict(n=n_act, M=M, vectors=nb, W=W, nonorthogonal=nonorth), group_key=key,
                                 est_work=work))
    return jobs
