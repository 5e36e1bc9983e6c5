"""Admission logic of the transport server (host only, no device calls)."""
import numpy as np

from kubomsd_amd.serving import TransportJob, TransportServer, estimate_job_bytes, synthetic_trace


class _Srv(TransportServer):
    def __init__(self, budget):  # no torch device needed for admission
        self.budget, self.used, self.class_edges = budget, 0, (100, 1000)
        self.queues = [__import__("collections").deque() for _ in range(3)]
        self.max_wait_s, self.cv = 5.0, __import__("threading").Condition()
        self.pending, self.closed, self.running, self.peak_used, self.peak_running = 0, False, 0, 0, 0
        self.t0 = 0.0
        self.group_max, self.groups_formed = 8, []
        self.slots = 4

    def clock(self):
        return self.now


def job(i, b, t, key=None):
    return TransportJob(i, None, None, arrival_s=t, est_bytes=b, group_key=key)


def test_oldest_fitting_job_first_and_small_jobs_pass_a_blocked_large_one():
    s = _Srv(1000)
    s.now = 1.0
    for j in (job(0, 900, 0.0), job(1, 50, 0.1), job(2, 40, 0.2), job(3, 700, 0.3)):
        s.submit(j)
    assert s._take()[0].job_id == 0          # oldest fits: admitted
    assert s._pick(1.0).job_id == 1       # 900 used: small ones pass the waiting 700
    assert s._take()[0].job_id == 1
    assert s._take()[0].job_id == 2
    assert s.used == 990
    # 700 does not fit until the big one releases
    assert s._pick(1.0) is None
    s._release(job(0, 900, 0.0))
    assert s._take()[0].job_id == 3


def test_jobs_arriving_together_split_longest_and_shortest_first():
    """Among jobs with the same arrival time slot 0 takes the one with the
    most estimated work and the other slots the one with the least (queue
    order on equal estimates); distinct arrival times keep first-come order
    whatever the estimates."""
    s = _Srv(1000)
    s.now = 1.0
    jobs = [TransportJob(i, None, None, arrival_s=0.0, est_bytes=100, est_work=w)
            for i, w in enumerate((1.0, 5.0, 3.0, 5.0, 1.0))]
    for j in jobs:
        s.submit(j)
    assert [s._take(0)[0].job_id, s._take(1)[0].job_id, s._take(2)[0].job_id, s._take(0)[0].job_id,
            s._take(1)[0].job_id] == [1, 0, 4, 3, 2]
    s = _Srv(1000)
    s.now = 1.0
    for j in (TransportJob(0, None, None, arrival_s=0.0, est_bytes=100, est_work=9.0),
              TransportJob(1, None, None, arrival_s=0.1, est_bytes=100, est_work=1.0)):
        s.submit(j)
    assert [s._take(1)[0].job_id for _ in range(2)] == [0, 1]
    # a large job that does not fit does not block smaller ones of the same arrival
    s = _Srv(1000)
    s.now = 1.0
    s.used = 500
    for j in (TransportJob(0, None, None, arrival_s=0.0, est_bytes=100, est_work=1.0),
              TransportJob(1, None, None, arrival_s=0.0, est_bytes=900, est_work=9.0)):
        s.submit(j)
    assert s._take(0)[0].job_id == 0
    # one slot: first in queue order among equal arrivals of equal estimates
    s = _Srv(1000)
    s.slots = 1
    s.now = 1.0
    for j in (TransportJob(0, None, None, arrival_s=0.0, est_bytes=100, est_work=2.0),
              TransportJob(1, None, None, arrival_s=0.0, est_bytes=100, est_work=1.0)):
        s.submit(j)
    assert s._take(0)[0].job_id == 1
    # the synthetic trace estimates work from the job shape
    trace = synthetic_trace(8, rate_per_s=8.0, n_range=(1e3, 1e4), seed=3)
    for j in trace:
        f = 10.0 if j.meta["nonorthogonal"] else 1.0
        assert j.est_work == j.meta["n"] * j.meta["M"] * j.meta["vectors"] * f


def test_starving_job_blocks_new_admissions():
    s = _Srv(1000)
    s.submit(job(0, 800, 0.0))
    s.now = 0.5
    assert s._take()[0].job_id == 0
    s.submit(job(1, 900, 1.0))  # cannot fit while job 0 runs
    s.submit(job(2, 10, 2.0))
    s.now = 2.0
    assert s._pick(s.now).job_id == 2   # not yet starving: small job passes
    s.now = 10.0
    assert s._pick(s.now) is None       # job 1 waited > max_wait_s: drain for it


def test_size_classes_and_trace_shape():
    s = _Srv(10 ** 12)
    assert [s.size_class(b) for b in (50, 100, 500, 5000)] == [0, 1, 1, 2]
    jobs = synthetic_trace(32, rate_per_s=8.0, n_range=(1e3, 1e4), seed=3)
    assert len(jobs) == 32 and jobs[0].arrival_s == 0.0
    assert np.all(np.diff([j.arrival_s for j in jobs]) >= 0)
    ns = [j.meta["n"] for j in jobs]
    assert 400 <= min(ns) and max(ns) <= 2 * 10 ** 4
    assert {j.meta["M"] for j in jobs} <= {256, 512, 1024, 2048, 4096}
    assert estimate_job_bytes(10 ** 6, 4 * 10 ** 6, 64, 1, "mixed", True) > \
        estimate_job_bytes(10 ** 6, 4 * 10 ** 6, 64, 1, "mixed", False)


def test_jobs_travel_to_worker_processes():
    """Serving jobs are sent to slot processes by pickling: the synthetic
    trace's jobs (model builders, configs) survive a round trip and build the
    same model."""
    import pickle
    jobs = synthetic_trace(3, rate_per_s=8.0, n_range=(1e3, 2e3), nonorthogonal_fraction=0.5, seed=4)
    for j in jobs:
        k = pickle.loads(pickle.dumps(j))
        assert k.job_id == j.job_id and k.cfg.n_moments == j.cfg.n_moments
        np.testing.assert_array_equal(k.cfg.energy_eV, j.cfg.energy_eV)
        a, b = j.model(), k.model()
        assert (a.H_cov != b.H_cov).nnz == 0 and (a.S != b.S).nnz == 0


class _FakeConn:
    def __init__(self, script):
        self.script = list(script)  # per send: "ok", "dead", "unpicklable", "dies"
        self.action = None

    def send(self, group):
        self.group = group
        self.action = self.script.pop(0)
        if self.action == "dead":
            raise BrokenPipeError("worker gone")
        if self.action == "unpicklable":
            raise TypeError("cannot pickle")

    def recv(self):
        if self.action == "dies":
            raise EOFError
        return [({"dos": "dos-of-job"}, None) for _ in self.group]

    def close(self):
        pass


class _FakeProc:
    exitcode = -11

    def join(self, timeout=None):
        pass

    def is_alive(self):
        return False

    def kill(self):
        pass


class _DispatchSrv(TransportServer):
    """Real scheduling and dispatch over fake worker pipes (no processes)."""

    def __init__(self, scripts, respawn_ok=True):
        super().__init__(slots=1, hbm_budget_bytes=10 ** 9)
        self.scripts = list(scripts)
        self.respawn_ok = respawn_ok
        self.spawned = 0

    def _spawn(self, slot):
        if self.spawned and not self.respawn_ok:
            raise RuntimeError("cannot start")
        self.spawned += 1
        return _FakeProc(), _FakeConn(self.scripts.pop(0))

    def start(self):
        if not self.procs:
            self.procs = [self._spawn(0)]


def test_dispatch_error_paths_release_budget_and_record_every_job():
    jobs = [TransportJob(i, None, None, arrival_s=0.0, est_bytes=100) for i in range(4)]
    # worker 1: job 0 ok, job 1 unpicklable, job 2: pipe dead -> respawn
    # worker 2: job 2 retried ok, job 3: worker dies while running it -> respawn
    srv = _DispatchSrv([["ok", "unpicklable", "dead"], ["ok", "dies"], []])
    res = srv.run(jobs, realtime=False)
    assert sorted(res) == [0, 1, 2, 3]
    assert res[0].error is None and res[0].dos == "dos-of-job"
    assert "could not be sent" in res[1].error
    assert res[2].error is None and res[2].dos == "dos-of-job"
    assert "exit code -11" in res[3].error
    assert srv.used == 0 and srv.spawned == 3


def test_lost_slot_reports_remaining_jobs():
    jobs = [TransportJob(i, None, None, arrival_s=0.0, est_bytes=100) for i in range(3)]
    srv = _DispatchSrv([["dies"]], respawn_ok=False)
    res = srv.run(jobs, realtime=False)
    assert sorted(res) == [0, 1, 2]
    assert "exit code" in res[0].error
    assert all("no serving slot" in res[j].error for j in (1, 2))
    assert srv.used == 100 * 0 and srv.pending == 0


def test_jobs_of_one_group_key_are_taken_together():
    """A slot takes the oldest admissible job and then, up to group_max, the
    waiting jobs with its group key in arrival order, each within the budget
    left; keyless jobs and other keys stay queued."""
    s = _Srv(1000)
    s.now = 1.0
    for j in (job(0, 100, 0.0, "a"), job(1, 100, 0.1, "b"), job(2, 100, 0.2, "a"), job(3, 100, 0.3),
              job(4, 700, 0.4, "a"), job(5, 100, 0.5, "a")):
        s.submit(j)
    s.group_max = 3
    assert [j.job_id for j in s._take()] == [0, 2, 4]
    assert s.used == 900 and s.running == 3
    # job 5 (key "a") no longer fits next to the admitted ones; "b" and the keyless job go alone
    assert [j.job_id for j in s._take()] == [1]
    s._release(job(4, 700, 0.4))
    assert [j.job_id for j in s._take()] == [3]
    assert [j.job_id for j in s._take()] == [5]
    assert s.groups_formed == [[0, 2, 4]]
    s.group_max = 1
    for j in (job(6, 10, 0.6, "c"), job(7, 10, 0.7, "c")):
        s.submit(j)
    assert [j.job_id for j in s._take()] == [6]


def test_grouped_dispatch_reports_each_job():
    jobs = [TransportJob(i, None, None, arrival_s=0.0, est_bytes=100, group_key="k") for i in range(3)]
    srv = _DispatchSrv([["ok"], []])
    res = srv.run(jobs, realtime=False)
    assert sorted(res) == [0, 1, 2] and all(res[j].error is None for j in range(3))
    assert srv.groups_formed == [[0, 1, 2]] and srv.used == 0
    # a worker that dies with a group reports the error on every member
    srv = _DispatchSrv([["dies"], []])
    res = srv.run(jobs, realtime=False)
    assert all("exit code -11" in res[j].error for j in range(3)) and srv.used == 0


def test_trace_group_keys():
    jobs = synthetic_trace(40, rate_per_s=8.0, n_range=(1e3, 1e4), nonorthogonal_fraction=0.3, seed=5)
    for j in jobs:
        if j.meta["nonorthogonal"] or j.meta["n"] * j.meta["vectors"] > 1 << 20:
            assert j.group_key is None
        else:
            assert j.group_key == (j.meta["M"], j.meta["vectors"], tuple(j.cfg.time_fs), j.cfg.precision)
    assert any(j.group_key is not None for j in jobs)
