"""Drop-in for the reference's compiled module kubo_msd._native
(cpp/native.cpp:613-720), running on the MI355X through libkmsd.so.

The same classes, constructor arguments and methods, with host NumPy arrays
in and out (complex128 blocks (N, nb), row-major), and the same error types:
invalid arguments raise ValueError, a failed metric solve raises
MetricFailure (a RuntimeError) carrying the per-column report as
.diagnostics.  Each call moves its block to the device, runs the library and
copies the result back, so the reference's adapter (src/kubo_msd/backends/
native.py:16-25) binds it by importing this module in place of
kubo_msd._native.  `threads` is accepted and recorded; the device work has no
OpenMP team.  All objects share one HIP context per process (device 0 or
KUBOMSD_DEVICE), created on first use.
"""
import os
import time

import numpy as np
import scipy.sparse as sp

from .backends import hip as _hip
from .errors import KuboMSDError, NumericalError
from .units import HBAR_EV_FS


class MetricFailure(RuntimeError):
    """A metric solve that failed the reference's acceptance (cpp/native.cpp:123-143)."""


_CTX = None


def _ctx():
    global _CTX
    if _CTX is None:
        _CTX = _hip.HipContext(int(os.environ.get("KUBOMSD_DEVICE", "0")))
    return _CTX


def _dev(B):
    a = np.asarray(B)
    if a.ndim != 2 or a.shape[0] < 1 or a.shape[1] < 1:
        raise ValueError("expected nonempty finite (N,batch) array")
    return _ctx().upload(np.ascontiguousarray(a, np.complex128))


def _host(t):
    return _ctx().to_host(t)


def _call(fn, *args):
    """Library errors as the reference's: failed solves -> MetricFailure with
    diagnostics, everything else -> ValueError."""
    try:
        return fn(*args)
    except NumericalError as exc:
        diag = getattr(exc, "diagnostics", None)
        if diag is not None:
            failure = MetricFailure(str(exc))
            failure.diagnostics = diag
            raise failure from exc
        raise ValueError(str(exc)) from exc
    except KuboMSDError as exc:
        raise ValueError(str(exc)) from exc


def build_info():
    info = dict(_hip.build_info())
    info.update(precision="complex128/float64", gil_released=True,
                implementation="MI355X HIP kernels (libkmsd.so) behind the kubo_msd._native interface")
    return info


class CSR:
    """Reference CSR(indptr:int64[n+1], indices:int64[nnz], data:c128[nnz], n, threads)."""

    def __init__(self, indptr, indices, data, n, threads=1):
        ptr = np.asarray(indptr, np.int64)
        idx = np.asarray(indices, np.int64)
        val = np.asarray(data, np.complex128)
        n = int(n)
        if n < 1 or ptr.shape != (n + 1,) or ptr[0] != 0 or ptr[-1] != len(idx) or len(val) != len(idx) or \
                np.any(np.diff(ptr) < 0) or (len(idx) and (idx.min() < 0 or idx.max() >= n)):
            raise ValueError("invalid CSR dimensions/threads")
        if not np.all(np.isfinite(val)):
            raise ValueError("CSR values must be finite")
        self.threads = int(threads)
        self.host = sp.csr_matrix((val, idx, ptr), shape=(n, n))
        self.matrix = _call(_hip.HipMatrix, _ctx(), self.host)
        self.n = n

    def matmul(self, B):
        return _host(_call(self.matrix.matmul, _dev(B)))


class Metric:
    """Reference Metric(S: CSR, diag: f64[n], blocks: [(a, b, L)], threads)."""

    def __init__(self, S, diag, blocks=(), threads=1):
        self.threads = int(threads)
        self.S = S
        self.core = _call(_hip.HipMetric, _ctx(), S.matrix, np.asarray(diag, np.float64),
                          [(int(a), int(b), np.asarray(L, np.complex128)) for a, b, L in blocks])
        self.seconds = 0.0

    def solve(self, B, rtol, atol, maxiter):
        t0 = time.perf_counter()
        try:
            Y, report = _call(self.core.solve, _dev(B), float(rtol), float(atol), int(maxiter))
        finally:
            self.seconds += time.perf_counter() - t0
        return _host(Y), report

    def sample(self, xi, coefficients, center, half):
        c = np.asarray(coefficients, np.float64)
        if c.ndim != 1:
            raise ValueError("sampling coefficients must be 1D")
        return _host(_call(self.core.sample, _dev(xi), c, float(center), float(half)))

    def set_residual_policy(self, policy, interval, gap):
        _call(self.core.set_residual_policy, policy, int(interval), float(gap))

    def set_direct(self, enabled, steps):
        _call(self.core.set_direct, bool(enabled), int(steps))

    def set_index_blocks(self, blocks):
        _call(self.core.set_index_blocks, [(np.asarray(i, np.int64), np.asarray(L, np.complex128))
                                           for i, L in blocks])

    def set_ic0(self, factor, permutation, scaling):
        _call(self.core.set_ic0, factor.host, permutation, scaling)

    def diagnostics(self):
        d = self.core.diagnostics()
        return {"calls": d["calls"], "max_iterations_sum": d["max_iterations_sum"], "seconds": self.seconds,
                "max_relative_residual": d["max_relative_residual"],
                "true_residual_refreshes": d["true_residual_refreshes"],
                "true_residual_checks": d["true_residual_checks"], "residual_restarts": d["residual_restarts"],
                "refinement_steps": d["refinement_steps"], "direct_solve_workers_observed": 1,
                "omp_threads_requested": self.threads, "omp_threads_observed": 1}


class Engine:
    """Reference Engine(H: CSR, metric: Metric, [(kind, [CSR...]) x 3], rtol,
    atol, maxiter, hbar)."""

    def __init__(self, H, metric, definitions, rtol, atol, maxiter, hbar=HBAR_EV_FS):
        definitions = list(definitions)
        if len(definitions) != 3:
            raise ValueError("invalid engine operators")
        self.metric = metric
        self.core = _call(_hip.HipEngine, _ctx(), H.matrix, metric.core,
                          [(int(k), [m.matrix for m in mats]) for k, mats in definitions], float(rtol),
                          float(atol), int(maxiter), float(hbar))

    @property
    def h_calls(self):
        return self.core.h_calls

    @property
    def v_calls(self):
        return self.core.v_calls

    def h_action(self, B):
        return _host(_call(self.core.h_action, _dev(B)))

    def v_cov(self, B, direction):
        return _host(_call(self.core.v_cov, _dev(B), np.asarray(direction, np.float64)))

    def v_cov_with_controls(self, B, direction, rtol, atol, maxiter):
        return _host(_call(self.core.v_cov_with_controls, _dev(B), np.asarray(direction, np.float64), rtol, atol,
                           maxiter))

    def series(self, B, coefficients, center, half, directions=()):
        c = np.asarray(coefficients, np.complex128)
        if c.ndim != 1:
            raise ValueError("Chebyshev coefficients must be 1D")
        return _host(_call(self.core.series, _dev(B), c, float(center), float(half), list(directions)))

    def chain(self, B, length, center, half):
        return _host(_call(self.core.chain, _dev(B), int(length), float(center), float(half)))

    def moments(self, B, length, center, half):
        return _host(_call(self.core.moments, _dev(B), int(length), float(center), float(half)))


def ic0_factor(a, shift):
    """Reference ic0_factor(CSR, shift): the IC(0) values on a's lower pattern."""
    return _call(_hip.ic0_factor, _ctx(), a.host, float(shift))
