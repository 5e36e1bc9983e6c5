"""ctypes binding of the FP64 C++ oracle (oracle/kpm_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.

The classes mirror the reference's pybind11 module ``kubo_msd._native``
(cpp/native.cpp:613-720): ``Csr.matmul``, ``Metric.solve/sample``,
``Engine.h_action/v_cov/series/chain/moments``.  Errors become
``OracleError`` (a ValueError) with per-column ``diagnostics`` on metric
failures, as ``backends/native.py:32-39`` does for the reference.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np
import scipy.sparse as sp

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libkpm_oracle.so"
HBAR_EV_FS = 0.6582119569509067

_lib = None


class OracleError(ValueError):
    pass


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise OracleError(f"oracle library missing: run make -C {HERE}")
        lib = C.CDLL(str(LIB_PATH))
        P, I64, D, I = C.c_void_p, C.c_int64, C.c_double, C.c_int
        lib.ko_last_error.restype = C.c_char_p
        lib.ko_csr_create.restype = P
        lib.ko_csr_create.argtypes = [I64, I64, P, P, P, I]
        lib.ko_csr_destroy.argtypes = [P]
        lib.ko_csr_matmul.argtypes = [P, P, P, I64, I64]
        lib.ko_metric_create.restype = P
        lib.ko_metric_create.argtypes = [P, P, I64, P, P, I]
        lib.ko_metric_destroy.argtypes = [P]
        lib.ko_metric_set_residual_policy.argtypes = [P, I, I, D]
        lib.ko_metric_solve.argtypes = [P, P, P, I64, I64, D, D, I, P, P, P, P, P]
        lib.ko_metric_sample.argtypes = [P, P, P, I64, I64, P, I64, D, D]
        lib.ko_metric_counters.argtypes = [P, P, P]
        lib.ko_metric_set_index_blocks.argtypes = [P, I64, P, P, P]
        lib.ko_metric_set_ic0.argtypes = [P, P, P, P, P, P]
        lib.ko_metric_set_direct.argtypes = [P, I, I]
        lib.ko_metric_refinement_steps.argtypes = [P]
        lib.ko_metric_refinement_steps.restype = I64
        lib.ko_ic0_factor.argtypes = [I64, P, P, P, D, P]
        lib.ko_engine_create.restype = P
        lib.ko_engine_create.argtypes = [P, P, P, P, D, D, I, D]
        lib.ko_engine_destroy.argtypes = [P]
        lib.ko_engine_set_controls.argtypes = [P, D, D, I]
        lib.ko_engine_counters.argtypes = [P, P, P]
        lib.ko_engine_last_report.argtypes = [P, P, P, P, P, P]
        lib.ko_h_action.argtypes = [P, P, P, I64, I64]
        lib.ko_v_cov.argtypes = [P, P, P, P, I64, I64]
        lib.ko_series.argtypes = [P, P, P, I64, I64, P, I64, D, D, P, I64]
        lib.ko_chain.argtypes = [P, P, P, I64, I64, I64, D, D]
        lib.ko_moments.argtypes = [P, P, P, I64, I64, I64, D, D]
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _block(B):
    B = np.ascontiguousarray(B, dtype=np.complex128)
    if B.ndim != 2:
        raise OracleError("expected nonempty finite (N,batch) array")
    return B


def _check(rc, report=None):
    if rc:
        err = OracleError(load().ko_last_error().decode())
        if report is not None:
            err.diagnostics = report
        raise err


def _report(b):
    return dict(converged=np.zeros(b, np.uint8), breakdown=np.zeros(b, np.uint8),
                iterations=np.zeros(b, np.int64), true_residual=np.zeros(b),
                relative_residual=np.zeros(b))


def _report_args(r):
    return [_ptr(r[k]) for k in ("converged", "breakdown", "iterations", "true_residual", "relative_residual")]


def _report_out(r):
    return dict(converged=[bool(x) for x in r["converged"]], breakdown=[bool(x) for x in r["breakdown"]],
                iterations=r["iterations"].tolist(), true_residual=r["true_residual"].tolist(),
                relative_residual=r["relative_residual"].tolist())


class Csr:
    """Sorted, duplicate-summed complex128 CSR (backends/native.py:42-46)."""

    def __init__(self, matrix, threads=1):
        lib = load()
        m = sp.csr_matrix(matrix, dtype=np.complex128, copy=True)
        m.sum_duplicates()
        m.sort_indices()
        self.n = m.shape[0]
        self._keep = (np.ascontiguousarray(m.indptr, np.int64), np.ascontiguousarray(m.indices, np.int64),
                      np.ascontiguousarray(m.data, np.complex128))
        self.h = lib.ko_csr_create(self.n, len(m.data), *[_ptr(a) for a in self._keep], threads)
        if not self.h:
            raise OracleError(lib.ko_last_error().decode())
        self.matrix = m

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ko_csr_destroy(self.h)

    def matmul(self, B):
        B = _block(B)
        out = np.empty_like(B)
        _check(load().ko_csr_matmul(self.h, _ptr(B), _ptr(out), *B.shape))
        return out


class Metric:
    """Batched PCG over S with diagonal or contiguous block-Jacobi factors."""

    def __init__(self, S, diagonal, blocks=(), threads=1):
        lib = load()
        self.S = S if isinstance(S, Csr) else Csr(S, threads)
        diag = np.ascontiguousarray(diagonal, np.float64)
        bptr = np.array([0] + [b for (_, b, _) in blocks] if blocks else [0], np.int64)
        fac = (np.concatenate([np.ascontiguousarray(L, np.complex128).ravel() for (_, _, L) in blocks])
               if blocks else np.zeros(1, np.complex128))
        self._keep = (diag, bptr, fac)
        self.h = lib.ko_metric_create(self.S.h, _ptr(diag), len(blocks), _ptr(bptr), _ptr(fac), threads)
        if not self.h:
            raise OracleError(lib.ko_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ko_metric_destroy(self.h)

    def set_residual_policy(self, policy, interval, gap):
        if policy not in ("fixed", "adaptive"):
            raise OracleError("invalid residual policy")
        _check(load().ko_metric_set_residual_policy(self.h, int(policy == "adaptive"), int(interval), float(gap)))

    def solve(self, B, rtol, atol, maxiter):
        B = _block(B)
        out = np.zeros_like(B)
        r = _report(B.shape[1])
        rc = load().ko_metric_solve(self.h, _ptr(B), _ptr(out), *B.shape, rtol, atol, int(maxiter), *_report_args(r))
        _check(rc, _report_out(r))
        return out, _report_out(r)

    def sample(self, xi, coeff, center, half):
        xi = _block(xi)
        out = np.empty_like(xi)
        c = np.ascontiguousarray(coeff, np.float64)
        _check(load().ko_metric_sample(self.h, _ptr(xi), _ptr(out), *xi.shape, _ptr(c), len(c), center, half))
        return out

    def set_index_blocks(self, blocks):
        """Cluster blocks [(rows, L)] replacing the contiguous ones (cpp/native.cpp:180-197)."""
        rows = np.concatenate([np.asarray(i, np.int64) for i, _ in blocks]) if blocks else np.zeros(1, np.int64)
        off = np.cumsum([0] + [len(i) for i, _ in blocks]).astype(np.int64)
        fac = (np.concatenate([np.ascontiguousarray(L, np.complex128).ravel() for _, L in blocks])
               if blocks else np.zeros(1, np.complex128))
        self._keep_blocks = (rows, off, fac)
        _check(load().ko_metric_set_index_blocks(self.h, len(blocks), _ptr(off), _ptr(rows), _ptr(fac)))

    def set_ic0(self, L, permutation, scaling):
        """IC(0) preconditioner from the factor CSR L (cpp/native.cpp:163-179)."""
        L = sp.csr_matrix(L, dtype=np.complex128)
        arrs = (np.ascontiguousarray(L.indptr, np.int64), np.ascontiguousarray(L.indices, np.int64),
                np.ascontiguousarray(L.data, np.complex128), np.ascontiguousarray(permutation, np.int64),
                np.ascontiguousarray(scaling, np.float64))
        if L.shape[0] != self.S.n or len(arrs[3]) != self.S.n or len(arrs[4]) != self.S.n:
            raise OracleError("invalid IC dimensions")
        self._keep_ic = arrs
        _check(load().ko_metric_set_ic0(self.h, *[_ptr(a) for a in arrs]))

    def set_direct(self, enabled, steps):
        _check(load().ko_metric_set_direct(self.h, int(bool(enabled)), int(steps)))

    def diagnostics(self):
        v = np.zeros(5, np.int64)
        mr = C.c_double()
        load().ko_metric_counters(self.h, _ptr(v), C.byref(mr))
        return dict(calls=int(v[0]), max_iterations_sum=int(v[1]), true_residual_refreshes=int(v[2]),
                    true_residual_checks=int(v[3]), residual_restarts=int(v[4]), max_relative_residual=mr.value,
                    refinement_steps=int(load().ko_metric_refinement_steps(self.h)))


def ic0_factor(lower, shift):
    """IC(0) values on the sorted lower pattern of `lower` (cpp/native.cpp:92-121, 614-618)."""
    L = sp.csr_matrix(lower, dtype=np.complex128)
    ptr = np.ascontiguousarray(L.indptr, np.int64)
    idx = np.ascontiguousarray(L.indices, np.int64)
    val = np.ascontiguousarray(L.data, np.complex128)
    out = np.empty_like(val)
    _check(load().ko_ic0_factor(L.shape[0], _ptr(ptr), _ptr(idx), _ptr(val), float(shift), _ptr(out)))
    return out


class Engine:
    """H / velocity / fused recurrences (cpp/native.cpp:488-610)."""

    def __init__(self, H, metric, definitions, rtol, atol, maxiter, hbar=HBAR_EV_FS):
        lib = load()
        self.H = H if isinstance(H, Csr) else Csr(H)
        self.metric = metric
        kinds = np.array([k for k, _ in definitions], np.int32)
        self._mats = []
        ptrs = (C.c_void_p * 9)()
        for a, (k, mats) in enumerate(definitions):
            for j, m in enumerate(mats):
                m = m if isinstance(m, Csr) else Csr(m)
                self._mats.append(m)
                ptrs[3 * a + j] = m.h
        self._kinds = kinds
        self.h = lib.ko_engine_create(self.H.h, metric.h, _ptr(kinds), ptrs, rtol, atol, int(maxiter), hbar)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ko_engine_destroy(self.h)

    def _fail(self, rc):
        if rc:
            b = 64
            r = _report(b)
            cnt = load().ko_engine_last_report(self.h, *_report_args(r))
            diag = {k: v[:cnt] for k, v in _report_out(r).items()} if cnt else None
            _check(rc, diag)

    def set_controls(self, rtol, atol, maxiter):
        load().ko_engine_set_controls(self.h, rtol, atol, int(maxiter))

    @property
    def h_calls(self):
        h, v = C.c_int64(), C.c_int64()
        load().ko_engine_counters(self.h, C.byref(h), C.byref(v))
        return h.value

    @property
    def v_calls(self):
        h, v = C.c_int64(), C.c_int64()
        load().ko_engine_counters(self.h, C.byref(h), C.byref(v))
        return v.value

    def h_action(self, B):
        B = _block(B)
        out = np.empty_like(B)
        self._fail(load().ko_h_action(self.h, _ptr(B), _ptr(out), *B.shape))
        return out

    def v_cov(self, B, direction):
        B = _block(B)
        out = np.empty_like(B)
        d = np.ascontiguousarray(direction, np.float64)
        self._fail(load().ko_v_cov(self.h, _ptr(d), _ptr(B), _ptr(out), *B.shape))
        return out

    def series(self, B, coeff, center, half, directions=()):
        B = _block(B)
        out = np.empty_like(B)
        c = np.ascontiguousarray(coeff, np.complex128)
        d = np.ascontiguousarray(np.asarray(directions, np.float64).reshape(-1, 3) if len(directions) else
                                 np.zeros((0, 3)))
        self._fail(load().ko_series(self.h, _ptr(B), _ptr(out), *B.shape, _ptr(c), len(c), center, half,
                                    _ptr(d), d.shape[0]))
        return out

    def chain(self, B, length, center, half):
        B = _block(B)
        out = np.empty((length,) + B.shape, np.complex128)
        self._fail(load().ko_chain(self.h, _ptr(B), _ptr(out), *B.shape, int(length), center, half))
        return out

    def moments(self, B, length, center, half):
        B = _block(B)
        out = np.empty((length, B.shape[1]))
        self._fail(load().ko_moments(self.h, _ptr(B), _ptr(out), *B.shape, int(length), center, half))
        return out


def load_reference_native():
    """The reference's own CPU kernel compiled into oracle/_ref (None if absent)."""
    import importlib.util
    import glob
    found = glob.glob(str(HERE / "_ref" / "_native*.so"))
    if not found:
        return None
    spec = importlib.util.spec_from_file_location("_native", found[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def threads_available():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1
