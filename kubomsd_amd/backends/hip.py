"""ctypes binding of libkmsd.so (include/kmsd.h) on PyTorch-ROCm device tensors.

The classes mirror the reference's compiled-kernel module ``kubo_msd._native``
(cpp/native.cpp:613-720): ``HipMatrix`` ~ ``CSR``, ``HipMetric`` ~ ``Metric``,
``HipEngine`` ~ ``Engine``, with the same method names, argument meaning and
error behaviour (``NumericalError`` carrying per-column ``diagnostics``,
message fragments "finite", "representable", "RESOURCE_BLOCKED").  Vector
blocks are torch tensors on the HIP device, complex128 (the reference's type)
or complex64 (FP32 storage, FP64 arithmetic and reductions).

There is no CPU fallback: a missing library or device raises
``BackendUnavailable``.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import scipy.sparse as sp

from ..errors import BackendUnavailable, KuboMSDError, NumericalError, ResourceBlocked
from ..units import HBAR_EV_FS

LIB_PATH = Path(__file__).resolve().parent.parent / "lib" / "libkmsd.so"
KMSD_OK, KMSD_EINVAL, KMSD_ENUMERIC, KMSD_ERESOURCE, KMSD_EDEVICE = range(5)
KMSD_C64, KMSD_C128 = 1, 2
KMSD_TUNE_F32_VALUES = 0x100
KMSD_TUNE_TILE = 0x200
KMSD_CSR, KMSD_BSR4, KMSD_AUTO = 0, 1, 2
KMSD_MU_FP32, KMSD_MU_BF16X3, KMSD_MU_FP64 = 0, 1, 2

_lib = None

# Exported entry points and their ctypes signatures (include/kmsd.h).
_P, _I64, _D, _I, _U64 = C.c_void_p, C.c_int64, C.c_double, C.c_int, C.c_uint64
SIGNATURES = {
    "kmsd_ctx_create": [_I, _U64, _P, _P],
    "kmsd_ctx_destroy": [_P],
    "kmsd_ctx_set_stream": [_P, _P],
    "kmsd_ctx_synchronize": [_P],
    "kmsd_ctx_memory": [_P, _P, _P],
    "kmsd_ctx_transfers": [_P, _P],
    "kmsd_ctx_set_tuning": [_P, C.c_char_p, _I64],
    "kmsd_autotune": [_P, _P, _I64, _I, _P],
    "kmsd_time_step": [_P, _P, _I64, _I, _P],
    "kmsd_last_error": [_P],
    "kmsd_build_info": [_P, _I64],
    "kmsd_matrix_create": [_P, _I, _I64, _I64, _P, _P, _P, _P],
    "kmsd_matrix_destroy": [_P],
    "kmsd_matrix_info": [_P, _P],
    "kmsd_matrix_matmul": [_P, _P, _P, _I64, _I],
    "kmsd_metric_create": [_P, _P, _P, _I64, _P, _P, _P],
    "kmsd_metric_destroy": [_P],
    "kmsd_metric_set_residual_policy": [_P, _I, _I, _D],
    "kmsd_metric_set_mixed": [_P, _I],
    "kmsd_metric_solve": [_P, _P, _P, _I64, _I, _D, _D, _I, _P],
    "kmsd_metric_sample": [_P, _P, _P, _I64, _I, _P, _I64, _D, _D],
    "kmsd_metric_counters": [_P, _P, _P],
    "kmsd_metric_inner": [_P, _P, _P, _P, _I64, _I],
    "kmsd_metric_set_index_blocks": [_P, _I64, _P, _P, _P],
    "kmsd_metric_set_ic0": [_P, _I64, _P, _P, _P, _P, _P],
    "kmsd_metric_set_direct": [_P, _I, _I],
    "kmsd_metric_refinement_steps": [_P, _P],
    "kmsd_metric_precondition": [_P, _P, _P, _I64, _I],
    "kmsd_ic0_factor": [_P, _I64, _I64, _P, _P, _P, _D, _P],
    "kmsd_engine_create": [_P, _P, _P, _P, _P, _D, _D, _I, _D, _P],
    "kmsd_engine_destroy": [_P],
    "kmsd_engine_set_controls": [_P, _D, _D, _I],
    "kmsd_engine_counters": [_P, _P, _P],
    "kmsd_engine_last_report": [_P, _P, _P],
    "kmsd_h_action": [_P, _P, _P, _I64, _I],
    "kmsd_v_cov": [_P, _P, _P, _P, _I64, _I],
    "kmsd_series": [_P, _P, _P, _I64, _I, _P, _I64, _D, _D, _P, _I64],
    "kmsd_chain_v": [_P, _P, _P, _P, _I64, _I64, _I, _D, _D],
    "kmsd_chain": [_P, _P, _P, _I64, _I64, _I, _D, _D],
    "kmsd_moments": [_P, _P, _P, _I64, _I64, _I, _D, _D],
    "kmsd_moments_doubling": [_P, _P, _P, _I64, _I64, _I, _D, _D],
    "kmsd_moments_noalloc": [_P, _P, _P, _I64, _I64, _I, _D, _D],
    "kmsd_moments_group": [_I64, _P, _P, _P, _I64, _I64, _I, _P, _P],
    "kmsd_mu_table": [_P, _P, _P, _P, _I64, _I64, _I64, _I64, _I, _I, _I],
    "kmsd_block_accumulate": [_P, _P, _P, _I64, _I64, _I64, _P, C.c_int32],
    "kmsd_time_gradient": [_P, _P, _P, _I64, _I64, _I64, _P, _D],
    "kmsd_bootstrap_ratio": [_P, _P, _P, _I64, _I64, _I64, _I64, _P, _I64, _P, _P, _P, _P],
    "kmsd_dos_validity": [_P, _P, _P, _I64, _D, _I64, _P, _P],
}


class _Report(C.Structure):
    _fields_ = [("converged", _P), ("breakdown", _P), ("iterations", _P),
                ("true_residual", _P), ("relative_residual", _P)]


def library():
    """Loads libkmsd.so; raises BackendUnavailable (never falls back)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise BackendUnavailable(
                f"HIP backend requires the native library {LIB_PATH}: run "
                "python -c 'import __graft_entry__; __graft_entry__.build()' (hipcc, gfx950)")
        lib = C.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = C.c_int
        lib.kmsd_last_error.restype = C.c_char_p
        _lib = lib
    return _lib


_TUNED = {}


def build_info():
    buf = C.create_string_buffer(512)
    library().kmsd_build_info(buf, 512)
    import json
    return json.loads(buf.value.decode())


def _torch():
    import torch
    return torch


def _dtype_code(t):
    torch = _torch()
    if t.dtype == torch.complex128:
        return KMSD_C128
    if t.dtype == torch.complex64:
        return KMSD_C64
    raise KuboMSDError("vector blocks must be complex128 or complex64 tensors")


def _np_ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class _ReportBuffers:
    def __init__(self, b):
        self.conv = np.zeros(b, np.uint8)
        self.brk = np.zeros(b, np.uint8)
        self.its = np.zeros(b, np.int64)
        self.res = np.zeros(b)
        self.rel = np.zeros(b)
        self.struct = _Report(*[_np_ptr(a) for a in (self.conv, self.brk, self.its, self.res, self.rel)])

    def as_dict(self, b=None):
        b = len(self.conv) if b is None else b
        return dict(converged=[bool(x) for x in self.conv[:b]], breakdown=[bool(x) for x in self.brk[:b]],
                    iterations=self.its[:b].tolist(), true_residual=self.res[:b].tolist(),
                    relative_residual=self.rel[:b].tolist())


class HipContext:
    """A HIP device + stream + HBM budget (include/kmsd.h kmsd_ctx_*)."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise BackendUnavailable("HIP backend requires a ROCm GPU (torch.cuda.is_available() is False)")
        self.lib = library()
        self.device = int(device)
        self.torch_device = torch.device("cuda", self.device)
        h = C.c_void_p()
        stream = torch.cuda.current_stream(self.torch_device).cuda_stream
        rc = self.lib.kmsd_ctx_create(self.device, int(hbm_budget_bytes), C.c_void_p(stream), C.byref(h))
        if rc:
            err = self.lib.kmsd_last_error(None).decode()
            if rc == KMSD_EINVAL:
                raise KuboMSDError(err)
            self._raise(rc, err)
        self.h = h
        # host-side transfer ledger (bytes moved by this package's own
        # torch copies; the library keeps its own, kmsd_ctx_transfers)
        self.host_transfers = {"h2d": 0, "d2h_data": 0, "d2h_status": 0}

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.kmsd_ctx_destroy(self.h)
            self.h = None

    def bind_stream(self):
        torch = _torch()
        self.lib.kmsd_ctx_set_stream(self.h, C.c_void_p(torch.cuda.current_stream(self.torch_device).cuda_stream))

    def error(self):
        return self.lib.kmsd_last_error(self.h).decode()

    @staticmethod
    def _raise(rc, msg, diagnostics=None):
        if rc == KMSD_ERESOURCE:
            err = ResourceBlocked(msg)
        elif rc == KMSD_EDEVICE:
            err = NumericalError("HIP device error: " + msg)
        else:
            err = NumericalError(msg)
        if diagnostics is not None and rc == KMSD_ENUMERIC:  # per-column reports of failed solves only
            err.diagnostics = diagnostics
        raise err

    def check(self, rc, diagnostics=None):
        if rc:
            self._raise(rc, self.error(), diagnostics)

    def set_tuning(self, key, value):
        self.check(self.lib.kmsd_ctx_set_tuning(self.h, key.encode(), int(value)))

    def autotune(self, matrix, nb, dtype, cache_path=None, f32_values=False, tile=False):
        """Launch shape (columns per lane, workgroups) for products of `matrix`
        on nb-column blocks of `dtype` (torch dtype), timed on the device by
        kmsd_autotune and remembered per (device, storage, n, nnz, nb, dtype)
        in this process and in a JSON file (KUBOMSD_TUNING_CACHE, default
        ~/.cache/kubomsd_amd/tuning.json), so later runs reuse the same shape
        and produce the same results.  f32_values: tune the product with FP32
        matrix values (the metric products of complex64 solves, which dominate
        nonorthogonal chains).  tile: also the tile size of the plane-tiled
        4x4-block walk (matrices with a plane stride).  Applies the shape to
        this context."""
        import json
        import os
        torch = _torch()
        code = KMSD_C64 if dtype == torch.complex64 else KMSD_C128
        if f32_values and code == KMSD_C64:
            code |= KMSD_TUNE_F32_VALUES
        if tile:
            code |= KMSD_TUNE_TILE
        info = matrix.info()
        key = "|".join(str(v) for v in (torch.cuda.get_device_name(self.device), info["format"], info["n"],
                                         info["nnz"], int(nb), code))
        path = cache_path or os.environ.get("KUBOMSD_TUNING_CACHE") or \
            os.path.join(os.path.expanduser("~"), ".cache", "kubomsd_amd", "tuning.json")
        choice = _TUNED.get(key)
        if choice is None:
            try:
                with open(path) as f:
                    choice = json.load(f).get(key)
            except (OSError, ValueError):
                choice = None
        if choice is None:
            self.bind_stream()
            out = np.zeros(3, np.int64)
            self.check(self.lib.kmsd_autotune(self.h, matrix.h, int(nb), code, _np_ptr(out)))
            choice = {"cpl": int(out[0]), "groups": int(out[1])}
            if tile:
                choice["tile_bytes"] = int(out[2])
            try:
                os.makedirs(os.path.dirname(path), exist_ok=True)
                try:
                    with open(path) as f:
                        table = json.load(f)
                except (OSError, ValueError):
                    table = {}
                table[key] = choice
                tmp = path + f".{os.getpid()}.tmp"
                with open(tmp, "w") as f:
                    json.dump(table, f, indent=1, sort_keys=True)
                os.replace(tmp, path)
            except OSError:
                pass  # an unwritable cache only costs a re-tune next run
        _TUNED[key] = choice
        self.set_tuning("cpl_c64" if dtype == torch.complex64 else "cpl_c128", choice["cpl"])
        self.set_tuning("groups", choice["groups"])
        if "tile_bytes" in choice:
            self.set_tuning("tile_bytes", choice["tile_bytes"])
        return dict(choice)

    def time_step(self, matrix, nb, dtype):
        """Microseconds of one fused Chebyshev step with `matrix` on an
        nb-column block of `dtype` at the current launch shape
        (kmsd_time_step)."""
        torch = _torch()
        code = KMSD_C64 if dtype == torch.complex64 else KMSD_C128
        self.bind_stream()
        us = C.c_double()
        self.check(self.lib.kmsd_time_step(self.h, matrix.h, int(nb), code, C.byref(us)))
        return us.value

    def tune_block_width(self, matrix, widths, dtype, margin=0.03):
        """Vector-block width among `widths` (ascending) with the least
        measured step time per column: a wider block is taken only when it is
        more than `margin` faster per column than the best narrower one, so
        ties keep the smaller workspace.  Remembered per (device, storage, n,
        nnz, widths, dtype) in this process.  Returns (width, {width: us per
        step})."""
        torch = _torch()
        widths = sorted({int(w) for w in widths})
        info = matrix.info()
        key = "|".join(str(v) for v in ("width", torch.cuda.get_device_name(self.device), info["format"],
                                         info["n"], info["nnz"], ",".join(map(str, widths)), str(dtype)))
        hit = _TUNED.get(key)
        if hit is not None:
            return hit["width"], dict(hit["us"])
        us = {w: self.time_step(matrix, w, dtype) for w in widths}
        best = widths[0]
        for w in widths[1:]:
            if us[w] / w < (1.0 - margin) * us[best] / best:
                best = w
        _TUNED[key] = {"width": best, "us": us}
        return best, dict(us)

    def upload(self, array, dtype=None):
        """Host array -> device tensor on this context (counted as H2D)."""
        torch = _torch()
        t = torch.from_numpy(np.ascontiguousarray(array)).to(self.torch_device)
        self.host_transfers["h2d"] += t.numel() * t.element_size()
        return t if dtype is None else t.to(dtype)

    def to_host(self, t, status=False):
        """Device tensor -> NumPy array, counted as array data or (status=True)
        as control information (flags, per-column norms, scalars)."""
        if not hasattr(t, "detach"):
            return np.asarray(t)
        out = t.detach().cpu().numpy()
        self.host_transfers["d2h_status" if status else "d2h_data"] += out.nbytes
        return out

    def status_value(self, t):
        """One device scalar read for a control decision (counted as status)."""
        return self.to_host(t.reshape(()), status=True).item()

    def transfers(self):
        """Bytes moved between host and device on this context: by the library
        (kmsd_ctx_transfers) and by this package's torch copies."""
        v = np.zeros(3, np.uint64)
        self.check(self.lib.kmsd_ctx_transfers(self.h, _np_ptr(v)))
        return dict(library_h2d=int(v[0]), library_d2h_data=int(v[1]), library_d2h_status=int(v[2]),
                    host_h2d=self.host_transfers["h2d"], host_d2h_data=self.host_transfers["d2h_data"],
                    host_d2h_status=self.host_transfers["d2h_status"])

    def memory(self):
        used, budget = C.c_uint64(), C.c_uint64()
        self.lib.kmsd_ctx_memory(self.h, C.byref(used), C.byref(budget))
        return dict(used_bytes=used.value, budget_bytes=budget.value)

    def synchronize(self):
        self.check(self.lib.kmsd_ctx_synchronize(self.h))

    def block(self, B, name="block"):
        """Validates a caller block: 2-D, contiguous, complex, on this device."""
        torch = _torch()
        if not isinstance(B, torch.Tensor):
            raise KuboMSDError(f"{name} must be a device tensor")
        if B.ndim != 2 or B.shape[0] < 1 or B.shape[1] < 1:
            raise KuboMSDError("expected nonempty finite (N,batch) array")
        if B.device != self.torch_device:
            raise KuboMSDError(f"{name} must live on {self.torch_device}")
        _dtype_code(B)
        return B.contiguous()


def _csr_host(matrix):
    m = sp.csr_matrix(matrix, dtype=np.complex128, copy=True)
    m.sum_duplicates()
    m.sort_indices()
    return m


class HipMatrix:
    """Device sparse matrix (reference CSR, cpp/native.cpp:54-90).  fmt picks
    the device storage: KMSD_AUTO (default: 4x4 blocks when they are at least
    75% full, else CSR), KMSD_CSR or KMSD_BSR4."""

    def __init__(self, ctx, matrix, fmt=KMSD_AUTO):
        self.ctx = ctx
        m = _csr_host(matrix)
        if m.shape[0] != m.shape[1]:
            raise KuboMSDError("invalid CSR dimensions/threads")
        self.n = m.shape[0]
        self.nnz = m.nnz
        ptr = np.ascontiguousarray(m.indptr, np.int64)
        idx = np.ascontiguousarray(m.indices, np.int64)
        val = np.ascontiguousarray(m.data, np.complex128)
        h = C.c_void_p()
        ctx.check(ctx.lib.kmsd_matrix_create(ctx.h, int(fmt), self.n, self.nnz, _np_ptr(ptr), _np_ptr(idx),
                                             _np_ptr(val), C.byref(h)))
        self.h = h
        self.host = m

    def info(self):
        v = np.zeros(8, np.int64)
        self.ctx.check(self.ctx.lib.kmsd_matrix_info(self.h, _np_ptr(v)))
        return dict(format={KMSD_CSR: "csr", KMSD_BSR4: "bsr4"}[int(v[0])], n=int(v[1]), nnz=int(v[2]),
                    stored=int(v[3]), device_bytes=int(v[4]), values=("real", "imag", "complex")[int(v[5])],
                    uniform_values=bool(v[6]), plane_stride=int(v[7]))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.kmsd_matrix_destroy(self.h)
            self.h = None

    def matmul(self, B):
        ctx = self.ctx
        B = ctx.block(B)
        if B.shape[0] != self.n:
            raise KuboMSDError("CSR input dimension mismatch")
        ctx.bind_stream()
        out = _torch().empty_like(B)
        ctx.check(ctx.lib.kmsd_matrix_matmul(self.h, C.c_void_p(B.data_ptr()), C.c_void_p(out.data_ptr()),
                                             B.shape[1], _dtype_code(B)))
        return out

    def __matmul__(self, B):
        return self.matmul(B)


class HipMetric:
    """Overlap metric with batched device PCG (reference Metric, cpp/native.cpp:150-480)."""

    def __init__(self, ctx, S, diagonal, blocks=()):
        self.ctx = ctx
        self.S = S if isinstance(S, HipMatrix) else HipMatrix(ctx, S)
        self.n = self.S.n
        diag = np.ascontiguousarray(diagonal, np.float64)
        if diag.shape != (self.n,):
            raise KuboMSDError("invalid metric diagonal/threads")
        blocks = list(blocks)
        bptr = np.array([0] + [int(b) for (_, b, _) in blocks], np.int64)
        fac = (np.concatenate([np.ascontiguousarray(L, np.complex128).ravel() for (_, _, L) in blocks])
               if blocks else np.zeros(1, np.complex128))
        for q, (a, b, L) in enumerate(blocks):
            if int(a) != bptr[q]:
                raise KuboMSDError("invalid block-Jacobi factor partition")
        h = C.c_void_p()
        ctx.check(ctx.lib.kmsd_metric_create(ctx.h, self.S.h, _np_ptr(diag), len(blocks), _np_ptr(bptr),
                                             _np_ptr(fac), C.byref(h)))
        self.h = h
        self.diagonal = diag

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.kmsd_metric_destroy(self.h)
            self.h = None

    def set_residual_policy(self, policy, interval, gap):
        if policy not in ("fixed", "adaptive"):
            raise KuboMSDError("invalid residual policy")
        self.ctx.check(self.ctx.lib.kmsd_metric_set_residual_policy(self.h, int(policy == "adaptive"),
                                                                   int(interval), float(gap)))

    def set_mixed(self, enabled):
        """complex128 solves by FP32 inner solves inside the FP64 refinement
        (the complex64 path), accepted on the same FP64 true residual."""
        self.ctx.check(self.ctx.lib.kmsd_metric_set_mixed(self.h, int(bool(enabled))))

    def solve(self, B, rtol, atol, maxiter):
        ctx = self.ctx
        B = ctx.block(B)
        ctx.bind_stream()
        out = _torch().empty_like(B)
        rep = _ReportBuffers(B.shape[1])
        rc = ctx.lib.kmsd_metric_solve(self.h, C.c_void_p(B.data_ptr()), C.c_void_p(out.data_ptr()), B.shape[1],
                                       _dtype_code(B), float(rtol), float(atol), int(maxiter), C.byref(rep.struct))
        ctx.check(rc, rep.as_dict())
        return out, rep.as_dict()

    def sample(self, xi, coeff, center, half):
        ctx = self.ctx
        xi = ctx.block(xi)
        ctx.bind_stream()
        out = _torch().empty_like(xi)
        c = np.ascontiguousarray(coeff, np.float64)
        ctx.check(ctx.lib.kmsd_metric_sample(self.h, C.c_void_p(xi.data_ptr()), C.c_void_p(out.data_ptr()),
                                             xi.shape[1], _dtype_code(xi), _np_ptr(c), len(c), float(center),
                                             float(half)))
        return out

    def inner(self, A, B):
        """Device complex128 (nb,): sum_i conj(A_ij) (S B)_ij."""
        ctx = self.ctx
        A = ctx.block(A)
        B = ctx.block(B)
        if A.shape != B.shape or A.dtype != B.dtype or A.shape[0] != self.n:
            raise KuboMSDError("inner product blocks must match")
        ctx.bind_stream()
        torch = _torch()
        out = torch.empty(A.shape[1], dtype=torch.complex128, device=A.device)
        ctx.check(ctx.lib.kmsd_metric_inner(self.h, C.c_void_p(A.data_ptr()), C.c_void_p(B.data_ptr()),
                                            C.c_void_p(out.data_ptr()), A.shape[1], _dtype_code(A)))
        return out

    def set_index_blocks(self, blocks):
        """Cluster-Jacobi blocks [(rows, L)] replacing the contiguous ones
        (reference Metric.set_index_blocks, cpp/native.cpp:180-197)."""
        blocks = list(blocks)
        rows = (np.concatenate([np.asarray(i, np.int64).ravel() for i, _ in blocks]) if blocks
                else np.zeros(1, np.int64))
        off = np.cumsum([0] + [np.asarray(i).size for i, _ in blocks]).astype(np.int64)
        for i, L in blocks:
            m = np.asarray(i).size
            if np.asarray(L).shape != (m, m):
                raise KuboMSDError("invalid cluster factor")
        fac = (np.concatenate([np.ascontiguousarray(L, np.complex128).ravel() for _, L in blocks]) if blocks
               else np.zeros(1, np.complex128))
        self.ctx.check(self.ctx.lib.kmsd_metric_set_index_blocks(self.h, len(blocks), _np_ptr(off), _np_ptr(rows),
                                                                _np_ptr(fac)))

    def set_ic0(self, L, permutation, scaling):
        """IC(0) preconditioner from the factor L (CSR, diagonal last in each
        row), the factor-row-to-original-row permutation and D per original
        row (reference Metric.set_ic0, cpp/native.cpp:163-179)."""
        L = sp.csr_matrix(L)
        p = np.ascontiguousarray(permutation, np.int64)
        z = np.ascontiguousarray(scaling, np.float64)
        if L.shape != (self.n, self.n) or p.shape != (self.n,) or z.shape != (self.n,):
            raise KuboMSDError("invalid IC dimensions")
        ptr = np.ascontiguousarray(L.indptr, np.int64)
        idx = np.ascontiguousarray(L.indices, np.int64)
        val = np.ascontiguousarray(L.data, np.complex128)
        self.ctx.check(self.ctx.lib.kmsd_metric_set_ic0(self.h, len(val), _np_ptr(ptr), _np_ptr(idx), _np_ptr(val),
                                                       _np_ptr(p), _np_ptr(z)))

    def precondition(self, R):
        """Z = M^-1 R, one application of the metric's preconditioner."""
        ctx = self.ctx
        R = ctx.block(R)
        ctx.bind_stream()
        out = _torch().empty_like(R)
        ctx.check(ctx.lib.kmsd_metric_precondition(self.h, C.c_void_p(R.data_ptr()), C.c_void_p(out.data_ptr()),
                                                  R.shape[1], _dtype_code(R)))
        return out

    def set_direct(self, enabled, steps):
        """Direct solves with the one full-size block factor plus `steps`
        refinement corrections (reference Metric.set_direct, cpp/native.cpp:198-205)."""
        self.ctx.check(self.ctx.lib.kmsd_metric_set_direct(self.h, int(bool(enabled)), int(steps)))

    def diagnostics(self):
        v = np.zeros(5, np.int64)
        mr = C.c_double()
        self.ctx.lib.kmsd_metric_counters(self.h, _np_ptr(v), C.byref(mr))
        steps = np.zeros(1, np.int64)
        self.ctx.lib.kmsd_metric_refinement_steps(self.h, _np_ptr(steps))
        return dict(calls=int(v[0]), max_iterations_sum=int(v[1]), true_residual_refreshes=int(v[2]),
                    true_residual_checks=int(v[3]), residual_restarts=int(v[4]),
                    max_relative_residual=mr.value, refinement_steps=int(steps[0]))


def ic0_factor(ctx, lower, shift):
    """IC(0) values on the sorted lower pattern of `lower`, computed on the
    device (reference ic0_factor, cpp/native.cpp:92-121, 614-618).  A failing
    pivot raises NumericalError naming the row."""
    L = sp.csr_matrix(lower)
    ptr = np.ascontiguousarray(L.indptr, np.int64)
    idx = np.ascontiguousarray(L.indices, np.int64)
    val = np.ascontiguousarray(L.data, np.complex128)
    out = np.empty_like(val)
    ctx.check(ctx.lib.kmsd_ic0_factor(ctx.h, L.shape[0], len(val), _np_ptr(ptr), _np_ptr(idx), _np_ptr(val),
                                      float(shift), _np_ptr(out)))
    return out


class HipEngine:
    """H, metric and velocity actions with fused Chebyshev drivers
    (reference Engine, cpp/native.cpp:488-610)."""

    def __init__(self, ctx, H, metric, definitions, rtol, atol, maxiter, hbar=HBAR_EV_FS):
        self.ctx = ctx
        self.H = H if isinstance(H, HipMatrix) else HipMatrix(ctx, H)
        self.metric = metric
        self.n = self.H.n
        kinds = np.array([int(k) for k, _ in definitions], np.int32)
        if kinds.shape != (3,):
            raise KuboMSDError("invalid engine operators")
        ptrs = (C.c_void_p * 9)()
        self._mats = []
        for a, (k, mats) in enumerate(definitions):
            for j, m in enumerate(mats):
                m = m if isinstance(m, HipMatrix) else HipMatrix(ctx, m)
                self._mats.append(m)
                ptrs[3 * a + j] = m.h
        h = C.c_void_p()
        ctx.check(ctx.lib.kmsd_engine_create(ctx.h, self.H.h, metric.h, _np_ptr(kinds), ptrs, float(rtol),
                                             float(atol), int(maxiter), float(hbar), C.byref(h)))
        self.h = h
        self.controls = (float(rtol), float(atol), int(maxiter))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.kmsd_engine_destroy(self.h)
            self.h = None

    def _counters(self):
        h, v = C.c_int64(), C.c_int64()
        self.ctx.lib.kmsd_engine_counters(self.h, C.byref(h), C.byref(v))
        return h.value, v.value

    @property
    def h_calls(self):
        return self._counters()[0]

    @property
    def v_calls(self):
        return self._counters()[1]

    def _check(self, rc):
        if rc:
            nb = C.c_int64()
            self.ctx.lib.kmsd_engine_last_report(self.h, None, C.byref(nb))
            diag = None
            if rc == KMSD_ENUMERIC and nb.value:
                rep = _ReportBuffers(nb.value)
                self.ctx.lib.kmsd_engine_last_report(self.h, C.byref(rep.struct), C.byref(nb))
                if not all(rep.conv):
                    diag = rep.as_dict()
            self.ctx.check(rc, diag)

    def _prep(self, B):
        B = self.ctx.block(B)
        if B.shape[0] != self.n:
            raise KuboMSDError("CSR input dimension mismatch")
        self.ctx.bind_stream()
        return B

    def h_action(self, B):
        B = self._prep(B)
        out = _torch().empty_like(B)
        self._check(self.ctx.lib.kmsd_h_action(self.h, C.c_void_p(B.data_ptr()), C.c_void_p(out.data_ptr()),
                                               B.shape[1], _dtype_code(B)))
        return out

    def v_cov(self, B, direction):
        B = self._prep(B)
        d = np.ascontiguousarray(direction, np.float64)
        if d.shape != (3,):
            raise KuboMSDError("direction must have three components")
        out = _torch().empty_like(B)
        self._check(self.ctx.lib.kmsd_v_cov(self.h, _np_ptr(d), C.c_void_p(B.data_ptr()),
                                            C.c_void_p(out.data_ptr()), B.shape[1], _dtype_code(B)))
        return out

    def v_cov_with_controls(self, B, direction, rtol, atol, maxiter):
        """Velocity action with temporary solve controls; the engine keeps its own."""
        saved = self.controls
        self.ctx.lib.kmsd_engine_set_controls(self.h, float(rtol), float(atol), int(maxiter))
        try:
            return self.v_cov(B, direction)
        finally:
            self.ctx.lib.kmsd_engine_set_controls(self.h, *saved)

    def series(self, B, coeff, center, half, directions=()):
        B = self._prep(B)
        c = np.ascontiguousarray(coeff, np.complex128)
        if c.ndim != 1:
            raise KuboMSDError("Chebyshev coefficients must be 1D")
        d = np.ascontiguousarray(np.asarray(directions, np.float64).reshape(-1, 3) if len(directions)
                                 else np.zeros((0, 3)))
        out = _torch().empty_like(B)
        self._check(self.ctx.lib.kmsd_series(self.h, C.c_void_p(B.data_ptr()), C.c_void_p(out.data_ptr()),
                                             B.shape[1], _dtype_code(B), _np_ptr(c), len(c), float(center),
                                             float(half), _np_ptr(d), d.shape[0]))
        return out

    def chain(self, B, length, center, half, out=None):
        B = self._prep(B)
        torch = _torch()
        if out is None:
            out = torch.empty((int(length),) + tuple(B.shape), dtype=B.dtype, device=B.device)
        self._check(self.ctx.lib.kmsd_chain(self.h, C.c_void_p(B.data_ptr()), C.c_void_p(out.data_ptr()),
                                            int(length), B.shape[1], _dtype_code(B), float(center), float(half)))
        return out

    def chain_v(self, B, length, center, half, direction, out=None):
        """(length, N, nb): v_cov(direction, T_k(H~) B) for k < length."""
        B = self._prep(B)
        torch = _torch()
        d = np.ascontiguousarray(direction, np.float64)
        if d.shape != (3,):
            raise KuboMSDError("direction must have three components")
        if out is None:
            out = torch.empty((int(length),) + tuple(B.shape), dtype=B.dtype, device=B.device)
        self._check(self.ctx.lib.kmsd_chain_v(self.h, _np_ptr(d), C.c_void_p(B.data_ptr()), C.c_void_p(out.data_ptr()),
                                              int(length), B.shape[1], _dtype_code(B), float(center), float(half)))
        return out

    def moments(self, B, length, center, half, doubling=False):
        """Device float64 (length, nb): mu[l, j] = Re <S B_j, T_l B_j>.
        doubling=True (S = I only) uses the Chebyshev product identities: about
        half the orders and no projection stream."""
        B = self._prep(B)
        torch = _torch()
        mu = torch.empty((int(length), B.shape[1]), dtype=torch.float64, device=B.device)
        fn = self.ctx.lib.kmsd_moments_doubling if doubling else self.ctx.lib.kmsd_moments
        self._check(fn(self.h, C.c_void_p(B.data_ptr()), C.c_void_p(mu.data_ptr()), int(length), B.shape[1],
                       _dtype_code(B), float(center), float(half)))
        return mu


def moments_group(engines, blocks, length, centers, halves):
    """Device float64 (length, nb) moments of every engine's chain, as
    HipEngine.moments would give them one by one, with the chains' orders
    in shared launches (kmsd_moments_group).  engines share one context;
    blocks are (n_q, nb) device tensors of one dtype."""
    torch = _torch()
    engines = list(engines)
    if not engines or len(blocks) != len(engines) or len(centers) != len(engines) or len(halves) != len(engines):
        raise KuboMSDError("one block, center and half per engine")
    ctx = engines[0].ctx
    prepped = [e._prep(B) for e, B in zip(engines, blocks)]
    nb, dt = prepped[0].shape[1], prepped[0].dtype
    if any(B.shape[1] != nb or B.dtype != dt for B in prepped):
        raise KuboMSDError("grouped chains need blocks of one width and dtype")
    mus = [torch.empty((int(length), nb), dtype=torch.float64, device=B.device) for B in prepped]
    k = len(engines)
    eh = (C.c_void_p * k)(*[e.h for e in engines])
    bp = (C.c_void_p * k)(*[B.data_ptr() for B in prepped])
    mp = (C.c_void_p * k)(*[m.data_ptr() for m in mus])
    cs = np.ascontiguousarray(centers, np.float64)
    hs = np.ascontiguousarray(halves, np.float64)
    ctx.bind_stream()
    rc = ctx.lib.kmsd_moments_group(k, eh, bp, mp, int(length), nb, _dtype_code(prepped[0]), _np_ptr(cs), _np_ptr(hs))
    if rc:
        engines[0]._check(rc)
    return mus


def mu_table(ctx, left, right, mode=KMSD_MU_FP64, table=None, accumulate=False):
    """table[b, j, q] = sum_i conj(left[j, i, b]) right[q, i, b] (complex128, device)."""
    torch = _torch()
    if left.ndim != 3 or right.ndim != 3 or left.shape[1:] != right.shape[1:] or left.dtype != right.dtype:
        raise KuboMSDError("left/right must be (m,N,nb) and (k,N,nb) blocks of one dtype")
    left = left.contiguous()
    right = right.contiguous()
    m, n, nb = left.shape
    k = right.shape[0]
    if table is None:
        table = torch.zeros((nb, m, k), dtype=torch.complex128, device=left.device)
    ctx.bind_stream()
    ctx.check(ctx.lib.kmsd_mu_table(ctx.h, C.c_void_p(left.data_ptr()), C.c_void_p(right.data_ptr()),
                                    C.c_void_p(table.data_ptr()), n, m, k, nb, _dtype_code(left), int(mode),
                                    int(bool(accumulate))))
    return table
