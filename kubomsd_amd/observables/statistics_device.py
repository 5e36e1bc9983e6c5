"""Device-resident statistics of the stochastic estimator (reference
src/kubo_msd/observables/statistics.py:5-23, observables/diffusion.py:4-6,
solvers/stochastic.py:52-119; the GPU edition keeps the block sums and the
bootstrap on the device: kubo_msd_gpu/src/kubo_msd/observables/
statistics_gpu.py:6-32, solvers/stochastic_gpu.py:53-145).

Block sums, the MSD time gradient, the paired block bootstrap and the DOS
validity masks run in libkmsd.so (kmsd_block_accumulate, kmsd_time_gradient,
kmsd_bootstrap_ratio, kmsd_dos_validity) on float64 device tensors; only the
resampling weights (integer counts drawn with numpy.random.default_rng(21491),
the reference's generator) come from the host.
"""
import ctypes as C

import numpy as np

from ..errors import KuboMSDError, NumericalError

FLAG_DOS, FLAG_NUMERATOR = 1, 2


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def resample_weights(nblocks, n_resamples=256):
    """Counts of each statistics block in each paired resample (the
    reference's draws: default_rng(21491).integers, bincount per row)."""
    rng = np.random.default_rng(21491)
    idx = rng.integers(nblocks, size=(n_resamples, nblocks))
    return np.array([np.bincount(row, minlength=nblocks) for row in idx], float)


class DeviceStatistics:
    """Statistics kernels bound to one HipContext (its device and stream)."""

    def __init__(self, ctx):
        import torch
        self.torch = torch
        self.ctx = ctx
        self.flags = torch.zeros(1, dtype=torch.int32, device=ctx.torch_device)
        self._weights = {}

    def _f64(self, t):
        torch = self.torch
        if not isinstance(t, torch.Tensor) or t.device != self.ctx.torch_device or t.dtype != torch.float64:
            raise KuboMSDError("device statistics need float64 tensors on the context's device")
        return t.contiguous()

    def clear_flags(self):
        self.flags.zero_()

    def check_flags(self):
        """Raises the reference's errors for nonfinite samples seen since the
        last clear (DOS first, as the reference checks it first)."""
        f = int(self.ctx.status_value(self.flags[0]))
        if f & FLAG_DOS:
            raise NumericalError("nonfinite stochastic DOS")
        if f & FLAG_NUMERATOR:
            raise NumericalError("nonfinite stochastic numerator")

    def accumulate(self, dst, src, flag):
        """dst += src[l] for every leading index l, in order (dst: one block
        row, src: per-sample rows of the same shape)."""
        src = self._f64(src)
        if not dst.is_contiguous():
            raise KuboMSDError("accumulation target must be contiguous")
        nelem = dst.numel()
        nloc = src.shape[0]
        if src.numel() != nloc * nelem:
            raise KuboMSDError("per-sample shape does not match the block shape")
        self.ctx.bind_stream()
        self.ctx.check(self.ctx.lib.kmsd_block_accumulate(self.ctx.h, _ptr(dst), _ptr(src), nloc, nelem, nelem,
                                                          _ptr(self.flags), flag))

    def msd_gradient(self, time_fs, msd, axis):
        """0.5 * numpy.gradient(msd, time_fs, axis, edge_order 2 (1 for two
        times)), bit for bit; None for fewer than two times."""
        if len(time_fs) < 2:
            return None
        msd = self._f64(msd)
        shape = msd.shape
        nt = shape[axis]
        outer = int(np.prod(shape[:axis], dtype=np.int64))
        inner = int(np.prod(shape[axis + 1:], dtype=np.int64))
        out = self.torch.empty_like(msd)
        t = np.ascontiguousarray(np.asarray(time_fs, float))
        self.ctx.bind_stream()
        self.ctx.check(self.ctx.lib.kmsd_time_gradient(self.ctx.h, _ptr(msd), _ptr(out), outer, nt, inner,
                                                       t.ctypes.data_as(C.c_void_p), .5))
        return out

    def weights(self, nblocks, n_resamples=256):
        key = (nblocks, n_resamples)
        if key not in self._weights:
            w = resample_weights(nblocks, n_resamples)
            self._weights[key] = self.torch.from_numpy(w).to(self.ctx.torch_device)
        return self._weights[key]

    def estimate_ratio(self, A, D, counts, n_resamples=256, valid=None, want_total=False):
        """Device estimate_ratio: (value, stderr or None, status[, total]).
        A: (blocks, ..., energy); D: (blocks, energy) or (blocks, 1).  valid
        (energy, bool) sets value and stderr to NaN where False."""
        torch = self.torch
        A = self._f64(A)
        D = self._f64(D)
        counts = np.asarray(counts, int)
        nb = len(counts)
        if A.shape[0] != nb or D.shape[0] != nb or np.any(counts <= 0):
            raise KuboMSDError("inconsistent statistical blocks/counts")
        ne = A.shape[-1]
        if D.ndim != 2 or (D.shape[1] != ne and D.shape != (nb, 1)):
            raise KuboMSDError("statistical denominator must be (blocks,energy)")
        status = {"method": "paired block bootstrap of sums", "n_blocks": nb, "counts": counts.tolist(),
                  "reliable_interval": nb >= 8, "bootstrap_seed": 21491, "n_resamples": n_resamples,
                  "input_accumulation": "sum; actual unequal counts retained"}
        value = torch.empty(A.shape[1:], dtype=torch.float64, device=A.device)
        se = torch.empty_like(value) if nb >= 2 else None
        total = torch.empty_like(value) if want_total else None
        w = self.weights(nb, n_resamples) if se is not None else None
        vmask = None
        if valid is not None:
            vmask = valid.to(device=A.device, dtype=torch.uint8).contiguous()
            if vmask.numel() != ne:
                raise KuboMSDError("validity mask must have one entry per energy")
        self.ctx.bind_stream()
        self.ctx.check(self.ctx.lib.kmsd_bootstrap_ratio(self.ctx.h, _ptr(A), _ptr(D), nb, value.numel(), ne,
                                                         D.shape[1], _ptr(w), n_resamples, _ptr(vmask),
                                                         _ptr(value), _ptr(se), _ptr(total)))
        return (value, se, status, total) if want_total else (value, se, status)

    def dos_validity(self, rho, se, threshold, nblocks):
        """(valid bool, reason uint32) device tensors of the DOS estimate."""
        torch = self.torch
        rho = self._f64(rho)
        se = self._f64(se) if se is not None else None
        valid = torch.empty(rho.shape, dtype=torch.uint8, device=rho.device)
        reason = torch.empty(rho.shape, dtype=torch.int32, device=rho.device)
        self.ctx.bind_stream()
        self.ctx.check(self.ctx.lib.kmsd_dos_validity(self.ctx.h, _ptr(rho), _ptr(se), rho.numel(), float(threshold),
                                                      int(nblocks), _ptr(valid), _ptr(reason)))
        return valid.bool(), reason
