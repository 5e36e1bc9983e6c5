"""MetricSolver: the reference's public metric-solve class
(src/kubo_msd/metric/solve.py:11-174) on the device.

``MetricSolver(S, config, backend=None, atom_slices=None)``
validates S (Hermitian, positive definite), builds the configured
preconditioner (diagonal, block-Jacobi, IC(0)) or the direct
solver, and solves S Y = B for device blocks B with the reference's per-column
acceptance.  ``backend`` is a ``HipContext`` (one is created on
``config.device`` when omitted).  Every solve and preconditioner application
runs in libkmsd.so; the configuration of a prepared solver cannot be changed
per solve.
"""
import numpy as np
import scipy.sparse as sp

from ..errors import NumericalError


class MetricSolver:
    def __init__(self, S, config, backend=None, atom_slices=None):
        from ..backends.hip import HipContext
        from ..operators.velocity import DeviceMetric, hermitian
        self.config = config
        self.backend = backend if backend is not None else HipContext(
            config.device, int(config.device_memory_limit_gb * 1024 ** 3))
        self.atom_slices = atom_slices
        self.host_S = hermitian(sp.csr_matrix(S), "overlap")
        self.metric = DeviceMetric(self.backend, self.host_S, config, atom_slices)
        self.s_lower, self.s_upper, self.bounds_report = (self.metric.s_lower, self.metric.s_upper,
                                                          self.metric.bounds_report)
        self.condition_estimate = self.s_upper / self.s_lower
        self.preconditioner_data = self.metric.preconditioner_data

    def _block(self, B):
        import torch
        if isinstance(B, torch.Tensor):
            B = B.to(self.backend.torch_device)
            return B if B.dtype in (torch.complex64, torch.complex128) else B.to(torch.complex128)
        return torch.from_numpy(np.ascontiguousarray(B, np.complex128)).to(self.backend.torch_device)

    def solve_s(self, B, config=None):
        """(Y, MetricSolveReport) with S Y = B per column; per-solve overrides
        may change tolerances and iteration limits only."""
        if config is not None:
            from types import SimpleNamespace
            config = SimpleNamespace(**{**self.config.as_dict(), **vars(config)})
            for field in ("metric_solver", "preconditioner", "metric_partition_file", "metric_ic_ordering",
                          "metric_ic_shift_max"):
                if getattr(config, field) != getattr(self.config, field):
                    raise NumericalError("per-solve override cannot change prepared " + field)
        B = self._block(B)
        if B.ndim != 2 or B.shape[0] != self.host_S.shape[0] or B.shape[1] < 1:
            raise NumericalError("metric RHS must be finite nonempty (N,batch)")
        return self.metric.solve_s(B, config)

    def precondition(self, R):
        """One application of the prepared preconditioner, M^-1 R."""
        return self.metric.device.precondition(self._block(R))

    def diagnostics(self):
        d = self.metric.diagnostics()
        d.update(condition_upper_or_estimate=self.condition_estimate,
                 kappa_times_epsilon=self.condition_estimate * np.finfo(float).eps)
        return d
