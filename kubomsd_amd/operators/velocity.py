"""Device operator bundle: H, S, velocities and their actions on the MI355X
(reference src/kubo_msd/operators/velocity.py:10-124, the seam the stochastic
solvers call through).

``OperatorBundle(model, config)`` uploads H, S (CSR, values FP64 and FP32) and
the Cartesian velocity definitions (explicit V, position X, or the composite
bond/xbar/delta form) once; every action below runs on device tensors:
  apply_h_action(B) = S^-1 H B        apply_v_cov(axis, B)  covariant velocity
  apply_v_action(axis, B) = S^-1 v B  inner(A, B) = sum_i conj(A) S B per column
  solve_s(B)                          ops.native: the fused HipEngine
Vectors are complex128 (precision='fp64') or complex64 (precision='mixed').
"""
import time
from types import SimpleNamespace

import numpy as np
import scipy.linalg as la
import scipy.sparse as sp

from ..backends.hip import HipContext, HipEngine, HipMatrix, HipMetric
from ..contracts import MetricSolveReport
from ..errors import KuboMSDError, NumericalError, ResourceBlocked
from ..metric.bounds import spd_bounds
from ..units import HBAR_EV_FS


def hermitian(matrix, name):
    m = sp.csr_matrix(matrix, dtype=np.complex128)
    if m.shape[0] != m.shape[1]:
        raise KuboMSDError(f"{name} must be square")
    if not np.all(np.isfinite(m.data)):
        raise KuboMSDError(f"{name} has nonfinite entries")
    scale = max(float(abs(m).max()) if m.nnz else 0.0, np.finfo(float).tiny)
    if m.nnz and float(abs(m - m.conj().T).max()) > 1e-12 * scale:
        raise KuboMSDError(f"{name} is not Hermitian")
    m.sum_duplicates()
    m.sort_indices()
    return m


def atom_block_cholesky(S, slices):
    """[(a, b, L)] with L the lower Cholesky factor of S[a:b, a:b] for each
    contiguous slice (reference metric/setup.py:73-76), factored in one
    batched call per block size instead of one per atom."""
    from ..metric.preconditioners import block_cholesky
    bounds = [(int(a), int(b)) for a, b in slices]
    if [i for a, b in bounds for i in range(a, b)] != list(range(S.shape[0])):
        raise NumericalError("atom_slices must partition orbitals")
    factors = block_cholesky(S, [np.arange(a, b) for a, b in bounds], "atom block")
    return [(a, b, L) for (a, b), L in zip(bounds, factors)]


def _gram(P, Q, rows=1 << 16):
    """P^H Q for tall blocks, summed over row chunks: library GEMMs with an
    inner dimension of millions run far below their speed on chunks."""
    out = None
    for r0 in range(0, P.shape[0], rows):
        part = P[r0:r0 + rows].conj().T @ Q[r0:r0 + rows]
        out = part if out is None else out + part
    return out


class DeviceMetric:
    """Host-side metric set-up plus the device solves (reference MetricSolver,
    src/kubo_msd/metric/solve.py:11-174 and metric/setup.py:40-91): batched
    PCG with the diagonal, atom-block Jacobi or IC(0)
    preconditioner, or the direct solver (metric_solver='cholesky': one dense
    factor of S plus refinement on the original equation)."""

    def __init__(self, ctx, S, config, atom_slices=None):
        if config.preconditioner == "full-cholesky":
            raise KuboMSDError("full-cholesky is supported only for CPU EXACT")
        self.config = config
        self.host_S = S
        n = S.shape[0]
        self.direct = config.metric_solver == "cholesky"
        self.factor_seconds = 0.0
        blocks = []
        if self.direct:
            # metric/cholesky.py:12-20: budget the dense factor before forming it.
            storage = S.data.nbytes + S.indices.nbytes + S.indptr.nbytes
            if 2 * storage + 4 * 16 * n * n > config.memory_limit_gb * 1024 ** 3:
                raise NumericalError("Cholesky factor/workspace exceeds memory budget")
            self.s_lower, self.s_upper, self.bounds_report = spd_bounds(S, allow_dense=False)
            t0 = time.perf_counter()
            try:
                L = la.cholesky(S.toarray(), lower=True, check_finite=False)
            except la.LinAlgError as exc:
                raise NumericalError("overlap is not positive definite (Cholesky factorization)") from exc
            if not np.isfinite(L).all():
                raise NumericalError("Cholesky produced nonfinite factor")
            self.factor_seconds = time.perf_counter() - t0
            self.factor_bytes = L.nbytes
            blocks = [(0, n, L)]
        else:
            self.s_lower, self.s_upper, self.bounds_report = spd_bounds(S)
        self.diagonal = S.diagonal().real.copy()
        if config.preconditioner == "block-jacobi":
            slices = atom_slices or [(j, j + 1) for j in range(n)]
            if [i for a, b in slices for i in range(a, b)] != list(range(n)):
                raise NumericalError("atom_slices must partition orbitals")
            blocks.extend(atom_block_cholesky(S, slices))
        self.blocks = blocks
        self.device = HipMetric(ctx, S, self.diagonal, blocks)
        self.preconditioner_data = None
        if self.direct:
            self.device.set_direct(True, config.metric_refine_maxiter)
        if config.preconditioner == "ic0":
            from ..metric.preconditioners import prepare_ic0_validated
            self.preconditioner_data = prepare_ic0_validated(ctx, S, ordering=config.metric_ic_ordering,
                                                             shift_max=config.metric_ic_shift_max)
            d = self.preconditioner_data
            self.device.set_ic0(d.factor, d.permutation, d.scaling)
        self.device.set_residual_policy(config.metric_residual_policy, config.metric_check_interval,
                                        config.metric_gap_ratio)
        # mixed precision: the complex128 solves (the time propagation's
        # state) also take the FP64 refinement with FP32 inner solves
        if config.precision == "mixed" and not self.direct:
            self.device.set_mixed(True)
        ctx.set_tuning("pcg_poll", max(1, int(config.metric_gpu_check_batch)) if
                       config.metric_gpu_check_batch > 1 else 4)
        # The device skips solves only for S = I with the diagonal preconditioner.
        self.identity = bool(S.nnz == n and np.all(S.diagonal() == 1) and config.preconditioner == "diagonal"
                             and not self.direct)

    def solve_s(self, B, config=None):
        cfg = self.config if config is None else config
        rtol = getattr(cfg, "rtol", None)
        rtol = cfg.metric_rtol if rtol is None else rtol
        out, rep = self.device.solve(B, max(rtol, 0), cfg.metric_atol, cfg.metric_maxiter)
        return out, MetricSolveReport(np.asarray(rep["converged"], bool), np.asarray(rep["iterations"], np.int64),
                                      np.asarray(rep["true_residual"]), np.asarray(rep["relative_residual"]),
                                      np.asarray(rep["breakdown"], bool))

    def diagnostics(self):
        d = self.device.diagnostics()
        d.update(kernel="hip", bounds=self.bounds_report, condition_upper_or_estimate=self.s_upper / self.s_lower,
                 solver=self.config.metric_solver,
                 resolved_preconditioner="not-applicable" if self.direct else self.config.preconditioner)
        if self.direct:
            d.update(factorizations=1, factor_bytes=int(self.factor_bytes), factor_seconds=self.factor_seconds,
                     factor_backend="host LAPACK (solves on the device)")
        if self.preconditioner_data is not None:
            d["preconditioner_setup"] = self.preconditioner_data.report
        return d


class OperatorBundle:
    def __init__(self, model, config, ctx=None):
        import torch
        self.config = config
        self.n = model.n_orbitals
        self.host_H = hermitian(model.H_cov, "Hamiltonian")
        self.host_S = hermitian(model.S, "overlap")
        if self.host_H.shape != self.host_S.shape:
            raise KuboMSDError("H/S dimensions disagree")
        budget = int(config.device_memory_limit_gb * 1024 ** 3)
        self.ctx = ctx or HipContext(config.device, budget)
        self.dtype = torch.complex128 if config.precision == "fp64" else torch.complex64
        self.metric = DeviceMetric(self.ctx, self.host_S, config, model.atom_slices)
        self.velocity = {a: hermitian(v, "velocity " + a) for a, v in model.velocity.items()}
        self.position = {a: hermitian(v, "position " + a) for a, v in model.position.items()}
        self.composite = {a: tuple(sp.csr_matrix(v, dtype=np.complex128) for v in parts)
                          for a, parts in model.composite.items()}
        self.directions = {a: v for a, v in zip(config.axes, config.axes_vectors)}
        for direction in config.axes_vectors:
            for axis, c in zip("xyz", direction):
                if c and axis not in self.velocity and axis not in self.position and axis not in self.composite:
                    raise KuboMSDError(f"missing physical velocity/position/dipole for {axis}")
        for axis, parts in self.composite.items():
            if len(parts) != 3:
                raise KuboMSDError("composite velocity requires bond/xbar/delta")
        definitions = []
        for axis in "xyz":
            if axis in self.velocity:
                definitions.append((0, [self.velocity[axis]]))
            elif axis in self.position:
                definitions.append((1, [self.position[axis]]))
            elif axis in self.composite:
                definitions.append((2, list(self.composite[axis])))
            else:
                definitions.append((-1, []))
        self.H = HipMatrix(self.ctx, self.host_H)
        self.tuning = None
        if config.autotune:
            # Launch shape for the dominant products (moment chains on one
            # sample batch); remembered per matrix signature, so reruns and
            # serving mixes use the same shape.
            # With a nonidentity metric the S products of the inner
            # (complex64: FP32-valued) solves dominate, so those are timed.
            try:
                if self.metric.identity:
                    self.tuning = self.ctx.autotune(self.H, config.sample_batch, self.dtype)
                else:
                    self.tuning = self.ctx.autotune(self.metric.device.S, config.sample_batch, self.dtype,
                                                    f32_values=True, tile=True)
            except ResourceBlocked:
                self.tuning = None
        self.native = HipEngine(self.ctx, self.H, self.metric.device, definitions, config.metric_rtol,
                                config.metric_atol, config.metric_maxiter, HBAR_EV_FS)
        self.S = self.metric.device.S
        self.sampler = None
        self.velocity_hermiticity = {}
        rng = np.random.default_rng(551)
        probe = np.eye(self.n, dtype=complex) if self.n <= 64 else \
            rng.normal(size=(self.n, 8)) + 1j * rng.normal(size=(self.n, 8))
        for axis in config.axes:
            self.velocity_hermiticity[axis] = self._check_velocity(axis, probe)

    # -- helpers -------------------------------------------------------------
    def asarray(self, a, dtype=None):
        import torch
        if isinstance(a, torch.Tensor):
            return a.to(device=self.ctx.torch_device, dtype=dtype or self.dtype).contiguous()
        return self.ctx.upload(np.asarray(a, np.complex128), dtype or self.dtype).contiguous()

    def _check_velocity(self, axis, host_probe):
        """Hermiticity of the constructed velocity in FP64 (reference
        velocity.py:57-88): relative error of P^H v P below 1e-10, retrying
        failed stochastic PCG probes with tighter solves."""
        import torch
        cfg = self.config
        rtol, atol = cfg.metric_rtol, cfg.metric_atol
        attempts = []
        probe = self.asarray(host_probe, torch.complex128)
        controls = None
        error = float("nan")
        for attempt in range(3):
            try:
                v = self.apply_v_cov(axis, probe, controls)
            except NumericalError as exc:
                attempts.append(dict(metric_rtol=rtol, metric_atol=atol, status="S_SOLVE_FAILED", reason=str(exc)))
                # keep the failed solve's per-column report next to the probe's attempts
                exc.diagnostics = {**(getattr(exc, "diagnostics", None) or {}),
                                   "velocity_probe": dict(axis=axis, threshold=1e-10, attempts=attempts)}
                raise
            V = self.ctx.to_host(_gram(probe, v), status=True)
            error = float(np.linalg.norm(V - V.conj().T) / max(np.linalg.norm(V), 1e-300))
            attempts.append(dict(metric_rtol=rtol, metric_atol=atol, relative_error=error))
            if np.isfinite(error) and error <= 1e-10:
                return dict(relative_error=error, threshold=1e-10, attempts=attempts,
                            mode="full" if self.n <= 64 else "8-vector bilinear probe")
            if not np.isfinite(error) or attempt == 2:
                break
            rtol = min(rtol * .1, 1e-12 if attempt == 0 else 1e-13)
            atol *= .1
            controls = SimpleNamespace(metric_rtol=rtol, metric_atol=atol, metric_maxiter=cfg.metric_maxiter)
        err = KuboMSDError(f"constructed velocity {axis} is not Hermitian (relative error {error:g}; "
                           f"probe attempts {attempts})")
        err.diagnostics = {"velocity_probe": dict(axis=axis, threshold=1e-10, attempts=attempts)}
        raise err

    # -- actions (reference velocity.py:90-117) --------------------------------
    @property
    def h_calls(self):
        return self.native.h_calls

    @property
    def v_calls(self):
        return self.native.v_calls

    def _dev(self, B):
        """Device block: torch tensors keep their complex dtype, host arrays
        (the reference's NumPy calls) become the precision's storage type."""
        import torch
        if isinstance(B, torch.Tensor) and B.is_complex():
            return self.asarray(B, B.dtype)
        return self.asarray(B)

    def solve_s(self, B, config=None):
        return self.metric.solve_s(self._dev(B), config)

    def apply_h_cov(self, B):
        return self.H.matmul(self._dev(B))

    def apply_h_action(self, B):
        return self.native.h_action(self._dev(B))

    def apply_v_cov(self, axis, B, config=None):
        direction = self.directions.get(axis, np.eye(3)["xyz".index(axis)] if axis in "xyz" else None)
        if direction is None:
            raise KuboMSDError("unknown direction")
        B = self._dev(B)
        if config is not None:
            return self.native.v_cov_with_controls(B, direction, config.metric_rtol, config.metric_atol,
                                                   config.metric_maxiter)
        return self.native.v_cov(B, direction)

    def apply_v_chain(self, axis, B, length, scale):
        """(length, N, nb): v_cov(axis, T_k(H~) B), H~ = (S^-1 H - c)/a with
        scale = (c, a) (kmsd_chain_v)."""
        direction = self.directions.get(axis, np.eye(3)["xyz".index(axis)] if axis in "xyz" else None)
        if direction is None:
            raise KuboMSDError("unknown direction")
        return self.native.chain_v(B, length, scale[0], scale[1], direction)

    def apply_v_action(self, axis, B):
        return self.solve_s(self.apply_v_cov(axis, B))[0]

    def inner(self, A, B):
        """Per-column <A, S B> on the device (complex128, kmsd_metric_inner)."""
        return self.metric.device.inner(self._dev(A), self._dev(B))


def build_operators(model, config, ctx=None):
    if config.backend != "hip":
        raise KuboMSDError("this package implements the HIP backend (backend='hip')")
    return OperatorBundle(model, config, ctx)
