"""NumPy restatement of the reference's stochastic KPM orchestration on the
FP64 C++ oracle engine.  TEST INFRASTRUCTURE ONLY (tests/ and bench.py's CPU
baseline import it; the product never does).

Follows, step for step:
  scaling / bessel_coefficients / projected_quadratic  src/kubo_msd/solvers/chebyshev.py:8-61
  propagate / stepwise_samples                          src/kubo_msd/solvers/stepwise.py:8-38
  spectral_samples                                      src/kubo_msd/solvers/spectral.py:15-53
  jackson / reconstruction_matrix                       src/kubo_msd/kernels/energy.py:12-24
  resolve_time_grid                                     src/kubo_msd/config.py:19-24
Pinned by tests/test_oracle_orchestration.py against the independent
eigen-basis closed forms in oracle/eigen.py (complete metric basis => exact).
"""
from types import SimpleNamespace

import numpy as np
from scipy.special import gammaln, jv

from . import native as O

HBAR = O.HBAR_EV_FS


def time_grid(tmax, dt):
    t = np.arange(int(np.floor(tmax / dt)) + 1, dtype=float) * dt
    if t[-1] < tmax and not np.isclose(t[-1], tmax, rtol=0, atol=8 * np.spacing(max(tmax, dt))):
        t = np.append(t, tmax)
    else:
        t[-1] = tmax
    return t


def jackson(M):
    m = np.arange(M)
    th = np.pi / (M + 1)
    g = ((M + 1 - m) * np.cos(m * th) + np.sin(m * th) / np.tan(th)) / (M + 1)
    g[0] = 1.0
    return g


def reconstruction_matrix(E, bounds, M):
    lo, hi = bounds
    c, a = (lo + hi) / 2, (hi - lo) / 2
    x = (np.asarray(E, float) - c) / a
    g = jackson(M)
    g[1:] *= 2
    return np.cos(np.arccos(x)[:, None] * np.arange(M)) * g / (np.pi * a * np.sqrt(1 - x * x)[:, None])


def scaling(cfg):
    lo, hi = cfg.spectral_bounds
    return (lo + hi) / 2, (hi - lo) / 2


def bessel_coefficients(dt, cfg, weighted_tail=False):
    c, a = scaling(cfg)
    z = a * dt / HBAR
    az = abs(z)
    if az == 0:
        return np.array([1 + 0j])
    for k in range(max(1, int(np.ceil(az))), cfg.propagation_max_order + 1):
        n = k + 1
        r = az / (2 * (n + 1))
        if r >= 1:
            continue
        bound = 2 * np.exp(min(n * np.log(az / 2) - gammaln(n + 1) + az * az / (4 * (n + 1)), 700)) / (1 - r)
        if weighted_tail:
            bound *= 4 * (n + 1) ** 2 / (1 - r) ** 2
        if bound <= cfg.propagation_tol:
            m = np.arange(k + 1)
            b = jv(m, z) * (-1j) ** m
            b[1:] *= 2
            return b * np.exp(-1j * c * dt / HBAR)
    raise ValueError("propagation Bessel tail not converged before max order")


class Ops:
    """Oracle operator bundle: engine + metric + explicit velocities (kind 0)
    or, for the axes in `position`, position matrices (kind 1: the velocity is
    the commutator (i/hbar)(H S^-1 X - X S^-1 H), native.cpp:506-533)."""

    def __init__(self, H, S, velocity, rtol=1e-12, threads=1, blocks=(), position=None):
        import scipy.sparse as sp
        self.host_H = sp.csr_matrix(H, dtype=complex)
        self.host_S = sp.csr_matrix(S, dtype=complex)
        self.n = self.host_H.shape[0]
        self.S = O.Csr(self.host_S, threads)
        self.metric = O.Metric(self.S, self.host_S.diagonal().real, blocks, threads)
        position = position or {}
        defs = [(0, [O.Csr(velocity[a], threads)]) if a in velocity else
                (1, [O.Csr(position[a], threads)]) if a in position else (-1, []) for a in "xyz"]
        self.native = O.Engine(O.Csr(self.host_H, threads), self.metric, defs, rtol, 0., 1000)
        self.rtol = rtol

    def inner(self, A, B):
        return np.sum(A.conj() * self.S.matmul(B), axis=0)

    def apply_v_cov(self, direction, B):
        return self.native.v_cov(B, direction)

    def solve(self, B):
        return self.metric.solve(B, self.rtol, 0., 1000)[0]


def config(**kw):
    base = dict(n_moments=64, propagation_tol=1e-12, propagation_max_order=4096, segment_fs=1.0,
                integration_dt_fs=0.01, include_vac=True)
    base.update(kw)
    base["energy_eV"] = np.asarray(base["energy_eV"], float)
    base["time_fs"] = np.asarray(base["time_fs"], float)
    base["axes_vectors"] = np.asarray(base.get("axes_vectors", np.eye(3)[:1]), float)
    return SimpleNamespace(**base)


def projected_quadratic(B, ops, cfg):
    c, a = scaling(cfg)
    G = reconstruction_matrix(cfg.energy_eV, cfg.spectral_bounds, cfg.n_moments)
    return (G @ ops.native.moments(B, cfg.n_moments, c, a)).T


def propagate(states, dt_fs, ops, cfg):
    B = np.array(states, complex)
    initial = ops.inner(B, B).real
    nseg = max(1, int(np.ceil(abs(dt_fs) / cfg.segment_fs)))
    coeff = bessel_coefficients(dt_fs / nseg, cfg)
    c, a = scaling(cfg)
    for _ in range(nseg):
        B = ops.native.series(B, coeff, c, a)
    final = ops.inner(B, B).real
    drift = float(np.max(np.abs(final - initial) / np.maximum(initial, np.finfo(float).tiny)))
    return B, drift


def stepwise_samples(ops, cfg, samples):
    initial = np.array(samples, complex)
    nb = initial.shape[1]
    na, nt, ne = len(cfg.axes_vectors), len(cfg.time_fs), len(cfg.energy_eV)
    dos = projected_quadratic(initial, ops, cfg)
    total = np.hstack([initial] + [np.zeros_like(initial)] * na)
    out = np.zeros((nb, na, nt, ne))
    c, a = scaling(cfg)
    for j, t in enumerate(cfg.time_fs):
        if j:
            span = t - cfg.time_fs[j - 1]
            count = max(1, int(np.ceil(span / cfg.segment_fs)))
            coeff = bessel_coefficients(span / count, cfg, weighted_tail=True)
            for _ in range(count):
                total = ops.native.series(total, coeff, c, a, cfg.axes_vectors)
        for ai in range(na):
            out[:, ai, j, :] = projected_quadratic(np.ascontiguousarray(total[:, (ai + 1) * nb:(ai + 2) * nb]),
                                                   ops, cfg)
    return dos, {"msd_A2/commutator": out}


def spectral_samples(ops, cfg, samples):
    samples = np.array(samples, complex)
    nb = samples.shape[1]
    na, ne, nt, L = len(cfg.axes_vectors), len(cfg.energy_eV), len(cfg.time_fs), cfg.n_moments
    dos = projected_quadratic(samples, ops, cfg)
    keys = ("msd_A2/vac_integral", "diffusion_A2_per_fs/vac_integral", "vac_A2_per_fs2/vac_integral")
    out = {k: np.zeros((nb, na, nt, ne)) for k in keys}
    times = np.unique(np.r_[time_grid(cfg.time_fs[-1], cfg.integration_dt_fs), cfg.time_fs])
    idx = {float(t): j for j, t in enumerate(cfg.time_fs)}
    G = reconstruction_matrix(cfg.energy_eV, cfg.spectral_bounds, L)
    c, a = scaling(cfg)
    c_anchor = samples
    v_anchor = [ops.solve(ops.apply_v_cov(d, samples)) for d in cfg.axes_vectors]
    anchor, next_anchor, tables = 0., -1., None
    M = np.zeros((nb, na, ne))
    D = np.zeros_like(M)
    previous, prev_t = None, 0.
    for t in times:
        if tables is None or t > next_anchor + 1e-14:
            new_anchor = 0. if tables is None else float(prev_t)
            if new_anchor > anchor:
                both, _ = propagate(np.hstack([c_anchor] + v_anchor), new_anchor - anchor, ops, cfg)
                c_anchor = both[:, :nb]
                v_anchor = [both[:, (i + 1) * nb:(i + 2) * nb] for i in range(na)]
            anchor = new_anchor
            next_anchor = min(max(anchor + cfg.segment_fs, float(t)), float(times[-1]))
            K = len(bessel_coefficients(next_anchor - anchor, cfg)) - 1
            left = ops.native.chain(np.ascontiguousarray(c_anchor), L + K, c, a)
            tables = []
            for ai, d in enumerate(cfg.axes_vectors):
                right = ops.native.chain(np.ascontiguousarray(v_anchor[ai]), K + 1, c, a)
                cov = ops.apply_v_cov(d, right.transpose(1, 0, 2).reshape(ops.n, -1)).reshape(
                    ops.n, K + 1, nb).transpose(1, 0, 2)
                tables.append(np.einsum("jnb,knb->bjk", left.conj(), cov))
        b = bessel_coefficients(float(t - anchor), cfg)
        Kt = len(b) - 1
        vac = np.empty((nb, na, ne))
        l = np.arange(L)[:, None]
        p = np.arange(Kt + 1)[None, :]
        for ai, table in enumerate(tables):
            q = table[:, :, :Kt + 1] @ b
            mom = np.sum(.5 * (q[:, l + p] + q[:, np.abs(l - p)]) * b.conj()[None, None, :], axis=2).real
            vac[:, ai, :] = mom @ G.T
        if previous is not None:
            h = float(t - prev_t)
            M = M + 2 * h * D + h * h / 3 * (2 * previous + vac)
            D = D + h / 2 * (previous + vac)
        if float(t) in idx:
            j = idx[float(t)]
            out["msd_A2/vac_integral"][:, :, j, :] = M
            out["diffusion_A2_per_fs/vac_integral"][:, :, j, :] = D
            out["vac_A2_per_fs2/vac_integral"][:, :, j, :] = vac
        previous, prev_t = vac, t
    return dos, out
