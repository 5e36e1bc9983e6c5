"""Shared builders: the same model fed to the FP64 oracle and the HIP engine."""
import numpy as np
import scipy.sparse as sp

from oracle import native as O
from kubomsd_amd.units import HBAR_EV_FS


def definitions(model, axes="xyz"):
    """(kind, [matrices]) per Cartesian axis, explicit velocities (kind 0)."""
    return [(0, [model.velocity[a]]) if a in model.velocity else (-1, []) for a in axes]


def oracle_engine(model, rtol=1e-10, atol=0.0, maxiter=1000, blocks=(), threads=1):
    S = O.Csr(model.S, threads)
    metric = O.Metric(S, model.S.diagonal().real, blocks, threads)
    return O.Engine(O.Csr(model.H, threads), metric,
                    [(k, [O.Csr(m, threads) for m in ms]) for k, ms in definitions(model)],
                    rtol, atol, maxiter, HBAR_EV_FS)


def hip_engine(ctx, model, rtol=1e-10, atol=0.0, maxiter=1000, blocks=()):
    from kubomsd_amd.backends.hip import HipEngine, HipMetric
    metric = HipMetric(ctx, model.S, model.S.diagonal().real, blocks)
    return HipEngine(ctx, model.H, metric, definitions(model), rtol, atol, maxiter, HBAR_EV_FS)


def random_block(n, b, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, b)) + 1j * rng.normal(size=(n, b))


def phases(n, b, seed):
    rng = np.random.default_rng(seed)
    return np.exp(2j * np.pi * rng.random((n, b)))


def to_dev(ctx, a, dtype="c128"):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a, np.complex128))
    if dtype == "c64":
        t = t.to(torch.complex64)
    return t.to(ctx.torch_device)


def host(t):
    import torch
    return t.detach().to("cpu").to(torch.complex128 if t.is_complex() else torch.float64).numpy()


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def dense_model(n=9, seed=903, kind="velocity"):
    """Small dense nonorthogonal model with complex entries."""
    from kubomsd_amd.synthetic import Model
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    S = np.eye(n) + .02 * (A.conj().T @ A)
    H = (A + A.conj().T) / 12
    X = np.diag(np.linspace(-.5, .8, n)).astype(complex) + .03 * (A + A.conj().T)
    return Model(sp.csr_matrix(H), sp.csr_matrix(S), {"x": sp.csr_matrix(X)}, np.zeros((n, 3)))
