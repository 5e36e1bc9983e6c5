"""D = 1/2 dMSD/dt and VAC integration (reference src/kubo_msd/observables/diffusion.py:4-14)."""
import numpy as np

from ..config import grid


def msd_gradient(time_fs, msd, axis=1):
    if len(time_fs) < 2:
        return None
    return .5 * np.gradient(msd, np.asarray(time_fs), axis=axis, edge_order=2 if len(time_fs) > 2 else 1)


def integrate_vac(time_fs, C, axis=1):
    """M(t) and D(t) from a piecewise-linear VAC C(t): D' = C, M' = 2 D."""
    t = grid(time_fs, "integration time", True)
    c = np.moveaxis(C, axis, 0)
    M = np.zeros_like(c)
    D = np.zeros_like(c)
    for j, h in enumerate(np.diff(t)):
        M[j + 1] = M[j] + 2 * h * D[j] + h * h / 3 * (2 * c[j] + c[j + 1])
        D[j + 1] = D[j] + h / 2 * (c[j] + c[j + 1])
    return np.moveaxis(M, 0, axis), np.moveaxis(D, 0, axis)
