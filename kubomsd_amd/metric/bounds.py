"""SPD checks and spectral bounds (reference src/kubo_msd/metric/bounds.py:7-58).

spd_bounds: Gershgorin when certifying, else small dense eigenvalues (n <= 128)
or the smallest Lanczos Ritz value minus 10 residuals.
certified_bounds: Gershgorin bounds of S^-1 H with no eigensolver.  The
generalized Ritz extrema and the user-bound validation run on the device
(metric/device_bounds.py).
"""
import numpy as np
import scipy.sparse as sp

from ..errors import NumericalError


def spd_bounds(S, *, allow_dense=True):
    S = sp.csr_matrix(S)
    diag = S.diagonal().real
    radii = np.asarray(abs(S).sum(axis=1)).ravel() - np.abs(diag)
    lo = float(np.min(diag - radii))
    hi = float(np.max(diag + radii))
    if lo > 0:
        return lo, hi, {"source": "Gershgorin", "certified": True}
    n = S.shape[0]
    if n <= 2 and not allow_dense:
        lo = float(diag[0]) if n == 1 else float((diag[0] + diag[1]) / 2 - np.hypot((diag[0] - diag[1]) / 2,
                                                                                    abs(S[0, 1])))
        source = "analytic one/two dimensional SPD check"
    elif n <= 128 and allow_dense:
        ev = np.linalg.eigvalsh(S.toarray())
        lo, hi = float(ev[0]), float(ev[-1])
        source = "small-matrix eigenvalues of S"
    else:
        v0 = np.random.default_rng(9021).normal(size=n)
        e, V = sp.linalg.eigsh(S, k=1, which="SA", tol=1e-12, v0=v0)
        residual = float(np.linalg.norm(S @ V[:, 0] - e[0] * V[:, 0]))
        lo = float(e[0] - 10 * residual)
        source = "smallest Ritz value of S minus 10 residuals (estimated)"
    if lo <= 0 or not np.isfinite([lo, hi]).all():
        raise NumericalError("overlap is not positive definite or SPD verification failed")
    return lo * (1 - 1e-10), hi * (1 + 1e-10), {"source": source, "certified": n <= 2 or (n <= 128 and allow_dense)}


def certified_bounds(host_H, host_S, s_lower, margin=1e-3, spectral_bounds=None):
    """Bounds that provably contain the generalized spectrum of (H, S) without
    an eigensolver: Gershgorin discs of H when S = I, otherwise
    +-||H||_inf / lambda_min(S) (the reference's conservative bound,
    bounds.py:28-32), widened by a relative margin.  spectral_bounds: user
    bounds, returned as given when they enclose those certified bounds
    (else NumericalError)."""
    H = sp.csr_matrix(host_H)
    S = sp.csr_matrix(host_S)
    n = H.shape[0]
    identity = S.nnz == n and np.all(S.diagonal() == 1)
    if identity:
        d = H.diagonal().real
        r = np.asarray(abs(H).sum(axis=1)).ravel() - np.abs(H.diagonal())
        lo, hi = float(np.min(d - r)), float(np.max(d + r))
        source = "Gershgorin discs of H (S = I)"
    else:
        b = float(np.asarray(abs(H).sum(axis=1)).max()) / s_lower
        lo, hi = -b, b
        source = "||H||_inf / lambda_min(S)"
    w = max(hi - lo, 1e-12)
    bounds = (lo - margin * w, hi + margin * w)
    conservative = bounds
    if spectral_bounds is not None:
        user = tuple(float(v) for v in spectral_bounds)
        if not (user[0] < lo and user[1] > hi):
            raise NumericalError("requested spectral bounds do not enclose the certified spectrum bounds")
        bounds = user
        source = "user bounds enclosing " + source
    return {"bounds": bounds, "source": source, "certified": True, "residuals": [],
            "conservative_bounds": conservative, "extrema": [lo, hi]}
