"""Jackson kernel and the Chebyshev-moment -> energy reconstruction matrix.

Same definitions as the reference (src/kubo_msd/kernels/energy.py:6-37):
  g_m = ((M+1-m) cos(m theta) + sin(m theta) cot(theta)) / (M+1), theta = pi/(M+1), g_0 = 1
  G[E, m] = cos(m arccos x_E) g_m (2 if m > 0 else 1) / (pi a sqrt(1 - x_E^2)),  x_E = (E - c)/a
with |x_E| < 1 required strictly.  These are fixed by the KPM method itself
(Weisse et al., Rev. Mod. Phys. 78, 275 (2006)), so the formulas coincide.
"""
import numpy as np

from ..errors import KuboMSDError


def validate_spectral_window(eps, bounds):
    e = np.asarray(eps, float)
    b = np.asarray(bounds, float)
    if b.shape != (2,) or not np.isfinite(b).all() or b[1] <= b[0]:
        raise KuboMSDError("invalid spectral bounds")
    if not np.isfinite(e).all() or np.any(e < b[0]) or np.any(e > b[1]):
        raise KuboMSDError("spectral bounds do not contain full spectrum")
    return float((b[0] + b[1]) / 2), float((b[1] - b[0]) / 2)


def jackson_coefficients(n_moments):
    M = int(n_moments)
    if M != n_moments or M < 1:
        raise KuboMSDError("n_moments must be positive integer")
    m = np.arange(M)
    theta = np.pi / (M + 1)
    g = ((M + 1 - m) * np.cos(m * theta) + np.sin(m * theta) / np.tan(theta)) / (M + 1)
    g[0] = 1.0
    return g


def reconstruction_matrix(E, bounds, M):
    """(ne, M) float64: DOS(E) = G @ mu."""
    c, a = validate_spectral_window([], bounds)
    x = (np.asarray(E, float) - c) / a
    if not np.isfinite(x).all() or np.any(np.abs(x) >= 1):
        raise KuboMSDError("Jackson output energies must lie strictly inside full spectral bounds")
    g = jackson_coefficients(M)
    g[1:] *= 2
    return np.cos(np.arccos(x)[:, None] * np.arange(M)) * g / (np.pi * a * np.sqrt(1 - x * x)[:, None])


def dos_floor(n_orbitals, width, cfg):
    """Validity floor on the DOS (src/kubo_msd/kernels/energy.py:35-37)."""
    w = cfg.sigma_eV if cfg.kernel == "gaussian" else (cfg.spectral_bounds[1] - cfg.spectral_bounds[0]) / (
        2 * cfg.n_moments)
    return max(cfg.rho_abs, cfg.rho_rel * n_orbitals / max(width, w))


def energy_weights(E, eps, cfg):
    """(len(E), len(eps)) kernel weights of levels eps at energies E
    (src/kubo_msd/kernels/energy.py:26-33): a normalized Gaussian of width
    sigma_eV, or the Jackson-damped Chebyshev reconstruction of delta(E - eps)
    with the resolved spectral bounds."""
    E = np.asarray(E, float)
    eps = np.asarray(eps, float)
    if not np.isfinite(E).all() or not np.isfinite(eps).all():
        raise KuboMSDError("nonfinite energy")
    if cfg.kernel == "gaussian":
        return np.exp(-.5 * ((E[:, None] - eps[None, :]) / cfg.sigma_eV) ** 2) / (np.sqrt(2 * np.pi) * cfg.sigma_eV)
    if cfg.spectral_bounds is None:
        raise KuboMSDError("Jackson requires resolved full spectral bounds")
    c, a = validate_spectral_window(eps, cfg.spectral_bounds)
    T = np.cos(np.arange(cfg.n_moments)[:, None] * np.arccos((eps - c) / a)[None, :])
    return reconstruction_matrix(E, cfg.spectral_bounds, cfg.n_moments) @ T


def kernel_resolution(cfg):
    """Energy resolution reported with every result
    (src/kubo_msd/kernels/energy.py:39-54): the Gaussian's constant FWHM, or
    per output energy the full width at half maximum of the Jackson kernel
    centred there (None where the half-height crossings are not isolated in
    the search bracket of +-8 pi a / M)."""
    from scipy.optimize import brentq, minimize_scalar
    if cfg.kernel == "gaussian":
        return {"type": "constant_fwhm", "fwhm_eV": float(2 * np.sqrt(2 * np.log(2)) * cfg.sigma_eV)}
    c, a = validate_spectral_window([], cfg.spectral_bounds)
    widths = []
    M = cfg.n_moments
    reach = 8 * np.pi * a / M
    g = jackson_coefficients(M)
    g[1:] *= 2
    m = np.arange(M)
    for E in cfg.energy_eV:
        left, right = max(c - a + 1e-10 * a, E - reach), min(c + a - 1e-10 * a, E + reach)
        validate_spectral_window([E], cfg.spectral_bounds)
        wE = g * np.cos(m * np.arccos((E - c) / a))  # the level's Chebyshev weights, fixed per E

        def f(x, wE=wE):  # energy_weights([x], [E], cfg)[0, 0]
            u = (x - c) / a
            return float(np.dot(np.cos(m * np.arccos(u)), wE) / (np.pi * a * np.sqrt(1 - u * u)))

        peak = minimize_scalar(lambda x: -f(x), bounds=(left, right), method="bounded").x
        half = f(peak) / 2
        try:
            lo = brentq(lambda x: f(x) - half, left, peak)
            hi = brentq(lambda x: f(x) - half, peak, right)
            widths.append(float(hi - lo))
        except ValueError:
            widths.append(None)
    return {"type": "local_jackson_fwhm", "fwhm_eV": widths,
            "null_reason": "no isolated half-height crossings in search bracket"}
