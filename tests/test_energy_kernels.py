"""Energy kernels and the reported resolution (src/kubo_msd/kernels/energy.py
of the reference), pinned by what the mathematics fixes: the Jackson kernel
of a level has unit weight on the spectral window (only T_0 survives the
Chebyshev-weighted integral), the Gaussian is normalized, and the reported
FWHM is the width of the kernel's half-maximum set found by dense sampling."""
import numpy as np
import pytest

from kubomsd_amd.config import RunConfig
from kubomsd_amd.errors import KuboMSDError
from kubomsd_amd.kernels.energy import energy_weights, kernel_resolution


def cfg(energy_eV=np.linspace(-8, 8, 9), **kw):
    base = dict(solver="stochastic-kpm", kernel="jackson", spectral_bounds=(-8.8, 8.8), n_moments=512)
    base.update(kw)
    return RunConfig(energy_eV, [0], **base)


@pytest.mark.parametrize("level", [-8.5, -3.0, 0.0, 6.2])
def test_jackson_kernel_has_unit_weight(level):
    c = cfg()
    a = 8.8
    k = 4096  # Gauss-Chebyshev nodes: exact for the kernel's polynomial degree < 2k
    theta = (np.arange(k) + .5) * np.pi / k
    x = a * np.cos(theta)
    w = energy_weights(x, [level], c)[:, 0]
    # integral over [-a, a] of K(x) dx = a * integral K(a u) sqrt(1-u^2) du / sqrt(1-u^2)
    integral = a * np.pi / k * np.sum(w * np.sqrt(1 - (x / a) ** 2))
    assert abs(integral - 1.0) < 1e-12


def test_gaussian_is_normalized_and_resolution_constant():
    c = RunConfig(np.linspace(-1, 1, 3), [0], solver="exact", kernel="gaussian", sigma_eV=0.05)
    x = np.linspace(-2, 2, 40001)
    w = energy_weights(x, [0.1], c)[:, 0]
    assert abs(np.trapezoid(w, x) - 1.0) < 1e-9
    r = kernel_resolution(c)
    assert r["type"] == "constant_fwhm" and abs(r["fwhm_eV"] - 2.354820045 * 0.05) < 1e-9


def test_jackson_fwhm_equals_sampled_half_maximum_width():
    c = cfg(energy_eV=np.array([-7.0, 0.0, 3.3]), n_moments=1024)
    r = kernel_resolution(c)
    assert r["type"] == "local_jackson_fwhm"
    for E, width in zip(c.energy_eV, r["fwhm_eV"]):
        x = np.linspace(E - 0.1, E + 0.1, 40001)  # spacing 5e-6
        w = np.concatenate([energy_weights(x[i:i + 4000], [E], c)[:, 0] for i in range(0, len(x), 4000)])
        above = x[w >= w.max() / 2]
        assert abs((above[-1] - above[0]) - width) < 1e-5


def test_energy_weights_reject_levels_outside_bounds():
    with pytest.raises(KuboMSDError, match="spectral bounds"):
        energy_weights([0.0], [9.0], cfg())


def test_output_window_keeps_full_spectral_scale():
    """tests/unit/test_jackson_spectral_scale.py: a narrow output window does
    not redefine the spectral scale; independent cosine-series evaluation."""
    energy = np.linspace(-0.5, 0.5, 11)
    levels = np.array([-8., -2., 0., 2., 15.])
    c = RunConfig(energy, [0., 0.1], kernel="jackson", n_moments=64, spectral_bounds=(-10., 20.))
    centre, radius = 5., 15.
    m = np.arange(64)
    ang = np.pi / 65
    damp = ((65 - m) * np.cos(m * ang) + np.sin(m * ang) / np.tan(ang)) / 65
    x = (energy - centre) / radius
    y = (levels - centre) / radius
    want = (np.cos(np.arccos(x)[:, None] * m) * damp * np.where(m == 0, 1, 2)) @ np.cos(m[:, None] * np.arccos(y)[None, :])
    want /= (np.pi * radius * np.sqrt(1 - x * x))[:, None]
    got = energy_weights(energy, levels, c)
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-13)
    assert not np.allclose(got[:, 0], got[:, 1])


@pytest.mark.parametrize("L", [1, 2, 64, 1024])
def test_jackson_area_and_term_count(L):
    """tests/unit/test_energy.py:18-28."""
    from kubomsd_amd.kernels.energy import jackson_coefficients
    g = jackson_coefficients(L)
    assert len(g) == L and g[0] == 1
    theta = (np.arange(4096) + .5) * np.pi / 4096
    E = np.cos(theta)[::-1]
    c = RunConfig(E, [0], kernel="jackson", spectral_bounds=(-1, 1), n_moments=L)
    w = energy_weights(E, [-.37, .2], c)
    np.testing.assert_allclose((w * np.sqrt(1 - E[:, None] ** 2)).sum(0) * np.pi / 4096, 1, atol=2e-12)


def test_full_spectrum_is_independent_of_local_output_and_no_clipping():
    """tests/unit/test_energy.py:5-15, 38-44."""
    from kubomsd_amd.kernels.energy import validate_spectral_window
    eps = np.array([-9., 0., 19.])
    assert validate_spectral_window(eps, (-10., 20.)) == (5., 15.)
    c = RunConfig([-.5, 0, .5], [0, 1], kernel="jackson", spectral_bounds=(-10, 20), n_moments=64)
    a = energy_weights(c.energy_eV, eps, c)
    b = energy_weights([-5., 0., 10.], eps, c)
    np.testing.assert_allclose(a[1], b[1], rtol=1e-13)
    assert not np.allclose(a[:, 0], a[:, 2])
    with pytest.raises(ValueError):
        validate_spectral_window([-2, 0], (-1, 1))
    c = RunConfig([0, 1], [0], kernel="jackson", spectral_bounds=(-1, 1))
    with pytest.raises(ValueError):
        energy_weights([0, 1], [0], c)
