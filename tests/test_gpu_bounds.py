"""Device spectral bounds (two-start Lanczos in the S inner product with
true Ritz residuals) against dense generalized eigenvalues, and the
reference's validation of user bounds (src/kubo_msd/metric/bounds.py:49-52,
kubo_msd_gpu/src/kubo_msd/metric/gpu_bounds.py:53-66)."""
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as la

from helpers import dense_model
from kubomsd_amd.synthetic import cubic_orbital_model, honeycomb_lattice

pytestmark = pytest.mark.gpu


def bundle(ctx, H, S, V, **kw):
    from kubomsd_amd.config import RunConfig
    from kubomsd_amd.contracts import ModelBundle
    from kubomsd_amd.operators.velocity import build_operators
    cfg = RunConfig(energy_eV=[0.0], time_fs=[0.0], solver="stochastic-kpm", kernel="jackson", backend="hip",
                    metric_rtol=1e-13, axes=("x",), autotune=False, **kw)
    return build_operators(ModelBundle(H, S, velocity=V), cfg, ctx)


@pytest.mark.parametrize("case", ["dense9", "honeycomb", "cubic"])
def test_ritz_extrema_match_dense(hip_ctx, case):
    from kubomsd_amd.metric.device_bounds import estimate_spectral_bounds_device, ritz_extrema
    if case == "dense9":
        m = dense_model(9, 903)
    elif case == "honeycomb":
        m = honeycomb_lattice(6, t=-2.7, W=1.0)
    else:
        m = cubic_orbital_model(2, seed=4411)
    ops = bundle(hip_ctx, m.H, m.S, {"x": m.velocity["x"]})
    ev = la.eigh(m.H.toarray(), m.S.toarray(), eigvals_only=True)
    r = ritz_extrema(ops)
    scale = max(abs(ev[0]), abs(ev[-1]))
    assert np.max(np.abs(r["extrema"][:, 0] - ev[0])) <= 1e-9 * scale
    assert np.max(np.abs(r["extrema"][:, 1] - ev[-1])) <= 1e-9 * scale
    rep = estimate_spectral_bounds_device(ops)
    lo, hi = rep["bounds"]
    assert lo < ev[0] and hi > ev[-1] and rep["extrema"][0] <= ev[0] + 1e-9 * scale


def test_user_bounds_validated(hip_ctx):
    from kubomsd_amd.errors import NumericalError
    from kubomsd_amd.metric.device_bounds import estimate_spectral_bounds_device
    m = honeycomb_lattice(6, t=-2.7, W=1.0)
    ops = bundle(hip_ctx, m.H, m.S, {"x": m.velocity["x"]})
    ev = np.linalg.eigvalsh(m.H.toarray())
    ok = estimate_spectral_bounds_device(ops, (ev[0] - .5, ev[-1] + .5))
    assert ok["bounds"] == (ev[0] - .5, ev[-1] + .5)
    with pytest.raises(NumericalError, match="enclose"):
        estimate_spectral_bounds_device(ops, (ev[0] + .1, ev[-1] + .5))


def test_two_level_example_breakdown(hip_ctx):
    """n = 2: the Krylov space closes after two steps (exact Ritz values)."""
    from kubomsd_amd.io.model import load_model
    from kubomsd_amd.metric.device_bounds import ritz_extrema
    model = load_model(Path(__file__).parent / "golden" / "examples" / "nonorthogonal" / "model.json")
    from kubomsd_amd.config import RunConfig
    from kubomsd_amd.operators.velocity import build_operators
    cfg = RunConfig(energy_eV=[0.0], time_fs=[0.0], solver="stochastic-kpm", kernel="jackson", backend="hip",
                    axes=("x",), autotune=False)
    ops = build_operators(model, cfg, hip_ctx)
    ev = la.eigh(ops.host_H.toarray(), ops.host_S.toarray(), eigvals_only=True)
    r = ritz_extrema(ops)
    np.testing.assert_allclose(r["extrema"], np.tile(ev, (2, 1)), rtol=0, atol=1e-12 * abs(ev).max())
    assert r["steps"] <= 2
