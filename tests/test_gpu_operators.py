"""Operator bundle behaviour the reference's unit tests pin
(tests/unit/test_operator_metric.py, test_velocity_probe_accuracy.py), on the
device: covariant actions of a nonorthogonal open-boundary position model,
composite velocity = open commutator at any origin, and the velocity
Hermiticity probe (retries with stricter solves that never leak into the
transport tolerance, rejection of a non-Hermitian velocity, diagnostics of a
failed stricter solve)."""
import numpy as np
import pytest
import scipy.linalg as la

from helpers import host

pytestmark = pytest.mark.gpu


def cfg(**kw):
    from kubomsd_amd.config import RunConfig
    base = dict(solver="stochastic-kpm", kernel="jackson", backend="hip", axes=("x",))
    base.update(kw)
    return RunConfig([0], [0], **base)


def ops_for(model, c, ctx):
    from kubomsd_amd.operators.velocity import build_operators
    return build_operators(model, c, ctx)


def test_nonorthogonal_open_position_and_covariant_action(hip_ctx):
    from kubomsd_amd.contracts import ModelBundle
    from kubomsd_amd.units import HBAR_EV_FS
    A = np.array([[1, .2j], [.3, 1.2]], complex)
    S = A.conj().T @ A
    H = A.conj().T @ np.diag([-.4, .8]) @ A
    X = A.conj().T @ np.array([[0, 1j], [-1j, 2]]) @ A
    ops = ops_for(ModelBundle(H, S, position={"x": X}), cfg(metric_rtol=1e-13), hip_ctx)
    want = 1j * (H @ la.solve(S, X) - X @ la.solve(S, H)) / HBAR_EV_FS
    B = np.array([[1 + 2j, 0], [.4, 1]], complex)
    np.testing.assert_allclose(host(ops.apply_v_cov("x", B)), want @ B, atol=1e-12)
    np.testing.assert_allclose(S @ host(ops.apply_h_action(B)), H @ B, atol=1e-12)


def test_composite_equals_open_commutator_and_origin_invariance(hip_ctx):
    from kubomsd_amd.contracts import ModelBundle
    H = np.array([[1, .3j], [-.3j, 2]])
    S = np.array([[1, .1], [.1, 1.2]])
    X = np.array([[.1, .4], [.4, 1.5]])
    r = np.array([0, 1])
    c = cfg(metric_rtol=1e-13)
    ref = host(ops_for(ModelBundle(H, S, position={"x": X}), c, hip_ctx).apply_v_cov("x", np.eye(2)))
    for origin in (0, 10):
        rr = r + origin
        XX = X + origin * S
        comp = {"x": (H * (rr[None, :] - rr[:, None]), XX - rr[:, None] * S, XX - S * rr[None, :])}
        got = host(ops_for(ModelBundle(H, S, composite=comp), c, hip_ctx).apply_v_cov("x", np.eye(2)))
        np.testing.assert_allclose(got, ref, atol=1e-12)


def ill_conditioned():
    from kubomsd_amd.contracts import ModelBundle
    rng = np.random.default_rng(51)
    n = 9
    U, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    S = (U * np.geomspace(.001, 1, n)) @ U.conj().T
    S = (S + S.conj().T) / 2
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = (A + A.conj().T) / 2
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    X = (A + A.conj().T) / 2
    return ModelBundle(H, S, position={"x": X})


@pytest.mark.parametrize("representation", ["position", "composite"])
def test_probe_retries_without_changing_transport_tolerance(hip_ctx, representation):
    from kubomsd_amd.contracts import ModelBundle
    from kubomsd_amd.metric.solve import MetricSolver
    c = cfg(metric_rtol=1e-6, metric_residual_policy="adaptive")
    model = ill_conditioned()
    if representation == "composite":
        X = model.position["x"]
        model = ModelBundle(model.H_cov, model.S, composite={"x": (np.zeros_like(X), X, X)})
    ops = ops_for(model, c, hip_ctx)
    report = ops.velocity_hermiticity["x"]
    assert report["attempts"][0]["relative_error"] > 1e-10
    assert report["relative_error"] <= 1e-10
    assert report["attempts"][-1]["metric_rtol"] < c.metric_rtol
    assert ops.config.metric_rtol == 1e-6
    B = np.arange(18).reshape(9, 2) * (1 + .2j) + 1
    want = host(MetricSolver(ops.host_S, c, hip_ctx).solve_s(ops.apply_h_cov(ops.asarray(B, None)))[0])
    np.testing.assert_allclose(host(ops.apply_h_action(B)), want, rtol=5e-12, atol=5e-12)


def test_probe_cannot_accept_a_nonhermitian_velocity(hip_ctx):
    from kubomsd_amd.contracts import ModelBundle
    model = ModelBundle(np.diag([1., 2]), np.eye(2), composite={"x": (np.eye(2), np.zeros((2, 2)), np.zeros((2, 2)))})
    with pytest.raises(ValueError, match="not Hermitian"):
        ops_for(model, cfg(metric_rtol=1e-6), hip_ctx)


def test_failed_stricter_solve_keeps_probe_and_column_diagnostics(hip_ctx):
    c = cfg(metric_rtol=1e-6, metric_residual_policy="adaptive", metric_maxiter=10)
    with pytest.raises(ValueError) as caught:
        ops_for(ill_conditioned(), c, hip_ctx)
    details = caught.value.diagnostics
    assert "relative_residual" in details and not all(details["converged"])
    probe = details["velocity_probe"]
    assert probe["axis"] == "x" and probe["threshold"] == 1e-10
    assert probe["attempts"][0]["relative_error"] > 1e-10
    assert probe["attempts"][-1]["status"] == "S_SOLVE_FAILED"
    assert probe["attempts"][-1]["metric_rtol"] == 1e-12
