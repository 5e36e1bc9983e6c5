"""The square velocity-correlation double moments (solvers/velocity_correlation.py)
against the FP64 oracle: A = chain(S^-1 V r), B = V chain(r) from the oracle
engine (bit-identical to the reference's compiled kernel) and the reference's
own contraction einsum('jnb,knb->bjk', conj(A), B) (src/kubo_msd/solvers/
spectral.py:37).  Tolerances (largest entry error over largest entry): FP64
tables on complex128 chains 1e-12; on complex64 chains 5e-7 (the chains'
storage rounding); FP32 MFMA 5e-6 and bf16x3 1e-4 on complex64 chains (the
product and accumulation rounding measured for the tables alone: 2e-6, 1e-4)."""
import numpy as np
import pytest

from helpers import dense_model, phases
from kubomsd_amd.synthetic import honeycomb_lattice
from oracle import kpm
from test_gpu_solvers import bundle, run_config

pytestmark = pytest.mark.gpu


def oracle_mu(H, S, V, r, M, bounds):
    ops = kpm.Ops(H, S, {"x": V}, rtol=1e-13)
    c, a = (bounds[0] + bounds[1]) / 2, (bounds[1] - bounds[0]) / 2
    d = np.array([1.0, 0, 0])
    A = ops.native.chain(np.ascontiguousarray(ops.solve(ops.apply_v_cov(d, r))), M, c, a)
    T = ops.native.chain(np.ascontiguousarray(r), M, c, a)
    n, nb = r.shape
    B = ops.apply_v_cov(d, T.transpose(1, 0, 2).reshape(n, -1)).reshape(n, M, nb).transpose(1, 0, 2)
    return np.einsum("jnb,knb->bjk", A.conj(), B)


@pytest.mark.parametrize("precision,mode,tol", [("fp64", "fp64", 1e-12), ("mixed", "fp64", 5e-7),
                                                ("mixed", "fp32", 5e-6), ("mixed", "bf16x3", 1e-4)])
@pytest.mark.parametrize("model", ["honeycomb", "nonorthogonal"])
def test_velocity_correlation_matches_oracle(hip_ctx, precision, mode, tol, model):
    from kubomsd_amd.solvers.velocity_correlation import velocity_correlation_moments
    if model == "honeycomb":
        m = honeycomb_lattice(12, 10, seed=12)
        H, S, V, bounds, M = m.H, m.S, m.velocity["x"], (-8.8, 8.8), 96
        r = phases(m.n, 3, 41)
    else:
        d = dense_model(9, 903)
        H, S, V = d.H, d.S, d.velocity["x"]
        ev = np.linalg.eigvals(np.linalg.solve(S.toarray(), H.toarray())).real
        bounds, M = (float(ev.min()) - .3, float(ev.max()) + .3), 24
        r = np.random.default_rng(5).normal(size=(9, 2)) + 0j
    cfg = run_config(energy_eV=[0.], time_fs=[0.], spectral_bounds=bounds, n_moments=M, axes=("x",),
                     precision=precision)
    ops = bundle(hip_ctx, H, S, {"x": V}, cfg)
    out = velocity_correlation_moments(ops, cfg, r, "x", mode=mode, reference=True)
    got = out["mu"].cpu().numpy()
    want = oracle_mu(H, S, V, r, M, bounds)
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    assert got.shape == (r.shape[1], M, M)
    assert err <= tol, err
    assert len(out["error_vs_fp64"]) == r.shape[1] and max(out["error_vs_fp64"]) <= tol
