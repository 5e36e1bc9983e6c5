"""Moment chains of several engines in shared launches (kmsd_moments_group):
every chain's moments are bit for bit those of its own kmsd_moments call,
for lattices of different sizes and ELL widths, a general CSR matrix, an
imaginary Hermitian matrix and a 4x4-block matrix (which runs alone), chain
lengths around the chunk and buffer-rotation boundaries, both precisions;
and equal to the FP64 oracle."""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import hip_engine, host, oracle_engine, phases, rel_err, to_dev
from kubomsd_amd.synthetic import Model, cubic_orbital_model, honeycomb_flake, square_lattice

pytestmark = pytest.mark.gpu


def models():
    rng = np.random.default_rng(5)
    n = 700
    A = sp.random(n, n, density=0.006, random_state=7, format="csr") + sp.random(n, n, density=0.006,
                                                                                    random_state=8, format="csr") * 1j
    Hc = ((A + A.conj().T) / 2 + sp.diags(rng.uniform(-1, 1, n))).tocsr()
    R = sp.random(n, n, density=0.01, random_state=9, format="csr")
    Hi = ((R - R.T) * 1j).tocsr()  # purely imaginary Hermitian
    eye = sp.identity(n, dtype=np.complex128, format="csr")
    cubic = cubic_orbital_model(4)
    cubic_orth = Model(cubic.H, sp.identity(cubic.n, dtype=np.complex128, format="csr"), {}, cubic.positions)
    return [honeycomb_flake(1000, W=1.0), honeycomb_flake(5000, W=2.0), square_lattice(40, W=1.0),
            honeycomb_flake(20000, W=0.5), Model(Hc, eye, {}, np.zeros((n, 3))), Model(Hi, eye, {}, np.zeros((n, 3))),
            cubic_orth]


def scales(ms):
    out = []
    for m in ms:
        b = float(np.abs(m.H).sum(axis=1).max()) * 1.02
        out.append((0.01, b + 0.02))
    return out


@pytest.mark.parametrize("dtype", ["c64", "c128"])
@pytest.mark.parametrize("nb", [4, 16])
def test_grouped_moments_equal_single_chains(hip_ctx, dtype, nb):
    from kubomsd_amd.backends.hip import moments_group
    ms = models()
    sc = scales(ms)
    engines = [hip_engine(hip_ctx, m) for m in ms]
    blocks = [to_dev(hip_ctx, phases(m.n, nb, 11 + q), dtype) for q, m in enumerate(ms)]
    for L in (1, 2, 3, 4, 17, 37, 300):
        want = [host(e.moments(B, L, c, a)) for e, B, (c, a) in zip(engines, blocks, sc)]
        h0 = [e.h_calls for e in engines]
        got = moments_group(engines, blocks, L, [c for c, _ in sc], [a for _, a in sc])
        for q, (g, w) in enumerate(zip(got, want)):
            np.testing.assert_array_equal(host(g), w, err_msg=f"chain {q}, L={L}")
            assert engines[q].h_calls - h0[q] == max(L - 1, 0)


def test_grouped_moments_match_oracle_and_reject_mixed_contexts(hip_ctx):
    from kubomsd_amd.backends.hip import HipContext, moments_group
    from kubomsd_amd.errors import KuboMSDError
    ms = models()[:4]
    sc = scales(ms)
    engines = [hip_engine(hip_ctx, m) for m in ms]
    Bh = [phases(m.n, 4, 3 + q) for q, m in enumerate(ms)]
    got = moments_group(engines, [to_dev(hip_ctx, B, "c128") for B in Bh], 64, [c for c, _ in sc],
                        [a for _, a in sc])
    for m, B, g, (c, a) in zip(ms, Bh, got, sc):
        assert rel_err(host(g), oracle_engine(m).moments(B, 64, c, a)) < 1e-11
    with pytest.raises(KuboMSDError, match="one width"):
        moments_group(engines[:2], [to_dev(hip_ctx, Bh[0], "c128"), to_dev(hip_ctx, Bh[1][:, :2], "c128")], 8,
                      [0, 0], [9, 9])
    with pytest.raises(KuboMSDError, match="distinct"):
        moments_group([engines[0], engines[0]], [to_dev(hip_ctx, Bh[0], "c128")] * 2, 8, [0, 0], [9, 9])



@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_stepwise_group_equals_alone_and_oracle(hip_ctx, precision):
    """Three stepwise requests of one size class (square lattices of
    different sizes and disorder) run with shared launches: each one's DOS
    and MSD numerators are bitwise what stepwise_samples gives it alone, and
    match the oracle's stepwise orchestration (fp64 within 1e-9, mixed
    within 3e-7)."""
    from kubomsd_amd.config import RunConfig
    from kubomsd_amd.contracts import ModelBundle
    from kubomsd_amd.operators.velocity import build_operators
    from kubomsd_amd.solvers.stepwise import stepwise_samples, stepwise_samples_group
    from oracle import kpm
    E = np.linspace(-4.5, 4.5, 21)
    cfg = RunConfig(energy_eV=E, time_fs=[0.0, 0.5, 1.0], solver="stochastic-kpm", kernel="jackson", backend="hip",
                    n_moments=64, axes=("x",), nsamples=4, sample_batch=4, include_vac=False, precision=precision,
                    metric_rtol=1e-13, segment_fs=1.0, spectral_bounds=(-5.6, 5.6))
    items, models = [], []
    for q, (L, W) in enumerate([(20, 2.0), (26, 1.0), (15, 3.0)]):
        m = square_lattice(L, W=W, seed=70 + q)
        ops = build_operators(ModelBundle(m.H, m.S, velocity=m.velocity), cfg, hip_ctx)
        items.append((ops, cfg, phases(m.n, 4, 80 + q)))
        models.append(m)
    grouped = stepwise_samples_group(items)
    tol = 1e-9 if precision == "fp64" else 3e-7
    kcfg = kpm.config(energy_eV=E, time_fs=cfg.time_fs, spectral_bounds=cfg.spectral_bounds, n_moments=64,
                      segment_fs=1.0)
    for (ops, _, B), m, (dos, nums, _) in zip(items, models, grouped):
        ados, anums, _ = stepwise_samples(ops, cfg, B)
        np.testing.assert_array_equal(host(dos), host(ados))
        np.testing.assert_array_equal(host(nums["msd_A2/commutator"]), host(anums["msd_A2/commutator"]))
        edos, enums = kpm.stepwise_samples(kpm.Ops(m.H, m.S, m.velocity, rtol=1e-13), kcfg, B)
        assert rel_err(host(dos), edos) <= tol
        assert rel_err(host(nums["msd_A2/commutator"]), enums["msd_A2/commutator"]) <= tol
