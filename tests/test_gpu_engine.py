"""HIP engine vs the FP64 oracle on identical inputs (orthogonal lattices).

Tolerances: complex128 blocks within 1e-11 relative to the largest entry
(reduction order differs from the oracle's serial row order; the reference's
own cpp-vs-reference bound is 2e-9..2e-11, tests/unit/test_native.py:128-143);
complex64 storage within 5e-7 (FP32 rounding of the input block and of every
stored iterate; measured 3e-8 .. 7e-8 for products, chains and moments, and
3e-7 for the 14-term augmented series, whose tolerance is ten times larger).
"""
import numpy as np
import pytest

from helpers import hip_engine, host, oracle_engine, phases, random_block, rel_err, to_dev
from kubomsd_amd.synthetic import honeycomb_lattice, square_lattice

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-11, "c64": 5e-7}


@pytest.fixture(scope="module")
def square():
    m = square_lattice(24, seed=11)
    return m, oracle_engine(m)


@pytest.fixture(scope="module")
def honey():
    m = honeycomb_lattice(12, 10, seed=12)
    return m, oracle_engine(m)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("nb", [1, 3, 16, 64, 80])
def test_matmul_matches_scipy(hip_ctx, square, dtype, nb):
    m, _ = square
    from kubomsd_amd.backends.hip import HipMatrix
    B = random_block(m.n, nb, 1)
    H = HipMatrix(hip_ctx, m.H)
    got = host(H.matmul(to_dev(hip_ctx, B, dtype)))
    want = m.H @ (B.astype(np.complex64) if dtype == "c64" else B)
    assert rel_err(got, want) < TOL[dtype]


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("nb", [2, 16, 33])
def test_moments_match_oracle(hip_ctx, square, dtype, nb):
    m, oe = square
    B = phases(m.n, nb, 2)
    c, a = 0.0, 5.1
    want = oe.moments(B, 40, c, a)
    he = hip_engine(hip_ctx, m)
    got = host(he.moments(to_dev(hip_ctx, B, dtype), 40, c, a))
    assert got.shape == want.shape
    assert rel_err(got, want) < TOL[dtype]
    assert he.h_calls == 39


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_moments_honeycomb_long_chain(hip_ctx, honey, dtype):
    m, oe = honey
    B = phases(m.n, 8, 3)
    c, a = 0.05, 8.8
    want = oe.moments(B, 300, c, a)
    he = hip_engine(hip_ctx, m)
    for chunk in (16, 7):
        hip_ctx.set_tuning("moment_chunk", chunk)
        got = host(he.moments(to_dev(hip_ctx, B, dtype), 300, c, a))
        assert rel_err(got, want) < (1e-10 if dtype == "c128" else TOL["c64"])
    hip_ctx.set_tuning("moment_chunk", 16)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_chain_matches_oracle(hip_ctx, square, dtype):
    m, oe = square
    B = random_block(m.n, 5, 4)
    want = oe.chain(B, 12, 0.2, 5.1)
    got = host(hip_engine(hip_ctx, m).chain(to_dev(hip_ctx, B, dtype), 12, 0.2, 5.1))
    assert rel_err(got, want) < TOL[dtype] * 10


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_h_action_and_v_cov(hip_ctx, honey, dtype):
    m, oe = honey
    B = random_block(m.n, 6, 5)
    he = hip_engine(hip_ctx, m)
    assert rel_err(host(he.h_action(to_dev(hip_ctx, B, dtype))), oe.h_action(B)) < TOL[dtype]
    for d in ([1.0, 0, 0], [0, 1.0, 0], [0.6, 0.8, 0]):
        got = host(he.v_cov(to_dev(hip_ctx, B, dtype), d))
        assert rel_err(got, oe.v_cov(B, d)) < TOL[dtype]
    assert he.v_calls == 3


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("ndirs", [0, 1, 2])
def test_augmented_series(hip_ctx, honey, dtype, ndirs):
    m, oe = honey
    nb = 4
    B0 = phases(m.n, nb, 6)
    B = np.hstack([B0] + [random_block(m.n, nb, 7 + d) * 0.1 for d in range(ndirs)])
    dirs = [[1.0, 0, 0], [0, 1.0, 0]][:ndirs]
    rng = np.random.default_rng(9)
    coeff = rng.normal(size=14) + 1j * rng.normal(size=14)
    coeff /= np.arange(1, 15) ** 2
    want = oe.series(B, coeff, 0.05, 8.8, dirs)
    got = host(hip_engine(hip_ctx, m).series(to_dev(hip_ctx, B, dtype), coeff, 0.05, 8.8, dirs))
    assert rel_err(got, want) < TOL[dtype] * 10


def test_nonfinite_input_rejected(hip_ctx, square):
    import torch
    from kubomsd_amd.errors import NumericalError
    m, _ = square
    B = phases(m.n, 2, 1)
    B[3, 1] = np.nan
    with pytest.raises(NumericalError, match="finite"):
        hip_engine(hip_ctx, m).moments(to_dev(hip_ctx, B), 4, 0.0, 5.1)


def test_budget_rejection(hip_ctx, square):
    from kubomsd_amd.backends.hip import HipContext, HipMatrix
    from kubomsd_amd.errors import ResourceBlocked
    small = HipContext(0, hbm_budget_bytes=4096)
    m, _ = square
    with pytest.raises(ResourceBlocked, match="RESOURCE_BLOCKED"):
        HipMatrix(small, m.H)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("L", [1, 2, 3, 64, 301])
def test_moment_doubling_matches_direct_oracle(hip_ctx, honey, dtype, L):
    """mu_2k = 2<T_k,T_k> - mu_0, mu_2k+1 = 2<T_k,T_k+1> - mu_1 equals the
    reference's direct recursion (exact identity; rounding differs)."""
    m, oe = honey
    B = phases(m.n, 8, 13)
    want = oe.moments(B, L, 0.05, 8.8)
    got = host(hip_engine(hip_ctx, m).moments(to_dev(hip_ctx, B, dtype), L, 0.05, 8.8, doubling=True))
    assert got.shape == want.shape
    assert rel_err(got, want) < TOL[dtype]


@pytest.mark.parametrize("nb", [1, 2, 3, 5])
@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_chain_v_matches_chain_then_velocity(hip_ctx, nb, dtype):
    """kmsd_chain_v (on-device regrouping around one velocity product) equals
    chain + per-order v_cov, bit for bit on pixels of 1..5 complex entries
    (tiled and direct transposes)."""
    from helpers import hip_engine
    from kubomsd_amd.synthetic import square_lattice
    m = square_lattice(12, t=-1.0, W=2.0)
    eng = hip_engine(hip_ctx, m)
    B = to_dev(hip_ctx, random_block(m.n, nb, 21), dtype)
    ch = eng.chain(B, 7, 0.0, 5.1)
    want = np.stack([host(eng.v_cov(ch[k].contiguous(), [1.0, 0, 0])) for k in range(7)])
    got = host(eng.chain_v(B, 7, 0.0, 5.1, [1.0, 0, 0]))
    assert got.shape == (7, m.n, nb)
    assert rel_err(got, want) < (1e-14 if dtype == "c128" else 1e-6)
