"""Pins the oracle's statistics (oracle/statistics.py) before it checks the
device statistics: the reference's own known answers, exactness of the
second-order gradient on quadratics, and the brute-force bootstrap; a planted
mistake in the oracle must fail these checks."""
import numpy as np
import pytest

from oracle import statistics as S


def test_reference_known_answers():
    # tests/unit/test_statistics.py:4-9: ratio of sums with unequal counts
    value, se = S.estimate_ratio(np.array([[2.], [90.]]), np.array([[1.], [30.]]), [1, 3])
    np.testing.assert_allclose(value, [92 / 31])
    assert se is not None
    # :11-16: proportional numerator and denominator -> zero error
    dos = np.arange(1., 9.)[:, None]
    value, se = S.estimate_ratio(3 * dos, dos, np.ones(8, int))
    np.testing.assert_allclose(value, 3)
    np.testing.assert_allclose(se, 0, atol=1e-13)
    # :18-21: one block -> no error estimate
    _, se = S.estimate_ratio(np.ones((1, 2)), np.ones((1, 2)), [3])
    assert se is None


def test_bootstrap_equals_brute_force():
    rng = np.random.default_rng(3)
    A = rng.normal(size=(5, 2, 3, 4))
    D = rng.uniform(1, 2, size=(5, 4))
    _, se = S.estimate_ratio(A, D, np.full(5, 2))
    np.testing.assert_allclose(se, S.bootstrap_brute_force(A, D), rtol=1e-12)
    Dc = np.full((5, 1), 2.0)
    _, se = S.estimate_ratio(A, Dc, np.full(5, 2))
    np.testing.assert_allclose(se, S.bootstrap_brute_force(A, Dc), rtol=1e-12)


def test_resample_weights_are_block_counts():
    w = S.resample_weights(7)
    assert w.shape == (256, 7)
    np.testing.assert_array_equal(w.sum(1), 7)
    assert np.all(w >= 0) and np.all(w == np.round(w))


@pytest.mark.parametrize("t", [np.array([0, .1, .31, 1.]), np.linspace(0, 2, 9), np.array([0., .5])])
def test_gradient_exact_on_quadratics(t):
    # tests/unit/test_vac_integrals.py:4-9: d/dt (2 t^2) / 2 = 2 t, and any
    # quadratic is differentiated exactly by the second-order formulas
    # (a straight line by the first-order edges of a two-point grid).
    M = np.broadcast_to((2 * t * t)[None, :, None], (1, len(t), 2))
    if len(t) == 2:
        M = np.broadcast_to((3 * t + 1)[None, :, None], (1, 2, 2))
        np.testing.assert_allclose(S.msd_gradient(t, M)[0, :, 0], 1.5, atol=1e-14)
    else:
        np.testing.assert_allclose(S.msd_gradient(t, M)[0, :, 0], 2 * t, atol=1e-13)
        q = 1 - 3 * t + .7 * t * t
        g = S.msd_gradient(t, q[None, :, None])[0, :, 0]
        np.testing.assert_allclose(g, .5 * (-3 + 1.4 * t), atol=1e-13)
    assert S.msd_gradient([0.], M[:, :1]) is None


def test_block_sums_and_validity():
    x = np.arange(10.)[:, None] * np.ones((1, 3))
    got = S.block_sums(x, range(10), 4, 3)
    np.testing.assert_array_equal(got[:, 0], [0 + 1 + 2 + 3, 4 + 5 + 6 + 7, 8 + 9])
    rho = np.array([1., 1e-9, np.nan, 2., 3.])
    se = np.array([.1, .1, .1, 2., np.inf])
    valid, reason = S.dos_validity(rho, se, 1e-6, 9)
    np.testing.assert_array_equal(valid, [True, False, False, False, False])
    np.testing.assert_array_equal(reason, [0, 1 | 2, 1, 2, 2])  # 1e-9 - 1.96*0.1 <= 0: uncertain too
    valid, reason = S.dos_validity(rho, None, 1e-6, 3)
    np.testing.assert_array_equal(reason, np.where([True, False, False, True, True], 0, 1) | 32)


def test_planted_mistakes_fail():
    """A dropped resample weight or a flipped gradient sign is caught."""
    rng = np.random.default_rng(4)
    A = rng.normal(size=(4, 3))
    D = rng.uniform(1, 2, size=(4, 3))
    _, se = S.estimate_ratio(A, D, np.ones(4, int))
    w = S.resample_weights(4)
    w[:, 0] = 0
    rA, rD = w @ A, w @ D
    wrong = np.std(rA / rD, axis=0, ddof=1)
    assert not np.allclose(wrong, S.bootstrap_brute_force(A, D), rtol=1e-6)
    assert np.allclose(se, S.bootstrap_brute_force(A, D), rtol=1e-12)
    t = np.array([0, .1, .31, 1.])
    wrong_g = -.5 * np.gradient(2 * t * t, t, edge_order=2)
    assert not np.allclose(wrong_g, 2 * t)
