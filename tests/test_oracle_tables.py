"""1D tables of the oracle against numpy and the closed forms they must satisfy."""
import numpy as np
import pytest
from numpy.polynomial import legendre as L

import oracle


@pytest.mark.parametrize("k", range(1, 9))
def test_tables(k):
    t = oracle.tables(k)
    n = k + 1
    x, w = L.leggauss(n)
    assert np.allclose(t["gp"], 0.5 * (x + 1), atol=1e-15)
    assert np.allclose(t["gw"], 0.5 * w, atol=1e-15)
    inner = np.sort(0.5 * (L.Legendre.basis(k).deriv().roots() + 1)) if k > 1 else np.array([])
    assert np.allclose(t["gl"], np.concatenate([[0.0], inner, [1.0]]), atol=1e-14)
    # partition of unity / derivative rows sum to zero
    assert np.allclose(t["S"].sum(axis=1), 1.0, atol=1e-14)
    assert np.allclose(t["Sd"].sum(axis=1), 0.0, atol=1e-12)
    assert abs(t["d0"].sum()) < 1e-11 and abs(t["d1"].sum()) < 1e-11
    # exactness on x^j, j <= k: interpolation of nodal values reproduces the polynomial
    for j in range(k + 1):
        f = t["gl"] ** j
        assert np.allclose(t["S"] @ f, t["gp"] ** j, atol=1e-13)
        df = j * t["gp"] ** (j - 1) if j > 0 else 0 * t["gp"]
        assert np.allclose(t["Sd"] @ f, df, atol=1e-11)
        assert abs(t["d1"] @ f - j) < 1e-10
        assert abs(t["d0"] @ f - (1.0 if j == 1 else 0.0)) < 1e-10


def test_random_vector_is_counter_based():
    v = oracle.random_vector(1000)
    assert np.all(np.abs(v) <= 1.0) and abs(v.mean()) < 0.1
    assert np.array_equal(oracle.random_vector(10), v[:10])
    assert not np.array_equal(oracle.random_vector(10, seed=1), v[:10])


def test_random_vector_matches_published_splitmix64():
    """The counter-based generator is SplitMix64: entry i for seed s is the (i+1)-th output of
    the sequential generator started from state s. Published outputs for state 1234567
    (Vigna's reference splitmix64.c; the first five values) pin the constants and shifts."""
    published = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                 4593380528125082431, 16408922859458223821]
    expected = np.array([2.0 * ((z >> 11) * 2.0 ** -53) - 1.0 for z in published])
    assert np.array_equal(oracle.random_vector(5, seed=1234567), expected)
