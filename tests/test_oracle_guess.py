"""CPU checks of the learned-guess oracle and the pressure right-hand side oracle.

* extrapolation constants: the reference's golden sums
  (tests/time_integration/extrapolation_constants_01.output) and exactness on polynomials of
  degree < order, which fixes every constant of extrapolation_constants.cpp:62-83;
* the numpy MLP backward pass against central finite differences;
* the weak DG divergence (divergence_operator.cpp:168-298 with central flux, walls u+ = -u-)
  against closed-form integrals of a single non-solenoidal mode.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from oracle import guess as og

HERE = os.path.dirname(__file__)


def test_extrapolation_golden_sums():
    gold = json.load(open(os.path.join(HERE, "golden", "extrapolation_sums.json")))["sums"]
    for order, sums in gold.items():
        for cur, s in zip(range(1, int(order) + 1), sums):
            assert sum(og.BETA[cur]) == pytest.approx(s, abs=1e-15)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_extrapolation_polynomial_exactness(order):
    # p(t) = t^j sampled at t = -1, -2, ..., -order extrapolates exactly to t = 0 for j < order
    for j in range(order):
        hist = [float((-(i + 1)) ** j) for i in range(order)]
        assert og.extrapolate(hist) == pytest.approx(1.0 if j == 0 else 0.0, abs=1e-13)
    # and not for j = order (the leading error term is nonzero)
    hist = [float((-(i + 1)) ** order) for i in range(order)]
    assert abs(og.extrapolate(hist)) > 0.5


def test_mlp_gradient_finite_differences():
    rng = np.random.default_rng(3)
    F, w, o, B = 12, 7, 5, 9
    n_par = w * F + w + w * w + w + o * w + o
    p = rng.uniform(-0.6, 0.6, n_par)
    X = rng.normal(size=(B, F))
    T = rng.normal(size=(B, o))
    loss, g = og.loss_and_grad(p, X, T, w, o)
    eps = 1e-6
    fd = np.empty(n_par)
    for i in range(n_par):
        e = np.zeros(n_par)
        e[i] = eps
        fd[i] = (og.loss_and_grad(p + e, X, T, w, o)[0] - og.loss_and_grad(p - e, X, T, w, o)[0]) / (2 * eps)
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.max(np.abs(g)))
    # the loss is the mean square of forward() - T
    Y, _ = og.forward(p, X, w, o)
    assert loss == pytest.approx(((Y - T) ** 2).mean(), rel=1e-14)


def test_features_are_differences():
    rng = np.random.default_rng(5)
    n3 = 8
    h = [rng.normal(size=4 * n3) for _ in range(3)]
    rhs = rng.normal(size=4 * n3)
    tgt = rng.normal(size=4 * n3)
    sc = [2.0, 0.5, 0.25, 4.0, 0.1]
    X, T = og.features(h, rhs, n3, sc, tgt)
    H = [v.reshape(-1, n3) for v in h]
    np.testing.assert_allclose(X[:, :n3] * 2.0, H[0], rtol=1e-15)
    np.testing.assert_allclose(X[:, n3:2 * n3] * 0.5, H[0] - H[1], rtol=1e-14)
    np.testing.assert_allclose(X[:, 2 * n3:3 * n3] * 0.25, H[0] - 2 * H[1] + H[2], rtol=1e-13)
    np.testing.assert_allclose(X[:, 3 * n3:] * 4.0, rhs.reshape(-1, n3), rtol=1e-15)
    np.testing.assert_allclose(T * 0.1, tgt.reshape(-1, n3) - (3 * H[0] - 3 * H[1] + H[2]),
                               rtol=1e-12, atol=1e-14)


def _single_mode(theta):
    # u = cos(2 pi x + theta) e_x (not solenoidal: div u = -2 pi sin(2 pi x + theta))
    return np.array([[1, 0, 0, 1, 0, 0, 1.0, theta, 0.0, 0.0]])


def _x2(op, k):
    """Nodal DG coefficients of v(x) = x^2 (exact for k >= 2 on GL nodes)."""
    tb = oracle.tables(k)
    gl = tb["gl"]
    nc = round((op.n // (k + 1) ** 3) ** (1 / 3))
    n = k + 1
    v = np.empty((nc, nc, nc, n, n, n))
    h = 1.0 / nc
    for cx in range(nc):
        v[:, :, cx, :, :, :] = ((cx + gl) * h)[None, None, :] ** 2
    return v.reshape(-1)


@pytest.mark.parametrize("k", [2, 4])
def test_divergence_rhs_closed_form_walls(k):
    # walls in x (Dirichlet pressure sides => no mean removal): b.v = -int u . grad(x^2)
    # = -int_0^1 2x cos(2 pi x + theta) dx = -sin(theta)/pi; any dropped sign or term breaks it
    theta = 0.7
    md = oracle.make_mesh(4, periodic=(0, 1, 1), boundary_type=(0,) * 6)
    op = oracle.LaplaceOperator(md, k)
    b = op.rhs_divergence_fourier(_single_mode(theta), 0.0)
    val = float(b @ _x2(op, k))
    # quadrature error of QGauss(k+1) on 4 cells: ~1e-5 at k = 2, ~1e-8 at k = 4
    assert val == pytest.approx(-math.sin(theta) / math.pi, rel=1e-4 if k == 2 else 1e-6)


@pytest.mark.parametrize("k", [2, 4])
def test_divergence_rhs_closed_form_periodic_face(k):
    # periodic in x: v = x^2 jumps from 1 to 0 across the periodic face, which adds
    # u_x(1) * 1 = cos(theta) to b.v (central flux of a continuous field)
    theta = 0.7
    md = oracle.make_mesh(4, periodic=(1, 1, 0), boundary_type=(0,) * 6)
    op = oracle.LaplaceOperator(md, k)
    b = op.rhs_divergence_fourier(_single_mode(theta), 0.0)
    val = float(b @ _x2(op, k))
    assert val == pytest.approx(-math.sin(theta) / math.pi + math.cos(theta),
                               rel=1e-4 if k == 2 else 1e-6)


def test_divergence_rhs_compatibility_and_time_phase():
    import exadg_amd  # only the host-side mode generator (no GPU call)
    modes = exadg_amd.fourier_modes(11, 1, 3)
    md = oracle.make_mesh(4, periodic=(1, 1, 0), boundary_type=(1,) * 6)
    op = oracle.LaplaceOperator(md, 3)
    b = op.rhs_divergence_fourier(modes, 0.3)
    assert abs(b.sum()) < 1e-13 * np.abs(b).sum()
    # omega t enters as a phase shift: t -> t + dt equals phi -> phi + omega dt
    m2 = modes.copy()
    m2[:, 7] += m2[:, 8] * 0.05
    np.testing.assert_allclose(op.rhs_divergence_fourier(m2, 0.3), op.rhs_divergence_fourier(modes, 0.35),
                               atol=1e-14 * np.abs(b).max())
