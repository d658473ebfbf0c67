"""Pin the oracle's multigrid components with closed forms and invariants.

* p-prolongation is exact on polynomials of the coarse degree; restriction is its transpose.
* the c-transfer prolongs a continuous Q1 function to its vertex values; restriction is the
  transpose (constrained nodes excluded).
* the continuous Q1 level equals a numpy dense assembly with Dirichlet rows as identity.
* the Chebyshev smoother's error propagation on every generalized eigenvector (A v = lam D v)
  equals the Chebyshev residual polynomial T_d((theta-lam)/delta)/T_d(theta/delta).
* the V-cycle with a Chebyshev coarse solver is a symmetric operator (needed by CG).
* MG-preconditioned CG reproduces the reference's golden sine errors at k = 6, 7.
"""
import json
import os

import numpy as np
import pytest
from numpy.polynomial import chebyshev as C
from scipy.linalg import eigh

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sine_errors.json")))


def _gl(n):
    from numpy.polynomial import legendre as L
    if n == 1:
        return np.array([0.5])
    return np.concatenate([[0.0], np.sort(0.5 * (L.Legendre.basis(n - 1).deriv().roots() + 1)), [1.0]])


def _dg_interp(mesh_cells, k, f, lower=0.0, upper=1.0):
    """nodal DG vector of f on a Cartesian box"""
    nx, ny, nz = mesh_cells
    h = (np.array([upper] * 3) - np.array([lower] * 3)) / np.array(mesh_cells)
    x = _gl(k + 1)
    out = []
    for cz in range(nz):
        for cy in range(ny):
            for cx in range(nx):
                for k3 in range(k + 1):
                    for j in range(k + 1):
                        for i in range(k + 1):
                            p = lower + h * (np.array([cx, cy, cz]) + np.array([x[i], x[j], x[k3]]))
                            out.append(f(p))
    return np.array(out)


def test_p_transfer_exact_and_transpose():
    cells = (2, 3, 2)
    m = oracle.make_mesh(cells)
    mg = oracle.Multigrid(m, [5, 2, 1], coarse_q1=True)
    f = lambda p: 1 + p[0] * p[1] - 2 * p[2] ** 2 + p[0] ** 2 * p[1] * p[2]  # degree <= 2 per dir
    coarse = _dg_interp(cells, 2, f)
    fine = mg.prolongate(3, coarse)  # level 3 = DG5, level 2 = DG2
    assert np.allclose(fine, _dg_interp(cells, 5, f), atol=1e-13)
    rng = np.random.default_rng(1)
    for lvl in (2, 3):
        xf = rng.standard_normal(mg.sizes[lvl])
        yc = rng.standard_normal(mg.sizes[lvl - 1])
        assert abs(mg.restrict(lvl, xf) @ yc - xf @ mg.prolongate(lvl, yc)) < 1e-11 * np.abs(xf).sum()


def test_c_transfer_trilinear_and_transpose():
    cells = (3, 2, 2)
    m = oracle.make_mesh(cells, boundary_type=(0, 1, 1, 1, 1, 1))  # Dirichlet only on x = 0
    mg = oracle.Multigrid(m, [1], coarse_q1=True)
    nn = [c + 1 for c in cells]
    h = 1.0 / np.array(cells)
    f = lambda p: p[0] * (1 + p[1]) * (2 - p[2])  # trilinear, zero on x = 0
    nodes = np.array([f(np.array([i, j, k]) * h) for k in range(nn[2]) for j in range(nn[1])
                      for i in range(nn[0])])
    dg = mg.prolongate(1, nodes)
    assert np.allclose(dg, _dg_interp(cells, 1, f), atol=1e-14)
    rng = np.random.default_rng(2)
    xf, yc = rng.standard_normal(mg.sizes[1]), rng.standard_normal(mg.sizes[0])
    yc_free = yc.copy()
    # constrained nodes (x = 0) are neither prolonged nor restricted into
    mask = np.array([i == 0 for k in range(nn[2]) for j in range(nn[1]) for i in range(nn[0])])
    yc_free[mask] = 0
    assert abs(mg.restrict(1, xf) @ yc_free - xf @ mg.prolongate(1, yc)) < 1e-12 * np.abs(xf).sum()
    assert np.all(mg.restrict(1, xf)[mask] == 0)


def test_q1_operator_dense_assembly():
    cells = (2, 3, 2)
    lo, hi = np.array([0.0, -1.0, 0.5]), np.array([1.0, 0.5, 2.0])
    m = oracle.make_mesh(cells, tuple(lo), tuple(hi), boundary_type=(0, 0, 1, 0, 1, 1))
    mg = oracle.Multigrid(m, [1], coarse_q1=True)
    h = (hi - lo) / np.array(cells)
    # 1D linear element matrices, exact
    K1 = lambda hh: np.array([[1, -1], [-1, 1]]) / hh
    M1 = lambda hh: np.array([[2, 1], [1, 2]]) * hh / 6
    Ke = np.kron(M1(h[2]), np.kron(M1(h[1]), K1(h[0]))) + np.kron(M1(h[2]), np.kron(K1(h[1]), M1(h[0]))) \
        + np.kron(K1(h[2]), np.kron(M1(h[1]), M1(h[0])))
    nn = [c + 1 for c in cells]
    N = np.prod(nn)
    A = np.zeros((N, N))
    for cz in range(cells[2]):
        for cy in range(cells[1]):
            for cx in range(cells[0]):
                ids = [(cx + (v & 1)) + nn[0] * ((cy + ((v >> 1) & 1)) + nn[1] * (cz + ((v >> 2) & 1)))
                       for v in range(8)]
                A[np.ix_(ids, ids)] += Ke
    idx = np.arange(N)
    ix, iy, iz = idx % nn[0], (idx // nn[0]) % nn[1], idx // (nn[0] * nn[1])
    cons = (ix == 0) | (ix == cells[0]) | (iy == cells[1])  # Dirichlet sides x-, x+, y+
    A[cons, :] = 0
    A[:, cons] = 0
    A[cons, cons] = 1
    x = np.random.default_rng(3).standard_normal(N)
    assert np.allclose(mg.level_vmult(0, x), A @ x, rtol=1e-13, atol=1e-13)


def test_chebyshev_error_polynomial():
    m = oracle.make_mesh((2, 1, 2), deformation=0.0)
    op = oracle.LaplaceOperator(m, 2)
    n = op.n
    A = np.column_stack([op.vmult(e) for e in np.eye(n)])
    D = np.diag(np.diag(A))
    lam, V = eigh(A, D)
    mg = oracle.Multigrid(m, [2, 1], coarse_q1=False, smoother_degree=5)
    for j in range(0, n, 7):
        v = V[:, j]
        x, theta, delta = mg.smoother_vmult(1, A @ v)
        e = v - x
        T5 = C.Chebyshev.basis(5)
        p = T5((theta - lam[j]) / delta) / T5(theta / delta)
        assert np.allclose(e, p * v, atol=1e-10 * np.abs(v).max())
    # Lanczos estimate lies inside the spectrum and near its top after 20 iterations
    assert 0.8 * lam.max() <= mg.lambda_max[1] <= lam.max() * (1 + 1e-12)


def test_vcycle_is_symmetric():
    m = oracle.make_mesh((3, 2, 2), deformation=0.1, mapping_degree=2)
    mg = oracle.Multigrid(m, [3, 1], coarse_q1=True, coarse_solver=1)
    rng = np.random.default_rng(4)
    x, y = rng.standard_normal(mg.sizes[-1]), rng.standard_normal(mg.sizes[-1])
    a, b = y @ mg.vmult(x), x @ mg.vmult(y)
    assert abs(a - b) <= 1e-10 * (abs(a) + abs(b))
    # positive definite on random vectors
    assert x @ mg.vmult(x) > 0


@pytest.mark.parametrize("mesh", ["cartesian", "curvilinear"])
@pytest.mark.parametrize("k", [6, 7])
def test_sine_golden_error_with_multigrid(mesh, k):
    m = oracle.make_mesh(8, -1.0, 1.0, deformation=0.15 if mesh == "curvilinear" else 0.0,
                         frequency=2, mapping_degree=3, boundary_type=(0, 1, 0, 0, 0, 0))
    op = oracle.LaplaceOperator(m, k)
    mg = oracle.Multigrid(m, oracle.bisect_degrees(k))
    x, info = oracle.cg_solve_mg(op, mg, op.rhs_sine(3 * np.pi))
    assert info["converged"] and info["iterations"] < 40
    assert f"{op.error_sine(3 * np.pi, x):.5e}" == f"{GOLD[mesh][str(k)]:.5e}"
