"""Pin the oracle's block-Jacobi pieces (row f2: block_jacobi_preconditioner.h,
OperatorBase::add_block_diagonal_matrices through the cell-based loop, operator_base.cpp:649-680,
and the Chebyshev smoother with PreconditionerSmoother::BlockJacobi).

* The cell blocks A_KK equal the diagonal blocks of the numpy dense assembly (Cartesian, curved,
  periodic including a cell that is its own periodic neighbour, mixed boundaries), and a
  plausible mistake (the symmetry term dropped) changes them.
* Block-Jacobi PCG solves the system (against a dense solve) in fewer iterations than point
  Jacobi.
* The block-Jacobi Chebyshev smoother's error propagation on every generalized eigenvector
  (A v = lam B v, B the block diagonal) is the Chebyshev residual polynomial, the V-cycle with it
  is symmetric positive definite, and MG-preconditioned CG reproduces the golden sine errors.
"""
import json
import os

import numpy as np
import pytest
from numpy.polynomial import chebyshev as C
from scipy.linalg import block_diag, eigh

import oracle
from tests import dense_sipg as ds

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sine_errors.json")))

CASES = [
    # cells, deformation, mapping degree, periodic, bc, degree
    ((2, 2, 1), 0.0, 1, (0, 0, 0), (0,) * 6, 2),
    ((2, 1, 2), 0.12, 2, (0, 0, 0), (0, 1, 0, 0, 1, 0), 2),
    ((3, 1, 1), 0.0, 1, (1, 1, 0), (0,) * 6, 1),
    ((1, 2, 2), 0.1, 3, (0, 1, 0), (0, 0, 0, 0, 1, 0), 3),
]


def _dense_blocks(A, m):
    nc = A.shape[0] // m
    return np.array([A[c * m:(c + 1) * m, c * m:(c + 1) * m] for c in range(nc)])


@pytest.mark.parametrize("case", CASES)
def test_cell_blocks_equal_dense_assembly(case):
    cells, dfm, md, per, bc, k = case
    op = oracle.LaplaceOperator(oracle.make_mesh(cells, deformation=dfm, mapping_degree=md,
                                                 periodic=per, boundary_type=bc), k)
    mesh = ds.Mesh(cells, 0.0, 1.0, dfm, 1, md, per, bc)
    A, _ = ds.assemble(mesh, k)
    m = (k + 1) ** 3
    B = op.block_diagonal()
    ref = _dense_blocks(A, m)
    # stored as [c, j, i] = (A_KK)_ij: symmetric, so compare directly and as a transpose
    assert np.abs(B - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(B - B.transpose(0, 2, 1)).max() <= 1e-12 * np.abs(ref).max()
    assert np.allclose(np.einsum("cii->ci", B).ravel(), op.diagonal(), rtol=0, atol=1e-13 * np.abs(ref).max())
    bad = _dense_blocks(ds.assemble(mesh, k, plant="drop_symmetry")[0], m)
    assert np.abs(B - bad).max() > 1e-3 * np.abs(ref).max()


def test_block_jacobi_pcg_solves_and_beats_point_jacobi():
    cells, k = (3, 2, 2), 2
    m = oracle.make_mesh(cells, deformation=0.1, mapping_degree=2, boundary_type=(0, 1, 0, 0, 0, 0))
    op = oracle.LaplaceOperator(m, k)
    A, _ = ds.assemble(ds.Mesh(cells, 0.0, 1.0, 0.1, 1, 2, (0, 0, 0), (0, 1, 0, 0, 0, 0)), k)
    b = oracle.random_vector(op.n)
    x, info = op.cg_solve_block_jacobi(b, rel_tol=1e-12)
    _, info_pj = op.cg_solve(b, rel_tol=1e-12)
    xs = np.linalg.solve(A, b)
    assert info["converged"] and np.linalg.norm(x - xs) <= 1e-9 * np.linalg.norm(xs)
    assert info["iterations"] < info_pj["iterations"]


def test_block_jacobi_chebyshev_error_polynomial():
    cells, k = (2, 2, 1), 1
    mo = oracle.make_mesh(cells, deformation=0.1, mapping_degree=2)
    A, _ = ds.assemble(ds.Mesh(cells, 0.0, 1.0, 0.1, 1, 2), k)
    op = oracle.LaplaceOperator(mo, k)
    B = block_diag(*op.block_diagonal())
    lam, V = eigh(A, B)
    mg = oracle.Multigrid(mo, [1], coarse_q1=True, smoother_degree=4, smoother_preconditioner=1)
    for j in range(0, A.shape[0], 5):
        v = V[:, j]
        x, theta, delta = mg.smoother_vmult(1, A @ v)
        T = C.Chebyshev.basis(4)
        p = T((theta - lam[j]) / delta) / T(theta / delta)
        assert np.allclose(v - x, p * v, atol=1e-10 * np.abs(v).max())


def test_block_jacobi_vcycle_symmetric_positive_definite():
    m = oracle.make_mesh((3, 2, 2), deformation=0.1, mapping_degree=2)
    mg = oracle.Multigrid(m, [3, 1], coarse_q1=True, coarse_solver=1, smoother_preconditioner=1)
    rng = np.random.default_rng(8)
    x, y = rng.standard_normal(mg.sizes[-1]), rng.standard_normal(mg.sizes[-1])
    a, b = y @ mg.vmult(x), x @ mg.vmult(y)
    assert abs(a - b) <= 1e-10 * (abs(a) + abs(b))
    assert x @ mg.vmult(x) > 0
    with pytest.raises(RuntimeError):  # block Jacobi on a continuous level (as the reference)
        oracle.Multigrid(m, [3, 1], mg_type="cpMG", smoother_preconditioner=1)


@pytest.mark.parametrize("mesh", ["cartesian", "curvilinear"])
def test_sine_golden_error_with_block_jacobi_smoother(mesh):
    k = 3
    m = oracle.make_mesh(8, -1.0, 1.0, deformation=0.15 if mesh == "curvilinear" else 0.0,
                         frequency=2, mapping_degree=3, boundary_type=(0, 1, 0, 0, 0, 0))
    op = oracle.LaplaceOperator(m, k)
    mg = oracle.Multigrid(m, oracle.bisect_degrees(k), smoother_preconditioner=1)
    x, info = oracle.cg_solve_mg(op, mg, op.rhs_sine(3 * np.pi))
    assert info["converged"] and info["iterations"] < 40
    assert f"{op.error_sine(3 * np.pi, x):.5e}" == f"{GOLD[mesh][str(k)]:.5e}"
