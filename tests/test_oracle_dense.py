"""Pin the oracle's SIPG operator and RHS against a numpy brute-force dense assembly.

Independent pieces: numpy Legendre machinery for nodes/bases, full tensor quadrature,
face-centric jump/average assembly. Also the algebraic invariants of the form: symmetry,
positive definiteness with Dirichlet boundaries, A 1 = 0 without Dirichlet boundaries.
"""
import numpy as np
import pytest

import oracle
from tests import dense_sipg

D, N = 0, 1
CASES = [
    # n_cells, lower, upper, deformation, frequency, mapping_degree, periodic, bc, degree
    ((2, 1, 2), 0.0, 1.0, 0.0, 1, 1, (0, 0, 0), (D,) * 6, 1),
    ((2, 2, 1), (0.0, -1.0, 0.5), (1.0, 0.5, 2.0), 0.0, 1, 1, (0, 0, 0), (D, N, D, N, D, D), 2),
    ((2, 2, 2), -1.0, 1.0, 0.15, 2, 3, (0, 0, 0), (D, N, D, D, D, D), 2),
    ((2, 1, 2), 0.0, 1.0, 0.1, 1, 2, (0, 0, 0), (D,) * 6, 3),
    ((2, 2, 1), 0.0, 1.0, 0.1, 1, 2, (1, 0, 1), (D,) * 6, 2),
    ((3, 1, 1), 0.0, 1.0, 0.0, 1, 1, (1, 1, 0), (D, D, D, D, N, N), 2),
    # curved, one cell along the periodic x: the cell is its own neighbour across that face
    ((1, 2, 2), 0.0, 1.0, 0.1, 1, 2, (1, 0, 0), (D,) * 6, 2),
]


def _pair(c):
    nc, lo, hi, dfm, fr, md, per, bc, k = c
    om = oracle.LaplaceOperator(oracle.make_mesh(nc, lo, hi, deformation=dfm, frequency=fr,
                                                 mapping_degree=md, periodic=per,
                                                 boundary_type=bc), k)
    dm = dense_sipg.Mesh(nc, lo, hi, deformation=dfm, frequency=fr, mapping_degree=md,
                         periodic=per, bc=bc)
    return om, dm, k


@pytest.mark.parametrize("case", CASES)
def test_oracle_vmult_equals_dense_assembly(case):
    om, dm, k = _pair(case)
    A, b = dense_sipg.assemble(dm, k, omega=2.5)
    assert A.shape[0] == om.n
    x = oracle.random_vector(om.n)
    y = om.vmult(x)
    assert np.linalg.norm(y - A @ x) <= 1e-12 * np.linalg.norm(A @ x)
    # symmetric bilinear form
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    # RHS (f, Dirichlet g, Neumann h) of u = prod sin(2.5 x_d)
    bo = om.rhs_sine(2.5)
    assert np.linalg.norm(bo - b) <= 1e-12 * np.linalg.norm(b)
    # diagonal
    assert np.allclose(om.diagonal(), np.diag(A), rtol=1e-12, atol=1e-12 * np.abs(A).max())
    has_dirichlet = any(case[7][s] == D and not case[6][s // 2] for s in range(6))
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    if has_dirichlet:
        assert ev.min() > 0  # SPD
    else:
        assert abs(ev.min()) <= 1e-10 * ev.max()
        assert np.abs(om.vmult(np.ones(om.n))).max() <= 1e-11 * np.abs(A).max()


@pytest.mark.parametrize("plant", ["drop_symmetry", "penalty_min"])
def test_dense_comparison_detects_planted_mistakes(plant):
    """Plausible mistakes planted in the dense assembly (the interior-face symmetry term
    -[u]{dv/dn} dropped; min instead of max of the cell penalties) are caught: the oracle
    matches the correct matrix to 1e-12 and misses the planted one by far more."""
    # cells of different penalty (boundary cells have larger tau) for "penalty_min"
    om, dm, k = _pair(((3, 1, 2), 0.0, 1.0, 0.0, 1, 1, (0, 0, 0), (D, N, D, D, N, D), 2))
    A, _ = dense_sipg.assemble(dm, k)
    Ap, _ = dense_sipg.assemble(dm, k, plant=plant)
    assert np.abs(Ap - A).max() > 1e-3 * np.abs(A).max()
    x = oracle.random_vector(om.n)
    y = om.vmult(x)
    assert np.linalg.norm(y - A @ x) <= 1e-12 * np.linalg.norm(y)
    assert np.linalg.norm(y - Ap @ x) > 1e-4 * np.linalg.norm(y)


def test_cartesian_penalty_closed_form():
    """tau_K = (3 + m/2)/h for m non-periodic boundary faces of a Cartesian cell."""
    h = 1.0 / 4
    om = oracle.LaplaceOperator(oracle.make_mesh(4), 2)
    tau = om.tau().reshape(4, 4, 4)
    for cz in range(4):
        for cy in range(4):
            for cx in range(4):
                m = sum((c == 0) + (c == 3) for c in (cx, cy, cz))
                assert abs(tau[cz, cy, cx] - (3 + m / 2) / h) < 1e-12
