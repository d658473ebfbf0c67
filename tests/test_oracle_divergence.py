"""Pin the oracle's weak DG divergence of a discrete velocity (the pressure-step right-hand side,
divergence_operator.cpp:168-298) against an independent numpy dense assembly, and pin that
assembly by closed forms:

* with every pressure side Dirichlet (velocity Neumann, flux u-.n), integration by parts is
  exact: (B u)_i = int (div u) q_i for a continuous polynomial velocity that the velocity space
  reproduces (QGauss(k_u+1) integrates the products exactly on affine cells);
* the pressure constant 1 = sum_i q_i sees only the boundary flux: 1^T B u = 0 for any u on a
  fully periodic box and on a box whose non-periodic sides are all walls.
"""
import numpy as np
import pytest

import oracle
from tests import dense_sipg

D, N = 0, 1
CASES = [
    # n_cells, deformation, mapping degree, periodic, pressure bc, k_p, k_u
    ((2, 1, 2), 0.0, 1, (0, 0, 0), (N,) * 6, 2, 3),
    ((2, 2, 1), 0.0, 1, (0, 0, 0), (D, N, D, N, N, D), 1, 2),
    ((2, 1, 2), 0.1, 2, (0, 0, 0), (N,) * 6, 2, 3),
    ((2, 2, 1), 0.08, 3, (1, 1, 0), (N,) * 6, 2, 2),
    ((1, 2, 2), 0.0, 1, (1, 0, 0), (D,) * 6, 3, 4),
]


def _meshes(c):
    nc, dfm, md, per, bc, kp, ku = c
    om = oracle.LaplaceOperator(oracle.make_mesh(nc, deformation=dfm, mapping_degree=md,
                                                 periodic=per, boundary_type=bc), kp)
    dm = dense_sipg.Mesh(nc, 0.0, 1.0, deformation=dfm, mapping_degree=md, periodic=per, bc=bc)
    return om, dm, kp, ku


@pytest.mark.parametrize("case", CASES)
def test_oracle_divergence_equals_dense_assembly(case):
    om, dm, kp, ku = _meshes(case)
    B = dense_sipg.assemble_divergence(dm, kp, ku)
    u = oracle.random_vector(B.shape[1], seed=17)
    b = om.divergence_apply(ku, u)
    assert np.linalg.norm(b - B @ u) <= 1e-12 * np.linalg.norm(B @ u)


def test_dense_divergence_integrates_by_parts_exactly():
    """All pressure sides Dirichlet: B u = int div(u) q for u = (x^2, y z, x) in DG_2^3."""
    kp, ku = 1, 2
    dm = dense_sipg.Mesh((2, 2, 1), 0.0, 1.0, bc=(D,) * 6)
    B = dense_sipg.assemble_divergence(dm, kp, ku)
    nu, npp = ku + 1, kp + 1
    gl_u = dense_sipg.gl_nodes(nu)
    u = []
    for cc in dm.cells:
        X = np.array([[gl_u[i], gl_u[j], gl_u[k]] for k in range(nu) for j in range(nu) for i in range(nu)])
        x, _ = dm.map(cc, X)
        u += [x[:, 0] ** 2, x[:, 1] * x[:, 2], x[:, 0]]
    u = np.concatenate(u)
    # int (div u) q_i = int (2 x + z) q_i with a Gauss rule exact for the product
    bp = dense_sipg.Basis1D(dense_sipg.gl_nodes(npp))
    gp, gw = dense_sipg.gauss(4)
    ref = []
    for cc in dm.cells:
        Q = np.array([[gp[i], gp[j], gp[k]] for k in range(4) for j in range(4) for i in range(4)])
        W = np.array([gw[i] * gw[j] * gw[k] for k in range(4) for j in range(4) for i in range(4)])
        x, J = dm.map(cc, Q)
        val, _ = dense_sipg._cell_basis(bp, Q)
        ref.append(val.T @ (W * np.linalg.det(J) * (2 * x[:, 0] + x[:, 2])))
    ref = np.concatenate(ref)
    assert np.linalg.norm(B @ u - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("periodic,bc", [((1, 1, 1), (N,) * 6), ((1, 0, 0), (N,) * 6),
                                         ((0, 0, 0), (N,) * 6)])
def test_constant_pressure_sees_no_flux_through_walls(periodic, bc):
    dm = dense_sipg.Mesh((2, 2, 2), 0.0, 1.0, deformation=0.1, mapping_degree=2,
                         periodic=periodic, bc=bc)
    B = dense_sipg.assemble_divergence(dm, 2, 3)
    u = oracle.random_vector(B.shape[1], seed=3)
    assert abs(np.ones(B.shape[0]) @ (B @ u)) <= 1e-12 * np.abs(B @ u).sum()
