"""Pin the oracle end-to-end on the reference's own golden outputs.

applications/poisson/sine/tests/{cartesian,curvilinear}.output hold the relative L2 error
(QGauss(k+3)) after the CG solve to 1e-10 for k = 1..7 on 8^3 cells. Matching all printed
digits pins geometry (MappingQ(3) of the DeformedCubeManifold), RHS, weak BCs, penalty and
the operator together.
"""
import json
import os

import numpy as np
import pytest

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sine_errors.json")))
FAST = [1, 2, 3, 4, 5]


def sine_problem(k, curvilinear):
    m = oracle.make_mesh(8, -1.0, 1.0, deformation=0.15 if curvilinear else 0.0, frequency=2,
                         mapping_degree=3, boundary_type=(0, 1, 0, 0, 0, 0))
    return oracle.LaplaceOperator(m, k)


@pytest.mark.parametrize("mesh", ["cartesian", "curvilinear"])
@pytest.mark.parametrize("k", FAST)
def test_sine_golden_error(mesh, k):
    op = sine_problem(k, mesh == "curvilinear")
    assert op.n == GOLD["n_dofs"][str(k)]
    b = op.rhs_sine(3 * np.pi)
    x, info = op.cg_solve(b, rel_tol=1e-10)
    assert info["converged"]
    err = op.error_sine(3 * np.pi, x)
    ref = GOLD[mesh][str(k)]
    assert f"{err:.5e}" == f"{ref:.5e}", (err, ref)


def test_golden_detects_wrong_mapping_degree():
    """A plausible mistake (mapping degree 1 instead of 3 on the curved mesh) fails the pin."""
    m = oracle.make_mesh(8, -1.0, 1.0, deformation=0.15, frequency=2, mapping_degree=1,
                         boundary_type=(0, 1, 0, 0, 0, 0))
    op = oracle.LaplaceOperator(m, 2)
    x, _ = op.cg_solve(op.rhs_sine(3 * np.pi))
    assert f"{op.error_sine(3 * np.pi, x):.5e}" != f"{GOLD['curvilinear']['2']:.5e}"
