"""Pin the oracle's hybrid multigrid over h-, p- and c-levels (the MultigridType hierarchies of
multigrid_preconditioner_base.cpp:97-323) with the reference's own examples, a dense assembly
and closed forms.

* The level lists equal the examples documented in initialize_levels
  (multigrid_preconditioner_base.cpp:104-135: h-levels [0 1 2], p-levels [1 3 7] for p-, ph-,
  h- and hp-MG) and the sine application's cphMG hierarchy at k = 3 on three meshes
  (sine/application.h:167): (h0,Q1,CG) (h1,Q1,CG) (h2,Q1,CG) (h2,Q3,CG) (h2,Q3,DG).
* The continuous Q_p level (p = 1 and p > 1) equals a numpy dense assembly (independent node
  numbering and quadrature) with Dirichlet rows and columns as the identity.
* Every transfer (h on DG and continuous levels, p on continuous levels, c at p > 1) prolongs
  the interpolant of a polynomial in the coarse space to the interpolant on the fine level, and
  its restriction is the transpose of its prolongation.
* PCG preconditioned by the cphMG V-cycle reproduces the reference's golden sine errors.
"""
import itertools
import json
import os

import numpy as np
import pytest

import oracle
from tests import dense_sipg as ds

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sine_errors.json")))


def _coarsening(levels):
    """(mesh coarsenings, degree, dg) per level; 0 coarsenings = the finest mesh"""
    H = max(h for h, _, _ in levels) + 1
    return [(H - 1 - h, p, dg) for h, p, dg in levels]


@pytest.mark.parametrize("name, H, expected", [
    # p-MG: h = 2, 2, 2 (the finest mesh only), p = 1, 3, 7
    ("pMG", 3, [(0, 1, True), (0, 3, True), (0, 7, True)]),
    # ph-MG: (h, p) = (0,1) (1,1) (2,1) (2,3) (2,7)
    ("phMG", 3, [(2, 1, True), (1, 1, True), (0, 1, True), (0, 3, True), (0, 7, True)]),
    # h-MG: (0,7) (1,7) (2,7)
    ("hMG", 3, [(2, 7, True), (1, 7, True), (0, 7, True)]),
    # hp-MG: (0,1) (0,3) (0,7) (1,7) (2,7)
    ("hpMG", 3, [(2, 1, True), (2, 3, True), (2, 7, True), (1, 7, True), (0, 7, True)]),
])
def test_level_lists_equal_the_reference_examples(name, H, expected):
    got = _coarsening(oracle.mg_level_list(name, [7, 3, 1], n_h_levels=H))
    assert got == expected


def test_sine_cphmg_hierarchy():
    got = oracle.mg_level_list("cphMG", [3, 1], n_h_levels=3)
    assert got == [(0, 1, False), (1, 1, False), (2, 1, False), (2, 3, False), (2, 3, True)]


def test_every_level_pair_changes_one_thing():
    for t in oracle.MG_TYPES[1:]:
        L = oracle.mg_level_list(t, [4, 2, 1], n_h_levels=3)
        for a, b in zip(L, L[1:]):
            assert sum(x != y for x, y in zip(a, b)) == 1, (t, L)
        # finest level: the operator itself, on the finest mesh
        assert L[-1] == (max(h for h, _, _ in L), 4, True), t


CG_CASES = [
    # cells, deformation, mapping degree, periodic, bc, degree
    ((2, 3, 2), 0.0, 1, (0, 0, 0), (0,) * 6, 1),
    ((2, 2, 3), 0.0, 1, (0, 0, 0), (0, 1, 0, 0, 1, 0), 3),
    ((2, 2, 2), 0.1, 2, (0, 0, 0), (0, 1, 0, 0, 0, 0), 2),
    ((3, 2, 1), 0.0, 1, (1, 0, 1), (0,) * 6, 2),
]


@pytest.mark.parametrize("case", CG_CASES)
def test_continuous_level_equals_dense_assembly(case):
    cells, dfm, md, per, bc, p = case
    m = oracle.make_mesh(cells, deformation=dfm, mapping_degree=md, periodic=per, boundary_type=bc)
    # cMG on a degree-p operator: level 0 is the continuous Q_p level
    mg = oracle.Multigrid(m, [p], mg_type="cMG")
    assert mg.dg == [False, True] and mg.degrees[0] == p
    A = ds.assemble_continuous(ds.Mesh(cells, 0.0, 1.0, dfm, 1, md, per, bc), p)
    assert A.shape[0] == mg.sizes[0]
    x = np.random.default_rng(3).standard_normal(A.shape[0])
    y = mg.level_vmult(0, x)
    assert np.linalg.norm(y - A @ x) <= 1e-12 * np.linalg.norm(A @ x)


def test_continuous_assembly_detects_a_wrong_node_map():
    # the comparison above fails for a plausible mistake: a local node order with y fastest
    cells, per = (3, 2, 1), (1, 0, 1)
    m = oracle.make_mesh(cells, periodic=per)
    mg = oracle.Multigrid(m, [2], mg_type="cMG")
    mesh = ds.Mesh(cells, 0.0, 1.0, 0.0, 1, 1, per, (0,) * 6)
    A_bad = ds.assemble_continuous(mesh, 2, plant="swap_xy")
    A = ds.assemble_continuous(mesh, 2)
    x = np.random.default_rng(5).standard_normal(A.shape[0])
    y = mg.level_vmult(0, x)
    assert np.linalg.norm(y - A @ x) <= 1e-12 * np.linalg.norm(A @ x)
    assert np.linalg.norm(y - A_bad @ x) > 1e-3 * np.linalg.norm(A @ x)


def _node_positions(cells, p, dg, lo=0.0, hi=1.0):
    """positions of the DoFs of a DG_p or continuous Q_p vector on the undeformed box"""
    cells = np.array(cells)
    h = (hi - lo) / cells
    x = ds.gl_nodes(p + 1)
    if dg:
        pts = []
        for cz, cy, cx in itertools.product(range(cells[2]), range(cells[1]), range(cells[0])):
            for k, j, i in itertools.product(range(p + 1), repeat=3):
                pts.append(lo + h * (np.array([cx, cy, cz]) + np.array([x[i], x[j], x[k]])))
        return np.array(pts)
    nn = cells * p + 1
    pts = []
    for gz, gy, gx in itertools.product(range(nn[2]), range(nn[1]), range(nn[0])):
        g = np.array([gx, gy, gz])
        c = np.minimum(g // p, cells - 1)
        pts.append(lo + h * (c + x[g - c * p]))
    return np.array(pts)


HIERARCHIES = [("cphMG", [4, 2, 1]), ("phMG", [3, 1]), ("hMG", [2]), ("hcMG", [2]),
               ("chMG", [3]), ("hpcMG", [3, 1])]


@pytest.mark.parametrize("name, degrees", HIERARCHIES)
def test_transfers_exact_on_coarse_polynomials_and_transposed(name, degrees):
    cells = (4, 2, 4)
    # all-Neumann box: no constrained nodes, so polynomials are in every level's space
    m = oracle.make_mesh(cells, boundary_type=(1,) * 6)
    mg = oracle.Multigrid(m, degrees, mg_type=name, n_h_levels=2, coarse_solver=1, singular=True)
    rng = np.random.default_rng(7)
    H = mg.h_levels[-1] + 1
    for l in range(1, mg.n_levels):
        hc, pc, dgc = mg.h_levels[l - 1], mg.degrees[l - 1], mg.dg[l - 1]
        hf, pf, dgf = mg.h_levels[l], mg.degrees[l], mg.dg[l]
        cc = tuple(c >> (H - 1 - hc) for c in cells)
        cf = tuple(c >> (H - 1 - hf) for c in cells)
        q2 = 1.0 if pc >= 2 else 0.0  # degree <= pc in each variable
        f = lambda q: (1 + 0.5 * q[..., 0] + 0.25 * q2 * q[..., 0] ** 2) * \
            (0.3 - q[..., 1] + 0.2 * q2 * q[..., 1] ** 2) * (2 + q[..., 2]) + \
            q[..., 0] * q[..., 1] * q[..., 2]
        uc = f(_node_positions(cc, pc, dgc))
        uf = f(_node_positions(cf, pf, dgf))
        assert np.allclose(mg.prolongate(l, uc), uf, atol=1e-12), (name, l)
        x, y = rng.standard_normal(mg.sizes[l - 1]), rng.standard_normal(mg.sizes[l])
        a, b = y @ mg.prolongate(l, x), x @ mg.restrict(l, y)
        assert abs(a - b) <= 1e-12 * (abs(a) + abs(b)), (name, l)


@pytest.mark.parametrize("name", ["cphMG", "pchMG", "hMG"])
def test_transfers_with_dirichlet_constraints_are_transposed(name):
    m = oracle.make_mesh((4, 4, 2), deformation=0.1, mapping_degree=2, boundary_type=(0, 1, 0, 0, 0, 1))
    mg = oracle.Multigrid(m, [2, 1], mg_type=name, n_h_levels=2, coarse_solver=1)
    rng = np.random.default_rng(11)
    for l in range(1, mg.n_levels):
        x, y = rng.standard_normal(mg.sizes[l - 1]), rng.standard_normal(mg.sizes[l])
        a, b = y @ mg.prolongate(l, x), x @ mg.restrict(l, y)
        assert abs(a - b) <= 1e-12 * (abs(a) + abs(b)), (name, l)
    # the V-cycle stays symmetric positive definite (needed by CG)
    x, y = rng.standard_normal(mg.sizes[-1]), rng.standard_normal(mg.sizes[-1])
    a, b = y @ mg.vmult(x), x @ mg.vmult(y)
    assert abs(a - b) <= 1e-10 * (abs(a) + abs(b))
    assert x @ mg.vmult(x) > 0


@pytest.mark.parametrize("mesh", ["cartesian", "curvilinear"])
@pytest.mark.parametrize("k", [2, 3])
def test_sine_golden_error_with_cphmg(mesh, k):
    # the sine application: 2^3 coarse cells refined twice (8^3), cphMG, Bisect, CG coarse 1e-3
    m = oracle.make_mesh(8, -1.0, 1.0, deformation=0.15 if mesh == "curvilinear" else 0.0,
                         frequency=2, mapping_degree=3, boundary_type=(0, 1, 0, 0, 0, 0))
    op = oracle.LaplaceOperator(m, k)
    mg = oracle.Multigrid(m, oracle.bisect_degrees(k), mg_type="cphMG", n_h_levels=3)
    assert mg.h_levels[0] == 0 and mg.sizes[0] == 3 ** 3  # 2^3 cells: 3^3 Q1 nodes
    x, info = oracle.cg_solve_mg(op, mg, op.rhs_sine(3 * np.pi))
    assert info["converged"] and info["iterations"] < 40
    assert f"{op.error_sine(3 * np.pi, x):.5e}" == f"{GOLD[mesh][str(k)]:.5e}"
