"""The no-allocation CG entry point (exadg_cg_solve_noalloc) against exadg_cg_solve.

Both run the same kernels in the same order; only the work vectors and the reduction buffers
differ (the library's static device workspace instead of the handle's own buffers). So on two
calls in a row (a larger curved mesh with the FP32 multigrid, then a smaller Cartesian one with
point Jacobi, which reuses the workspace the first call left behind) the iteration counts and
the solutions agree, within 1e-13 relative (the kernels are deterministic, so the difference
is expected to be zero). An operator above EXADG_NOALLOC_MAX_DOFS and a graph-replaying
multigrid are rejected with EXADG_ERR_ARGUMENT before any kernel runs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = 0


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _solve_pair(product, torch, op, prec, seed):
    b = product.random_vector(op.n, seed=seed)
    x_ref = torch.zeros_like(b)
    x_na = torch.zeros_like(b)
    info_ref = product.cg_solve(op, b, x_ref, prec=prec, rel_tol=1e-10)
    info_na = product.cg_solve(op, b, x_na, prec=prec, rel_tol=1e-10, noalloc=True)
    torch.cuda.synchronize()
    return info_ref, info_na, x_ref.cpu().numpy(), x_na.cpu().numpy()


def test_noalloc_cg_matches_cg_on_two_calls(torch_cuda, product):
    torch = torch_cuda
    # call 1: curved k=4 mesh, FP32 multigrid 4 -> 2 -> 1 -> Q1 (Chebyshev coarse solve)
    print("noalloc call 1: curved 6x5x4 k=4, FP32 MG", flush=True)
    mp = product.mesh_desc((6, 5, 4), 0.0, 1.0, deformation=0.1, frequency=1, mapping_degree=4)
    op1 = product.LaplaceOperator(mp, product.bc_desc((D,) * 6), 4)
    desc = product.mg_desc(4, degrees=[4, 2, 1], precision=product.FP32, coarse_q1=True,
                           coarse_solver=1, coarse_rel_tol=1e-3)
    prec1 = product.Preconditioner(op1, product.PRECOND_MULTIGRID, desc)
    i_ref, i_na, x_ref, x_na = _solve_pair(product, torch, op1, prec1, seed=11)
    err = _rel(x_na, x_ref)
    print(f"  iterations {i_ref['iterations']} / {i_na['iterations']}, rel.diff {err:.2e}, "
          f"r_true {i_na['r_true']:.3e} (tol {i_na['tolerance']:.3e})", flush=True)
    assert i_ref["converged"] and i_na["converged"]
    assert i_na["iterations"] == i_ref["iterations"]
    assert err <= 1e-13
    assert i_na["r_true"] <= 10 * i_na["tolerance"]

    # call 2: a smaller Cartesian k=2 mesh with point Jacobi; the workspace still holds call 1
    print("noalloc call 2: Cartesian 3x4x3 k=2, point Jacobi", flush=True)
    mp2 = product.mesh_desc((3, 4, 3), 0.0, 1.0)
    op2 = product.LaplaceOperator(mp2, product.bc_desc((D,) * 6), 2)
    assert op2.n < op1.n
    prec2 = product.Preconditioner(op2, product.PRECOND_POINT_JACOBI)
    i_ref, i_na, x_ref, x_na = _solve_pair(product, torch, op2, prec2, seed=12)
    err = _rel(x_na, x_ref)
    print(f"  iterations {i_ref['iterations']} / {i_na['iterations']}, rel.diff {err:.2e}",
          flush=True)
    assert i_ref["converged"] and i_na["converged"]
    assert i_na["iterations"] == i_ref["iterations"]
    assert err <= 1e-13


def test_noalloc_cg_rejects_what_it_cannot_hold(torch_cuda, product):
    torch = torch_cuda
    # above the maximum: 33^3 cells at k=4 = 4,492,125 DoF > 2^22
    print("noalloc: oversize operator", flush=True)
    op = product.LaplaceOperator(product.mesh_desc((33, 33, 33), 0.0, 1.0),
                                 product.bc_desc((D,) * 6), 4)
    assert op.n > product.NOALLOC_MAX_DOFS
    b = torch.ones(op.n, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    info = product.cg_solve(op, b, x, noalloc=True, raise_on_error=False)
    assert info["status"] == product.EXADG_ERR_ARGUMENT
    assert "EXADG_NOALLOC_MAX_DOFS" in product.lib().exadg_last_error().decode()
    del op, b, x
    # a multigrid that replays HIP graphs
    print("noalloc: graph-replaying multigrid", flush=True)
    op = product.LaplaceOperator(product.mesh_desc((3, 3, 3), 0.0, 1.0),
                                 product.bc_desc((D,) * 6), 3)
    desc = product.mg_desc(3, degrees=[3, 1], precision=product.FP32, coarse_q1=True,
                           coarse_solver=1, capture_graph=True)
    prec = product.Preconditioner(op, product.PRECOND_MULTIGRID, desc)
    b = product.random_vector(op.n, seed=3)
    x = torch.zeros_like(b)
    info = product.cg_solve(op, b, x, prec=prec, noalloc=True, raise_on_error=False)
    assert info["status"] == product.EXADG_ERR_ARGUMENT
