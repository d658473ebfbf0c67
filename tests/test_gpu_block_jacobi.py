"""GPU parity of the block-Jacobi preconditioner and smoother (row f2) against the CPU oracle.

* exadg_preconditioner_vmult with EXADG_PRECOND_BLOCK_JACOBI equals the exact solve with the
  oracle's cell blocks A_KK (numpy) to 1e-12: on affine meshes through fast diagonalization
  (k = 1..8, periodic with a cell that is its own neighbour, mixed boundaries), on curved meshes
  through the dense inverses (k <= 4); ADDITIVE_SCHWARZ gives the same vector on the DG space;
  curved degree 5 is refused with EXADG_ERR_ARGUMENT.
* Block-Jacobi PCG reaches the oracle's solution with the oracle's iteration count (+-1).
* The multigrid V-cycle with the block-Jacobi Chebyshev smoother equals the oracle's to 1e-9 in
  FP64 (eigenvalue estimates to 1e-8) and to 1e-4 in FP32, and PCG with it reproduces the
  golden sine errors.
"""
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sine_errors.json")))
D, N = 0, 1

CASES = [
    # name, cells, deformation, mapping degree, periodic, bc, degree
    ("cart_k1", (3, 2, 2), 0.0, 1, (0, 0, 0), (D,) * 6, 1),
    ("cart_k2", (2, 3, 2), 0.0, 1, (0, 0, 0), (D, N, D, D, D, D), 2),
    ("cart_k3", (2, 2, 3), 0.0, 1, (0, 0, 0), (D,) * 6, 3),
    ("cart_k4", (2, 2, 2), 0.0, 1, (0, 0, 0), (D, D, N, D, D, N), 4),
    ("curved_k2", (3, 2, 2), 0.1, 2, (0, 0, 0), (D, N, D, D, D, D), 2),
    ("curved_k4", (2, 2, 2), 0.08, 4, (0, 0, 0), (D,) * 6, 4),
    ("periodic_self_k2", (1, 3, 2), 0.0, 1, (1, 0, 0), (D, D, D, N, D, D), 2),
    ("periodic_curved_k3", (2, 2, 2), 0.1, 3, (1, 1, 0), (D,) * 6, 3),
    ("cart_k5", (2, 2, 1), 0.0, 1, (0, 0, 0), (D, N, D, D, D, D), 5),
    ("cart_k6_periodic", (1, 2, 2), 0.0, 1, (1, 0, 0), (D,) * 6, 6),
    ("cart_k8", (2, 1, 1), 0.0, 1, (0, 0, 0), (D, D, D, D, N, D), 8),
]


def _ops(product, c):
    name, cells, dfm, md, per, bc, k = c
    mo = oracle.make_mesh(cells, deformation=dfm, mapping_degree=md, periodic=per, boundary_type=bc)
    mp = product.mesh_desc(cells, deformation=dfm, mapping_degree=md, periodic=per)
    return product.LaplaceOperator(mp, product.bc_desc(bc), k), oracle.LaplaceOperator(mo, k)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_block_jacobi_apply_matches_exact_block_solve(case, torch_cuda, product):
    torch = torch_cuda
    print(f"case {case[0]}", flush=True)
    op, op_o = _ops(product, case)
    B = op_o.block_diagonal()
    m = B.shape[1]
    r = oracle.random_vector(op.n)
    z_ref = np.concatenate([np.linalg.solve(B[c].T, r[c * m:(c + 1) * m]) for c in range(B.shape[0])])
    rs = torch.from_numpy(r).cuda()
    for kind in (product.PRECOND_BLOCK_JACOBI, product.PRECOND_ADDITIVE_SCHWARZ):
        prec = product.Preconditioner(op, kind)
        zs = torch.empty_like(rs)
        prec.vmult(zs, rs)
        torch.cuda.synchronize()
        err = np.linalg.norm(zs.cpu().numpy() - z_ref) / np.linalg.norm(z_ref)
        print(f"  kind {kind}: rel.err {err:.2e}", flush=True)
        assert err <= 1e-12


def test_block_jacobi_refuses_curved_degree_5(torch_cuda, product):
    op = product.LaplaceOperator(product.mesh_desc((2, 2, 2), deformation=0.1, mapping_degree=2),
                                 product.bc_desc((D,) * 6), 5)
    with pytest.raises(product.ExaDGError) as e:
        product.Preconditioner(op, product.PRECOND_BLOCK_JACOBI)
    assert e.value.code == product.EXADG_ERR_ARGUMENT


def test_block_jacobi_pcg_matches_oracle(torch_cuda, product):
    torch = torch_cuda
    case = ("curved_k3", (3, 3, 2), 0.1, 3, (0, 0, 0), (D, N, D, D, D, D), 3)
    op, op_o = _ops(product, case)
    b = oracle.random_vector(op.n)
    x_o, info_o = op_o.cg_solve_block_jacobi(b, rel_tol=1e-10)
    prec = product.Preconditioner(op, product.PRECOND_BLOCK_JACOBI)
    bs = torch.from_numpy(b).cuda()
    xs = torch.zeros_like(bs)
    info = product.cg_solve(op, bs, xs, prec=prec, rel_tol=1e-10)
    err = np.linalg.norm(xs.cpu().numpy() - x_o) / np.linalg.norm(x_o)
    print(f"block-Jacobi PCG: {info['iterations']} vs oracle {info_o['iterations']} iterations, "
          f"rel.diff {err:.2e}", flush=True)
    assert info["converged"] and abs(info["iterations"] - info_o["iterations"]) <= 1
    assert err <= 1e-8


MG_CASES = [
    ("curved_k3", (3, 3, 2), 0.1, 3, (0, 0, 0), (D, N, D, D, D, D), [3, 1]),
    ("cart_k4", (3, 2, 2), 0.0, 1, (0, 0, 0), (D,) * 6, [4, 2, 1]),
]


@pytest.mark.parametrize("precision", ["FP64", "FP32"])
@pytest.mark.parametrize("smoother", [1, 2], ids=["block_jacobi", "additive_schwarz"])
@pytest.mark.parametrize("case", MG_CASES, ids=[c[0] for c in MG_CASES])
def test_block_jacobi_vcycle_matches_oracle(case, smoother, precision, torch_cuda, product):
    torch = torch_cuda
    name, cells, dfm, md, per, bc, degs = case
    print(f"MG {name} smoother {smoother} {precision}", flush=True)
    mo = oracle.make_mesh(cells, deformation=dfm, mapping_degree=md, periodic=per, boundary_type=bc)
    mp = product.mesh_desc(cells, deformation=dfm, mapping_degree=md, periodic=per)
    op = product.LaplaceOperator(mp, product.bc_desc(bc), degs[0])
    desc = product.mg_desc(degs[0], degrees=degs, precision=getattr(product, precision),
                           coarse_solver=1, smoother_preconditioner=smoother)
    prec = product.Preconditioner(op, product.PRECOND_MULTIGRID, desc)
    mg_o = oracle.Multigrid(mo, degs, coarse_solver=1, smoother_preconditioner=smoother)
    info = prec.mg_info()
    if precision == "FP64":
        for l in range(1, info["n_levels"]):
            assert abs(info["lambda_max"][l] - mg_o.lambda_max[l]) <= 1e-8 * mg_o.lambda_max[l], l
    r = oracle.random_vector(op.n)
    z_ref = mg_o.vmult(r)
    rs = torch.from_numpy(r).cuda()
    zs = torch.empty_like(rs)
    prec.vmult(zs, rs)
    torch.cuda.synchronize()
    err = np.linalg.norm(zs.cpu().numpy() - z_ref) / np.linalg.norm(z_ref)
    print(f"  V-cycle rel.err {err:.2e}", flush=True)
    assert err <= (1e-9 if precision == "FP64" else 1e-4)


@pytest.mark.parametrize("mesh", ["cartesian", "curvilinear"])
def test_sine_golden_error_with_block_jacobi_smoother_on_gpu(mesh, torch_cuda, product):
    torch = torch_cuda
    k = 3
    print(f"sine {mesh} k={k} block-Jacobi smoother", flush=True)
    mp = product.mesh_desc(8, -1.0, 1.0, deformation=0.15 if mesh == "curvilinear" else 0.0,
                           frequency=2, mapping_degree=3)
    op = product.LaplaceOperator(mp, product.bc_desc((D, N, D, D, D, D)), k)
    desc = product.mg_desc(k, smoother_preconditioner=product.SMOOTHER_BLOCK_JACOBI)
    prec = product.Preconditioner(op, product.PRECOND_MULTIGRID, desc)
    b = torch.empty(op.n, dtype=torch.float64, device="cuda")
    op.rhs_sine(3 * np.pi, b)
    x = torch.zeros_like(b)
    info = product.cg_solve(op, b, x, prec=prec, rel_tol=1e-10)
    err = op.error_sine(3 * np.pi, x)
    print(f"  {info['iterations']} iterations, error {err:.5e}", flush=True)
    assert info["converged"] and info["iterations"] < 40
    assert f"{err:.5e}" == f"{GOLD[mesh][str(k)]:.5e}"
