"""GPU parity of the solver layer (C ABI) against the CPU oracle.

* The FP64 V-cycle equals the oracle's V-cycle to 1e-9 when the coarse solve is a fixed
  polynomial (Chebyshev) or is converged tightly. Per-level Chebyshev eigenvalue estimates
  agree to 1e-8.
* The FP32 V-cycle agrees with the FP64 oracle to 1e-4 relative. Its only effect on the outer
  FP64 CG is the iteration count.
* PCG (none / point Jacobi / FP64 MG / FP32 MG) reaches rel. residual 1e-10 and reproduces the
  reference's golden sine errors (6 digits). Jacobi iteration counts match the oracle's to +-1.
* Error paths: NaN right-hand side -> EXADG_ERR_NOT_FINITE; iteration cap -> EXADG_ERR_NO_CONVERGE.
"""
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sine_errors.json")))
D, N = 0, 1


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


MG_CASES = [
    # name, n_cells, lower, upper, deformation, freq, mapping_degree, bc, degrees, coarse_q1,
    # coarse solver (0 CG, 1 Chebyshev), coarse tol
    ("cart_k3", (4, 3, 3), 0.0, 1.0, 0.0, 1, 1, (D,) * 6, [3, 1], True, 1, 1e-3),
    ("curved_k4", (3, 3, 3), 0.0, 1.0, 0.1, 1, 3, (D, N, D, D, D, D), [4, 2, 1], True, 1, 1e-3),
    ("cart_k5_cg", (3, 2, 3), -1.0, 1.0, 0.0, 1, 1, (D, N, D, D, D, D), [5, 3, 2, 1], True, 0, 1e-12),
    ("curved_k2_dg1coarse", (4, 2, 2), 0.0, 1.0, 0.08, 1, 2, (D,) * 6, [2, 1], False, 1, 1e-3),
]


def _mg_pair(product, c, precision):
    name, nc, lo, hi, dfm, fr, md, bc, degs, q1, cs, ctol = c
    k = degs[0]
    mo = oracle.make_mesh(nc, lo, hi, deformation=dfm, frequency=fr, mapping_degree=md,
                          boundary_type=bc)
    mp = product.mesh_desc(nc, lo, hi, deformation=dfm, frequency=fr, mapping_degree=md)
    op = product.LaplaceOperator(mp, product.bc_desc(bc), k)
    desc = product.mg_desc(k, degrees=degs, precision=precision, coarse_q1=q1, coarse_solver=cs,
                           coarse_rel_tol=ctol)
    prec = product.Preconditioner(op, product.PRECOND_MULTIGRID, desc)
    mg_o = oracle.Multigrid(mo, degs, coarse_q1=q1, coarse_solver=cs, coarse_rel_tol=ctol)
    return op, prec, mg_o


@pytest.mark.parametrize("case", MG_CASES, ids=[c[0] for c in MG_CASES])
def test_vcycle_fp64_matches_oracle(case, torch_cuda, product):
    torch = torch_cuda
    op, prec, mg_o = _mg_pair(product, case, product.FP64)
    info = prec.mg_info()
    assert info["n_levels"] == mg_o.n_levels
    for l in range(1, info["n_levels"]):
        assert abs(info["lambda_max"][l] - mg_o.lambda_max[l]) <= 1e-8 * mg_o.lambda_max[l], l
    if case[10] == 1:
        assert info["chebyshev_degree"][0] == mg_o.cheb_degree[0]
    r = oracle.random_vector(op.n)
    z_ref = mg_o.vmult(r)
    rs = torch.from_numpy(r).cuda()
    zs = torch.empty_like(rs)
    prec.vmult(zs, rs)
    torch.cuda.synchronize()
    err = _rel(zs.cpu().numpy(), z_ref)
    print(f"{case[0]}: V-cycle FP64 rel.err {err:.2e}", flush=True)
    assert err <= 1e-9


@pytest.mark.parametrize("case", MG_CASES[:2], ids=[c[0] for c in MG_CASES[:2]])
def test_vcycle_fp32_close_to_oracle(case, torch_cuda, product):
    torch = torch_cuda
    op, prec, mg_o = _mg_pair(product, case, product.FP32)
    r = oracle.random_vector(op.n)
    z_ref = mg_o.vmult(r)
    rs = torch.from_numpy(r).cuda()
    zs = torch.empty_like(rs)
    prec.vmult(zs, rs)
    torch.cuda.synchronize()
    err = _rel(zs.cpu().numpy(), z_ref)
    print(f"{case[0]}: V-cycle FP32 rel.err vs FP64 oracle {err:.2e}", flush=True)
    assert err <= 1e-4


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", ["curved_k4", "periodic_singular_k4"])
def test_vcycle_graph_replays_match_eager(case, precision, torch_cuda, product):
    """With the Chebyshev coarse solver the V-cycle is replayed as a HIP graph from the second
    call on. Every replay (with eager work and synchronizations in between, and a second
    (dst, src) pair that needs its own graph) equals the eager V-cycle."""
    torch = torch_cuda
    prec_t = product.FP64 if precision == "fp64" else product.FP32
    if case == "curved_k4":
        mesh = product.mesh_desc(3, deformation=0.1, mapping_degree=3)
        bc, singular = product.bc_desc((D, N, D, D, D, D)), False
    else:
        mesh = product.mesh_desc(4, periodic=(1, 1, 0))
        bc, singular = product.bc_desc((N,) * 6), True
    op = product.LaplaceOperator(mesh, bc, 4)
    descs = [product.mg_desc(4, precision=prec_t, coarse_solver=product.COARSE_CHEBYSHEV,
                             singular=singular, capture_graph=g) for g in (True, False)]
    pg = product.Preconditioner(op, product.PRECOND_MULTIGRID, descs[0])
    pe = product.Preconditioner(op, product.PRECOND_MULTIGRID, descs[1])
    r = torch.from_numpy(oracle.random_vector(op.n)).cuda()
    r2 = torch.from_numpy(oracle.random_vector(op.n, seed=7)).cuda()
    ze, ze2 = torch.empty_like(r), torch.empty_like(r)
    pe.vmult(ze, r)
    pe.vmult(ze2, r2)
    torch.cuda.synchronize()
    zg, zg2, y = torch.empty_like(r), torch.empty_like(r), torch.empty_like(r)
    scale = float(ze.abs().max())
    for rep in range(5):
        zg.fill_(float("nan"))
        pg.vmult(zg, r)
        op.vmult(y, r)  # eager work between replays
        torch.cuda.synchronize()
        assert float((zg - ze).abs().max()) <= 1e-14 * scale, rep
        if rep >= 2:
            zg2.zero_()
            pg.vmult(zg2, r2)
            torch.cuda.synchronize()
            assert float((zg2 - ze2).abs().max()) <= 1e-14 * float(ze2.abs().max()), rep
    # a solve through the graph takes the eager solve's iterations
    b = torch.zeros_like(r)
    op.vmult(b, r)
    if singular:
        product.subtract_mean(b)
    xg, xe = torch.zeros_like(b), torch.zeros_like(b)
    ig = product.cg_solve(op, b, xg, pg, rel_tol=1e-10, rel_mode=1)
    ie = product.cg_solve(op, b, xe, pe, rel_tol=1e-10, rel_mode=1)
    assert ig["iterations"] == ie["iterations"] and ig["converged"]
    xscale = float(xe.abs().max())
    assert float((xg - xe).abs().max()) <= 1e-12 * xscale
    # another (dst, src) pair captured after the solve's own graph, then a second solve that
    # replays the solve's graph: each graph returns the solution vector it was captured with
    z3 = torch.full_like(r, float("nan"))
    pg.vmult(z3, r)
    torch.cuda.synchronize()
    assert float((z3 - ze).abs().max()) <= 1e-14 * scale
    xg2 = torch.zeros_like(b)
    ig2 = product.cg_solve(op, b, xg2, pg, rel_tol=1e-10, rel_mode=1)
    assert ig2["iterations"] == ig["iterations"]
    assert float((xg2 - xg).abs().max()) <= 1e-14 * xscale


def _sine(product, k, curvilinear):
    dfm = 0.15 if curvilinear else 0.0
    mp = product.mesh_desc(8, -1.0, 1.0, deformation=dfm, frequency=2, mapping_degree=3)
    bc = product.bc_desc((D, N, D, D, D, D))
    mo = oracle.make_mesh(8, -1.0, 1.0, deformation=dfm, frequency=2, mapping_degree=3,
                          boundary_type=(D, N, D, D, D, D))
    return product.LaplaceOperator(mp, bc, k), oracle.LaplaceOperator(mo, k)


@pytest.mark.parametrize("mesh", ["cartesian", "curvilinear"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_pcg_multigrid_golden_error(mesh, precision, torch_cuda, product):
    torch = torch_cuda
    k = 3
    op, ref = _sine(product, k, mesh == "curvilinear")
    prec = product.Preconditioner(op, product.PRECOND_MULTIGRID,
                                  product.mg_desc(k, precision=product.FP64 if precision == "fp64"
                                                  else product.FP32))
    b = torch.zeros(op.n, dtype=torch.float64, device="cuda")
    op.rhs_sine(3 * np.pi, b)
    x = torch.zeros_like(b)
    info = product.cg_solve(op, b, x, prec, rel_tol=1e-10)
    torch.cuda.synchronize()
    assert info["converged"] and info["r_final"] <= 1e-10 * info["r0"]
    xs = x.cpu().numpy()
    # true residual in FP64 by the oracle operator
    bh = b.cpu().numpy()
    assert np.linalg.norm(bh - ref.vmult(xs)) <= 2e-10 * np.linalg.norm(bh)
    err = ref.error_sine(3 * np.pi, xs)
    err_gpu = op.error_sine(3 * np.pi, x)
    print(f"{mesh} {precision}: {info['iterations']} its, err {err:.5e} (GPU {err_gpu:.5e})", flush=True)
    assert f"{err:.5e}" == f"{GOLD[mesh][str(k)]:.5e}"
    assert f"{err_gpu:.5e}" == f"{GOLD[mesh][str(k)]:.5e}"
    assert abs(err_gpu - err) <= 1e-10 * err
    assert info["iterations"] <= 25


def test_pcg_jacobi_matches_oracle_iterations(torch_cuda, product):
    torch = torch_cuda
    op, ref = _sine(product, 2, True)
    prec = product.Preconditioner(op, product.PRECOND_POINT_JACOBI)
    b = torch.zeros(op.n, dtype=torch.float64, device="cuda")
    op.rhs_sine(3 * np.pi, b)
    x = torch.zeros_like(b)
    info = product.cg_solve(op, b, x, prec, rel_tol=1e-10)
    xo, info_o = ref.cg_solve(b.cpu().numpy(), rel_tol=1e-10)
    assert abs(info["iterations"] - info_o["iterations"]) <= 1
    assert _rel(x.cpu().numpy(), xo) <= 1e-8
    # unpreconditioned CG also converges
    x2 = torch.zeros_like(b)
    info2 = product.cg_solve(op, b, x2, None, rel_tol=1e-8)
    assert info2["converged"]


def test_pcg_cartesian_k3_8c_dirichlet_box(torch_cuda, product):
    """[0,1]^3, 8^3 cells, k=3, u = prod sin(pi x), Dirichlet; P = DG3 -> DG1 -> CG-Q1, FP32 MG."""
    torch = torch_cuda
    op = product.LaplaceOperator(product.mesh_desc(8), product.bc_desc(), 3)
    ref = oracle.LaplaceOperator(oracle.make_mesh(8), 3)
    prec = product.Preconditioner(op, product.PRECOND_MULTIGRID, product.mg_desc(3))
    b = torch.zeros(op.n, dtype=torch.float64, device="cuda")
    op.rhs_sine(np.pi, b)
    x = torch.zeros_like(b)
    info = product.cg_solve(op, b, x, prec, rel_tol=1e-10, rel_mode=1)
    assert info["converged"] and info["r_final"] <= 1e-10 * info["b_norm"]
    xo, _ = oracle.cg_solve_mg(ref, oracle.Multigrid(oracle.make_mesh(8), [3, 1]),
                               b.cpu().numpy(), rel_tol=1e-12)
    assert _rel(x.cpu().numpy(), xo) <= 1e-8
    assert ref.error_sine(np.pi, x.cpu().numpy()) < 1e-3


def test_solver_error_paths(torch_cuda, product):
    torch = torch_cuda
    op = product.LaplaceOperator(product.mesh_desc(3), product.bc_desc(), 2)
    b = torch.ones(op.n, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    info = product.cg_solve(op, b, x, None, max_iter=1, raise_on_error=False)
    assert info["status"] == 4 and not info["converged"]
    b[5] = float("nan")
    x.zero_()
    info = product.cg_solve(op, b, x, None, raise_on_error=False)
    assert info["status"] == 3 and info["not_finite"]
    with pytest.raises(product.ExaDGError):
        product.cg_solve(op, b, b, None)
    # FP32 operator cannot drive the FP64 outer solver
    op32 = product.LaplaceOperator(product.mesh_desc(3), product.bc_desc(), 2, product.FP32)
    with pytest.raises(product.ExaDGError):
        product.Preconditioner(op32, product.PRECOND_POINT_JACOBI)
