"""GPU parity of the hybrid multigrid over h-, p- and c-levels (C ABI) against the CPU oracle.

* The level list of every hierarchy equals the oracle's (multigrid_preconditioner_base.cpp:97-323).
* The FP64 V-cycle equals the oracle's V-cycle to 1e-9 (Chebyshev coarse solve, a fixed
  polynomial), and every level's Chebyshev eigenvalue estimate agrees to 1e-8: continuous Q_p
  levels (p = 1 and p > 1), c-transfers at p > 1, h-transfers on DG and continuous levels, on
  Cartesian, curved, mixed-boundary and periodic boxes.
* The FP32 V-cycle agrees with the FP64 oracle to 1e-4.
* PCG with the reference's cphMG hierarchy (sine/application.h:167) reproduces the golden sine
  errors on the GPU, with the FP64 true residual within the tolerance.
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
    # name, cells, deformation, mapping degree, periodic, bc, mg type, degrees, meshes
    ("cphMG_curved_k3", (4, 4, 4), 0.1, 3, (0, 0, 0), (D, N, D, D, D, D), "cphMG", [3, 1], 3),
    ("cphMG_cart_k4", (4, 2, 4), 0.0, 1, (0, 0, 0), (D,) * 6, "cphMG", [4, 2, 1], 2),
    ("pchMG_curved_k2", (4, 4, 2), 0.08, 2, (0, 0, 0), (D, D, N, D, D, D), "pchMG", [2, 1], 2),
    ("phMG_periodic_k3", (4, 4, 4), 0.0, 1, (1, 0, 0), (D, D, D, N, D, D), "phMG", [3, 1], 2),
    ("hMG_curved_k2", (4, 4, 4), 0.1, 2, (0, 0, 0), (D,) * 6, "hMG", [2], 3),
    ("hpMG_cart_k3", (4, 4, 2), 0.0, 1, (0, 0, 0), (D, N, D, D, D, D), "hpMG", [3, 1], 2),
    ("chMG_periodic_k2", (4, 2, 4), 0.0, 1, (0, 1, 0), (D,) * 6, "chMG", [2], 2),
    ("hcMG_curved_k2", (4, 4, 4), 0.1, 2, (0, 0, 0), (D, D, D, D, N, D), "hcMG", [2], 2),
    ("cpMG_curved_k5", (3, 2, 2), 0.1, 5, (0, 0, 0), (D, N, D, D, D, D), "cpMG", [5, 3, 2, 1], 1),
]


def _pair(product, c, precision):
    name, cells, dfm, md, per, bc, t, degs, H = c
    k = degs[0]
    mo = oracle.make_mesh(cells, deformation=dfm, mapping_degree=md, periodic=per, boundary_type=bc)
    mp = product.mesh_desc(cells, deformation=dfm, mapping_degree=md, periodic=per)
    op = product.LaplaceOperator(mp, product.bc_desc(bc), k)
    desc = product.mg_desc(k, degrees=degs, precision=precision, coarse_solver=1,
                           coarse_rel_tol=1e-3, mg_type=t, n_h_levels=H)
    prec = product.Preconditioner(op, product.PRECOND_MULTIGRID, desc)
    mg_o = oracle.Multigrid(mo, degs, coarse_solver=1, coarse_rel_tol=1e-3, mg_type=t, n_h_levels=H)
    return op, prec, mg_o


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_hybrid_vcycle_fp64_matches_oracle(case, torch_cuda, product):
    torch = torch_cuda
    print(f"case {case[0]}", flush=True)
    op, prec, mg_o = _pair(product, case, product.FP64)
    levels = prec.mg_levels()
    assert [(h, p, dg) for h, p, dg, _ in levels] == \
        list(zip(mg_o.h_levels, mg_o.degrees, mg_o.dg))
    assert [n for *_, n in levels] == mg_o.sizes
    info = prec.mg_info()
    for l in range(1, info["n_levels"]):
        assert abs(info["lambda_max"][l] - mg_o.lambda_max[l]) <= 1e-8 * mg_o.lambda_max[l], l
    assert info["chebyshev_degree"][0] == mg_o.cheb_degree[0]
    r = oracle.random_vector(op.n)
    z_ref = mg_o.vmult(r)
    rs = torch.from_numpy(r).cuda()
    zs = torch.empty_like(rs)
    prec.vmult(zs, rs)
    torch.cuda.synchronize()
    err = np.linalg.norm(zs.cpu().numpy() - z_ref) / np.linalg.norm(z_ref)
    print(f"  levels {[(h, p, 'DG' if dg else 'CG') for h, p, dg, _ in levels]}: "
          f"V-cycle FP64 rel.err {err:.2e}", flush=True)
    assert err <= 1e-9


@pytest.mark.parametrize("case", [CASES[0], CASES[4]], ids=[CASES[0][0], CASES[4][0]])
def test_hybrid_vcycle_fp32_close_to_oracle(case, torch_cuda, product):
    torch = torch_cuda
    print(f"case {case[0]} FP32", flush=True)
    op, prec, mg_o = _pair(product, case, product.FP32)
    r = oracle.random_vector(op.n)
    z_ref = mg_o.vmult(r)
    rs = torch.from_numpy(r).cuda()
    zs = torch.empty_like(rs)
    prec.vmult(zs, rs)
    torch.cuda.synchronize()
    err = np.linalg.norm(zs.cpu().numpy() - z_ref) / np.linalg.norm(z_ref)
    print(f"  V-cycle FP32 rel.err {err:.2e}", flush=True)
    assert err <= 1e-4


@pytest.mark.parametrize("mesh", ["cartesian", "curvilinear"])
@pytest.mark.parametrize("k", [3, 5])
@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_sine_golden_error_with_cphmg_on_gpu(mesh, k, precision, torch_cuda, product):
    # the sine application: 2^3 coarse cells refined twice, cphMG with Bisect, CG coarse 1e-3
    torch = torch_cuda
    print(f"sine {mesh} k={k} cphMG {precision}", flush=True)
    bc = (D, N, D, D, D, D)
    mp = product.mesh_desc(8, -1.0, 1.0, deformation=0.15 if mesh == "curvilinear" else 0.0,
                           frequency=2, mapping_degree=3)
    op = product.LaplaceOperator(mp, product.bc_desc(bc), k)
    desc = product.mg_desc(k, degrees=oracle.bisect_degrees(k), mg_type="cphMG", n_h_levels=0,
                           precision=getattr(product, precision))
    prec = product.Preconditioner(op, product.PRECOND_MULTIGRID, desc)
    levels = prec.mg_levels()
    assert levels[0][:3] == (0, 1, False) and levels[0][3] == 27  # 2^3 cells: 3^3 Q1 nodes
    assert levels[-2][1:3] == (k, False) and levels[-1][1:3] == (k, True)
    b = torch.empty(op.n, dtype=torch.float64, device="cuda")
    op.rhs_sine(3 * np.pi, b)
    x = torch.zeros_like(b)
    info = product.cg_solve(op, b, x, prec=prec, rel_tol=1e-10)
    err = op.error_sine(3 * np.pi, x)
    print(f"  {info['iterations']} iterations, error {err:.5e}, r_true/tol "
          f"{info['r_true'] / info['tolerance']:.2f}", flush=True)
    assert info["converged"] and info["iterations"] < 40
    assert info["r_true"] <= 10 * info["tolerance"]
    assert f"{err:.5e}" == f"{GOLD[mesh][str(k)]:.5e}"
