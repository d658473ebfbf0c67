"""GPU parity of the weak DG divergence of a discrete velocity (the pressure-step right-hand
side) against the oracle, the Fourier interpolation of the synthetic velocity, and the
dual-splitting pressure step driven through the C ABI."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
D, N = 0, 1

CASES = [
    # name, n_cells, deformation, mapping degree, periodic, pressure bc, k_p, k_u
    ("cart_walls", (3, 2, 2), 0.0, 1, (0, 0, 0), (N,) * 6, 2, 3),
    ("cart_outflow", (2, 3, 2), 0.0, 1, (0, 0, 0), (D, N, N, N, N, D), 3, 4),
    ("curved_walls", (3, 2, 3), 0.1, 3, (0, 0, 0), (N,) * 6, 3, 4),
    ("periodic_xy", (3, 3, 2), 0.0, 1, (1, 1, 0), (N,) * 6, 4, 5),
    ("curved_periodic", (2, 3, 2), 0.08, 4, (1, 0, 1), (D,) * 6, 4, 4),
    ("single_periodic", (1, 2, 2), 0.0, 1, (1, 0, 0), (N,) * 6, 1, 2),
    ("curved_k5", (2, 2, 2), 0.1, 5, (1, 1, 0), (N,) * 6, 5, 6),
]


def _pair(product, c):
    name, nc, dfm, md, per, bc, kp, ku = c
    op = product.LaplaceOperator(product.mesh_desc(nc, deformation=dfm, mapping_degree=md,
                                                   periodic=per), product.bc_desc(bc), kp)
    ref = oracle.LaplaceOperator(oracle.make_mesh(nc, deformation=dfm, mapping_degree=md,
                                                  periodic=per, boundary_type=bc), kp)
    return op, product.DivergenceOperator(op, ku), ref


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_divergence_matches_oracle(case, torch_cuda, product):
    torch = torch_cuda
    op, div, ref = _pair(product, case)
    assert div.n_pressure == op.n == ref.n
    u = oracle.random_vector(div.n_velocity, seed=31)
    b_ref = ref.divergence_apply(case[7], u)
    b = torch.full((op.n,), float("nan"), dtype=torch.float64, device="cuda")
    div.apply(b, torch.from_numpy(u).cuda())
    torch.cuda.synchronize()
    err = np.linalg.norm(b.cpu().numpy() - b_ref) / np.linalg.norm(b_ref)
    print(f"{case[0]}: divergence rel.err {err:.2e}", flush=True)
    assert err <= 1e-12


def test_fourier_interpolation_at_the_nodes(torch_cuda, product):
    """Node values equal the Fourier sum evaluated in numpy at the node positions; on an affine
    mesh the nodes are the mapped Gauss-Lobatto points, cell by cell, x fastest."""
    torch = torch_cuda
    op = product.LaplaceOperator(product.mesh_desc((2, 3, 2), periodic=(1, 1, 0)),
                                 product.bc_desc((N,) * 6), 2)
    div = product.DivergenceOperator(op, 3)
    modes = product.fourier_modes(4, 1, 3)
    u = torch.empty(div.n_velocity, dtype=torch.float64, device="cuda")
    div.interpolate_fourier(modes, 0.23, u)
    x = div.velocity_points().cpu().numpy()
    ph = 2 * np.pi * (modes[:, :3] @ x) + (modes[:, 7] + modes[:, 8] * 0.23)[:, None]
    ref = np.einsum("mk,m,mp->kp", modes[:, 3:6], modes[:, 6], np.cos(ph))  # (3, points)
    n3 = 4 ** 3
    got = u.cpu().numpy().reshape(-1, 3, n3).transpose(1, 0, 2).reshape(3, -1)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    gl = oracle.tables(3)["gl"]
    xc = x.reshape(3, 2, 3, 2, 4, 4, 4)  # (comp, cz, cy, cx, pz, py, px)
    for cx in range(2):
        assert np.allclose(xc[0, 0, 0, cx, 0, 0, :], (cx + gl) / 2, atol=1e-15)


def test_pressure_step_solves_the_pressure_poisson_problem(torch_cuda, product):
    """One dual-splitting pressure step on a periodic-x/y box with walls in z: the right-hand side
    is -gamma0/dt times the oracle's divergence (mean removed), the extrapolated guess is used,
    and the solution reaches the FP64 true residual 1e-10 ||b|| with zero mean."""
    torch = torch_cuda
    op = product.LaplaceOperator(product.mesh_desc(4, periodic=(1, 1, 0)), product.bc_desc((N,) * 6), 3)
    prec = product.Preconditioner(op, product.PRECOND_MULTIGRID,
                                  product.mg_desc(3, singular=True, coarse_solver=product.COARSE_CHEBYSHEV))
    div = product.DivergenceOperator(op, 4)
    ref = oracle.LaplaceOperator(oracle.make_mesh(4, periodic=(1, 1, 0), boundary_type=(N,) * 6), 3)
    modes = product.fourier_modes(8, 1, 3)
    dt, gamma0 = 0.01, 1.5
    u = torch.empty(div.n_velocity, dtype=torch.float64, device="cuda")
    div.interpolate_fourier(modes, 0.1, u)
    hist = [torch.from_numpy(oracle.random_vector(op.n, seed=s)).cuda() * 1e-3 for s in (1, 2)]
    p = torch.zeros(op.n, dtype=torch.float64, device="cuda")
    rhs = torch.empty_like(p)
    info = product.pressure_step(op, prec, div, u, -gamma0 / dt, hist, p, rhs)
    torch.cuda.synchronize()
    b_ref = -gamma0 / dt * ref.divergence_apply(4, u.cpu().numpy())
    b_ref -= b_ref.mean()
    assert np.linalg.norm(rhs.cpu().numpy() - b_ref) <= 1e-12 * np.linalg.norm(b_ref)
    assert info["converged"] and info["r_true"] <= 1e-10 * info["b_norm"] * 1.01
    ph = p.cpu().numpy()
    assert abs(ph.mean()) <= 1e-14 * np.abs(ph).max()
    r = ref.vmult(ph) - b_ref
    assert np.linalg.norm(r) <= 2e-10 * np.linalg.norm(b_ref)
    # the extrapolated guess: order 2 from the history, the same solve from it by hand
    x = 2 * hist[0] - hist[1]
    info2 = product.cg_solve(op, rhs, x, prec, rel_tol=1e-10, rel_mode=1)
    product.subtract_mean(x)
    torch.cuda.synchronize()
    assert info2["iterations"] == info["iterations"]
    assert float((x - p).abs().max()) <= 1e-13 * float(p.abs().max())
