"""GPU parity of the SIPG vmult, diagonal and RHS against the CPU oracle (through the C ABI).

Bar: FP64 relative error <= 1e-12; FP32 <= 2e-6 relative to ||A x||.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

D, N = 0, 1


def _cases():
    cases = []
    # (name, n_cells, lower, upper, deformation, frequency, mapping_degree, periodic, bc)
    for k in range(1, 9):
        cases.append(("cart", k, (3, 2, 4), (0.0, -0.5, 0.2), (1.0, 0.5, 1.4), 0.0, 1, 1,
                      (0, 0, 0), (D, N, D, D, N, D)))
        # MappingQ(k), the mapping of the curved operator sweep and the mixed-precision solve
        cases.append(("curved", k, (3, 3, 2), 0.0, 1.0, 0.1, 1, k, (0, 0, 0),
                      (D, D, N, D, D, D)))
    cases.append(("periodic_xy", 4, (4, 3, 3), 0.0, 1.0, 0.0, 1, 1, (1, 1, 0), (D,) * 4 + (N, N)))
    cases.append(("periodic_curved", 3, (3, 2, 3), 0.0, 1.0, 0.08, 1, 3, (1, 0, 1), (D,) * 6))
    cases.append(("single_periodic", 2, (1, 2, 2), 0.0, 1.0, 0.0, 1, 1, (1, 0, 0), (D,) * 6))
    # curved with one cell along a periodic direction: the diagonal falls back to probing there
    cases.append(("single_periodic_curved", 2, (1, 2, 2), 0.0, 1.0, 0.1, 1, 2, (1, 0, 0),
                  (D,) * 6))
    cases.append(("sine_curv", 3, (4, 4, 4), -1.0, 1.0, 0.15, 2, 3, (0, 0, 0), (D, N, D, D, D, D)))
    return cases


CASES = _cases()


def _build(product, c, precision):
    name, k, nc, lo, hi, dfm, fr, md, per, bc = c
    m_o = oracle.make_mesh(nc, lo, hi, deformation=dfm, frequency=fr, mapping_degree=md,
                           periodic=per, boundary_type=bc)
    m_p = product.mesh_desc(nc, lo, hi, deformation=dfm, frequency=fr, mapping_degree=md,
                            periodic=per)
    op_p = product.LaplaceOperator(m_p, product.bc_desc(bc), k, precision=precision)
    op_o = oracle.LaplaceOperator(m_o, k)
    return op_p, op_o


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-k{c[1]}" for c in CASES])
def test_vmult_fp64_matches_oracle(case, torch_cuda, product):
    torch = torch_cuda
    op_p, op_o = _build(product, case, product.FP64)
    assert op_p.n == op_o.n
    assert op_p.is_affine == (case[5] == 0.0)
    x = oracle.random_vector(op_o.n)
    y_ref = op_o.vmult(x)
    xs = torch.from_numpy(x).cuda()
    ys = torch.full_like(xs, float("nan"))
    op_p.vmult(ys, xs)
    torch.cuda.synchronize()
    err = _rel(ys.cpu().numpy(), y_ref)
    print(f"{case[0]} k={case[1]} n={op_o.n} rel.err={err:.3e}", flush=True)
    assert err <= 1e-12
    # vmult_add accumulates: dst = y + A x
    op_p.vmult_add(ys, xs)
    torch.cuda.synchronize()
    assert _rel(ys.cpu().numpy(), 2 * y_ref) <= 1e-12


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-k{c[1]}" for c in CASES])
def test_vmult_fp32_matches_oracle(case, torch_cuda, product):
    """FP32 operator (the multigrid levels), vmult and vmult_add, on every mesh case."""
    torch = torch_cuda
    op_p, op_o = _build(product, case, product.FP32)
    x = oracle.random_vector(op_o.n)
    y_ref = op_o.vmult(x)
    xs = torch.from_numpy(x.astype(np.float32)).cuda()
    ys = torch.empty_like(xs)
    op_p.vmult(ys, xs)
    torch.cuda.synchronize()
    err = _rel(ys.cpu().numpy().astype(np.float64), y_ref)
    print(f"fp32 {case[0]} k={case[1]} rel.err={err:.3e}", flush=True)
    assert err <= 2e-6
    op_p.vmult_add(ys, xs)
    torch.cuda.synchronize()
    assert _rel(ys.cpu().numpy().astype(np.float64), 2 * y_ref) <= 2e-6


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-k{c[1]}" for c in CASES])
def test_diagonal_fp32_matches_oracle(case, torch_cuda, product):
    """FP32 diagonal (the Jacobi part of the FP32 smoothers), every mesh case: the closed forms
    on affine and curved cells, and probing on a curved cell that is its own periodic
    neighbour."""
    torch = torch_cuda
    op_p, op_o = _build(product, case, product.FP32)
    d_ref = op_o.diagonal()
    d = torch.zeros(op_o.n, dtype=torch.float32, device="cuda")
    op_p.calculate_diagonal(d)
    torch.cuda.synchronize()
    assert _rel(d.cpu().numpy().astype(np.float64), d_ref) <= 2e-6


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-k{c[1]}" for c in CASES])
def test_diagonal_matches_oracle(case, torch_cuda, product):
    torch = torch_cuda
    op_p, op_o = _build(product, case, product.FP64)
    d_ref = op_o.diagonal()
    d = torch.zeros(op_o.n, dtype=torch.float64, device="cuda")
    op_p.calculate_diagonal(d)
    torch.cuda.synchronize()
    assert _rel(d.cpu().numpy(), d_ref) <= 1e-12
    op_p.calculate_inverse_diagonal(d)
    torch.cuda.synchronize()
    assert _rel(d.cpu().numpy(), 1.0 / d_ref) <= 1e-12


@pytest.mark.parametrize("case", CASES[1::4], ids=[f"{c[0]}-k{c[1]}" for c in CASES[1::4]])
def test_rhs_matches_oracle(case, torch_cuda, product):
    torch = torch_cuda
    op_p, op_o = _build(product, case, product.FP64)
    omega = 3 * np.pi
    b_ref = op_o.rhs_sine(omega)
    b = torch.zeros(op_o.n, dtype=torch.float64, device="cuda")
    op_p.rhs_sine(omega, b)
    torch.cuda.synchronize()
    assert _rel(b.cpu().numpy(), b_ref) <= 1e-12


def test_vmult_rejects_aliasing(torch_cuda, product):
    torch = torch_cuda
    op = product.LaplaceOperator(product.mesh_desc(2), product.bc_desc(), 2)
    x = torch.zeros(op.n, dtype=torch.float64, device="cuda")
    with pytest.raises(product.ExaDGError) as e:
        op.vmult(x, x)
    assert e.value.code == 1


def test_vmult_full_size_affine_k4(torch_cuda, product):
    """64^3 cells, k=4 (32.8M DoF): exact parity with the oracle and symmetry x^T A y = y^T A x."""
    torch = torch_cuda
    op_p = product.LaplaceOperator(product.mesh_desc(64), product.bc_desc(), 4)
    op_o = oracle.LaplaceOperator(oracle.make_mesh(64), 4)
    assert op_p.n == 32768000
    x = oracle.random_vector(op_o.n)
    y_ref = op_o.vmult(x)
    xs = torch.from_numpy(x).cuda()
    ys = torch.empty_like(xs)
    op_p.vmult(ys, xs)
    torch.cuda.synchronize()
    y = ys.cpu().numpy()
    err = _rel(y, y_ref)
    print(f"64^3 k=4 rel.err={err:.3e}", flush=True)
    assert err <= 1e-12
    z = oracle.random_vector(op_o.n, seed=12345)
    zs = torch.from_numpy(z).cuda()
    ws = torch.empty_like(zs)
    op_p.vmult(ws, zs)
    torch.cuda.synchronize()
    a = float(torch.dot(zs, ys))
    b = float(torch.dot(xs, ws))
    assert abs(a - b) <= 1e-11 * (abs(a) + abs(b))


@pytest.mark.parametrize("k,curved", [(2, False), (4, False), (3, True), (6, True)])
def test_error_sine_vs_oracle(torch_cuda, product, k, curved):
    """exadg_laplace_error_sine equals the oracle's QGauss(k+3) error norm (calculate_error)."""
    torch = torch_cuda
    dfm, md = (0.12, 3) if curved else (0.0, 1)
    mp = product.mesh_desc((3, 4, 2), -1.0, 1.0, deformation=dfm, frequency=2, mapping_degree=md)
    mo = oracle.make_mesh((3, 4, 2), -1.0, 1.0, deformation=dfm, frequency=2, mapping_degree=md)
    op = product.LaplaceOperator(mp, product.bc_desc(), k)
    ref = oracle.LaplaceOperator(mo, k)
    u = oracle.random_vector(op.n) * 0.1 + 1.0
    e_gpu = op.error_sine(1.3 * np.pi, torch.from_numpy(u).cuda())
    e_ref = ref.error_sine(1.3 * np.pi, u)
    assert abs(e_gpu - e_ref) <= 1e-12 * e_ref


def test_random_vector_matches_oracle(torch_cuda, product):
    """exadg_vector_random is the counter-based SplitMix64 vector of the oracle, bit for bit, at
    any length (the grid-stride loop covers lengths beyond one grid), in FP64 and FP32."""
    for n, seed in ((1, 7), (1000, oracle.DEFAULT_SEED), (3_000_017, 12345)):
        v = product.random_vector(n, seed=seed)
        assert np.array_equal(v.cpu().numpy(), oracle.random_vector(n, seed=seed))
    v32 = product.random_vector(1000, precision=product.FP32)
    assert np.array_equal(v32.cpu().numpy(), oracle.random_vector(1000).astype(np.float32))
    assert not np.array_equal(product.random_vector(64, seed=1).cpu().numpy(),
                              product.random_vector(64, seed=2).cpu().numpy())
    assert product.random_vector(0).numel() == 0
