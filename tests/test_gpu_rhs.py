"""GPU parity of the right-hand side assembly (device kernel) against the CPU oracle.

* exadg_laplace_rhs_sine (f, Dirichlet g and Neumann h of u = prod sin(omega x_d), evaluated on
  the device) equals the oracle's face-centric host assembly to 1e-12 on every mesh case of the
  vmult tests (Cartesian, curved, periodic, mixed boundary types).
* The general interface: the sine data evaluated by the caller at the points the library names
  (exadg_laplace_rhs_points) and passed as values gives the same vector; on an affine mesh the
  points are the analytic Gauss points in the documented order.
* Linearity: rhs(f1 + 2 f2, g1 + 2 g2) = rhs(f1, g1) + 2 rhs(f2, g2); no data gives b = 0.
"""
import numpy as np
import pytest

import oracle
from tests.test_gpu_vmult import CASES, _build, _rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-k{c[1]}" for c in CASES])
def test_rhs_sine_on_device_matches_oracle(case, torch_cuda, product):
    torch = torch_cuda
    op_p, op_o = _build(product, case, product.FP64)
    omega = 2.7
    b = torch.full((op_o.n,), float("nan"), dtype=torch.float64, device="cuda")
    op_p.rhs_sine(omega, b)
    torch.cuda.synchronize()
    err = _rel(b.cpu().numpy(), op_o.rhs_sine(omega))
    print(f"rhs {case[0]} k={case[1]} rel.err {err:.2e}", flush=True)
    assert err <= 1e-12


def _sides(case):
    """(side, number of boundary points) in the library's boundary-point order."""
    name, k, nc, lo, hi, dfm, fr, md, per, bc = case
    n = k + 1
    out = []
    for side in range(6):
        d = side // 2
        if per[d]:
            continue
        t1, t2 = (1, 2) if d == 0 else (0, 2) if d == 1 else (0, 1)
        out.append((side, nc[t1] * nc[t2] * n * n))
    return out


def _point_data(torch, case, xc, xb, omega):
    """f = 3 omega^2 u at the cell points; g = u on Dirichlet sides, h = grad u . n on Neumann
    sides (the boundary of these meshes is the flat unit box side, normal +-e_d)."""
    bc = case[9]
    s = torch.sin(omega * xc)
    f = 3 * omega ** 2 * s[0] * s[1] * s[2]
    vals, off = [], 0
    for side, m in _sides(case):
        x = xb[:, off:off + m]
        sx, cx = torch.sin(omega * x), torch.cos(omega * x)
        d, sign = side // 2, (1.0 if side % 2 else -1.0)
        if bc[side] == 0:
            vals.append(sx[0] * sx[1] * sx[2])
        else:
            grad = [cx[0] * sx[1] * sx[2], sx[0] * cx[1] * sx[2], sx[0] * sx[1] * cx[2]]
            vals.append(sign * omega * grad[d])
        off += m
    g = torch.cat(vals) if vals else torch.zeros(0, dtype=torch.float64, device="cuda")
    return f.contiguous(), g.contiguous()


@pytest.mark.parametrize("case", CASES[::3], ids=[f"{c[0]}-k{c[1]}" for c in CASES[::3]])
def test_rhs_from_point_values_matches_oracle(case, torch_cuda, product):
    torch = torch_cuda
    op_p, op_o = _build(product, case, product.FP64)
    ncp, nbp = op_p.rhs_n_points()
    n = case[1] + 1
    assert ncp == op_o.n and nbp == sum(m for _, m in _sides(case))
    xc, xb = op_p.rhs_points()
    omega = 2.7
    f, g = _point_data(torch, case, xc, xb, omega)
    b = torch.empty(op_o.n, dtype=torch.float64, device="cuda")
    op_p.rhs(b, f, g if nbp > 0 else None)
    torch.cuda.synchronize()
    assert _rel(b.cpu().numpy(), op_o.rhs_sine(omega)) <= 1e-12
    if case[5] == 0.0:  # affine: the analytic Gauss points in the documented order
        name, k, nc, lo, hi = case[:5]
        lo = np.broadcast_to(np.asarray(lo, dtype=float), (3,))
        hi = np.broadcast_to(np.asarray(hi, dtype=float), (3,))
        gp = oracle.tables(k)["gp"]
        h = (hi - lo) / np.asarray(nc)
        xh = xc.cpu().numpy().reshape(3, nc[2], nc[1], nc[0], n, n, n)
        for d in range(3):
            cidx = np.arange(nc[d])
            ref = lo[d] + h[d] * (cidx[:, None] + gp[None, :])  # (cells_d, points_d)
            shape = [1] * 6
            shape[2 - d] = nc[d]
            shape[5 - d] = n
            got = xh[d]
            exp = np.broadcast_to(ref.reshape([nc[d] if i == 2 - d else n if i == 5 - d else 1
                                               for i in range(6)]), got.shape)
            assert np.abs(got - exp).max() <= 1e-14 * max(1.0, np.abs(hi).max())


def test_rhs_is_linear_in_the_data(torch_cuda, product):
    torch = torch_cuda
    case = CASES[3]
    op_p, _ = _build(product, case, product.FP64)
    ncp, nbp = op_p.rhs_n_points()
    gen = torch.Generator(device="cuda").manual_seed(3)
    f1, f2 = (torch.rand(ncp, dtype=torch.float64, device="cuda", generator=gen) for _ in range(2))
    g1, g2 = (torch.rand(nbp, dtype=torch.float64, device="cuda", generator=gen) for _ in range(2))
    b1, b2, b3 = (torch.empty(op_p.n, dtype=torch.float64, device="cuda") for _ in range(3))
    op_p.rhs(b1, f1, g1)
    op_p.rhs(b2, f2, g2)
    op_p.rhs(b3, f1 + 2 * f2, g1 + 2 * g2)
    torch.cuda.synchronize()
    assert float(torch.linalg.norm(b3 - b1 - 2 * b2)) <= 1e-13 * float(torch.linalg.norm(b3))
    op_p.rhs(b3)
    torch.cuda.synchronize()
    assert float(b3.abs().max()) == 0.0
