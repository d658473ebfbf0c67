"""Snapshot I/O of device vectors: a bit-exact round trip (FP64 solutions, FP32 network
parameters), and a restart of a pressure sequence from a snapshot of its solution history that
continues exactly as the uninterrupted run (deterministic solver)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_device_round_trip(tmp_path, torch_cuda, product):
    torch = torch_cuda
    path = str(tmp_path / "dev.snapshot")
    g = torch.Generator(device="cuda").manual_seed(1)
    P = [torch.rand(70001, dtype=torch.float64, device="cuda", generator=g) for _ in range(4)]
    product.snapshot_write(path, P, time=0.04, step=4, degree=3)
    Q = [torch.empty_like(p) for p in P]
    info = product.snapshot_read(path, Q)
    assert info["step"] == 4 and info["degree"] == 3
    for a, b in zip(P, Q):
        assert torch.equal(a, b)
    net = product.GuessNet(2, history=3, width=32, seed=4)
    w = net.params()
    product.snapshot_write(path, [w], user=[1.0, 2.0, 3.0, 4.0, 5.0], keep=2)
    w2 = torch.zeros_like(w)
    assert product.snapshot_read(path, [w2])["user"][:5] == [1.0, 2.0, 3.0, 4.0, 5.0]
    assert torch.equal(w, w2)
    assert product.snapshot_read_info(path + ".old.1")["n"] == 70001


def test_restart_continues_the_sequence_exactly(tmp_path, torch_cuda, product):
    from exadg_amd import pressure_sequence as ps
    torch = torch_cuda
    op, prec = ps.setup_box(product, 4, 3, product.FP32)
    modes = product.fourier_modes(9, 1, 3)
    n_steps, cut, dt = 7, 4, 0.01

    def step(P, Bv, s):
        op.rhs_divergence_fourier(modes, s * dt, Bv[s])
        o = min(s, 3)
        if o == 0:
            P[s].zero_()
        else:
            product.extrapolate([P[s - 1 - i] for i in range(o)], P[s])
        info = product.cg_solve(op, Bv[s], P[s], prec, rel_tol=1e-10, rel_mode=1)
        product.subtract_mean(P[s])
        return info["iterations"]

    def fresh():
        return (torch.zeros((n_steps, op.n), dtype=torch.float64, device="cuda"),
                torch.zeros((n_steps, op.n), dtype=torch.float64, device="cuda"))

    P, Bv = fresh()
    for s in range(n_steps):
        step(P, Bv, s)
    # interrupted run: steps 0..cut-1, snapshot of the history, restart into new memory
    P1, B1 = fresh()
    for s in range(cut):
        step(P1, B1, s)
    path = str(tmp_path / "pressure.snapshot")
    product.snapshot_write(path, [P1[cut - 1], P1[cut - 2], P1[cut - 3]], time=(cut - 1) * dt,
                           step=cut - 1, degree=3)
    P2, B2 = fresh()
    info = product.snapshot_read(path, [P2[cut - 1], P2[cut - 2], P2[cut - 3]])
    assert info["step"] == cut - 1
    for s in range(cut, n_steps):
        step(P2, B2, s)
    for s in range(cut, n_steps):
        assert torch.equal(P2[s], P[s]), s
