"""A sequence of pressure-Poisson solves with a learned initial guess.

The sequence mimics the pressure step of the dual-splitting scheme
(time_int_bdf_dual_splitting.cpp:489-540): at every time step t_s = s dt the right-hand side is
the weak DG divergence of a solenoidal Fourier velocity field with no-slip walls, the initial
guess is extrapolated from the previous solutions (orders ramp up 1, 2, 3 at start-up as in
ExtrapolationConstants), and CG + hybrid p-multigrid runs to ||r|| <= rel_tol ||b||. The
operator is singular on a periodic-x/y, walled-z box (Neumann pressure on the walls), so the
right-hand side and every solution have their mean removed.

Arms on the held-out steps, all from the same exact history of previous solutions:
  zero guess, order-3 extrapolation (the reference's guess), learned guess (GuessNet).
The learned guess is trained on the first steps only (no leakage).
"""
import time

import numpy as np


def setup_box(ex, n_cells, degree, precision_mg, coarse_solver=None, deformation=0.0,
              mapping_degree=1, op=None):
    """Periodic-x/y, Neumann-z box on [0,1]^3 (optionally deformed by the DeformedCubeManifold
    map, MappingQ(mapping_degree)) with CG + hybrid p-MG; the coarse Q1 problem is solved by
    Chebyshev (MGCoarseChebyshev) unless stated otherwise. An existing operator may be passed
    to build only another preconditioner."""
    if coarse_solver is None:
        coarse_solver = ex.COARSE_CHEBYSHEV
    if op is not None:
        return op, ex.Preconditioner(op, ex.PRECOND_MULTIGRID,
                                     ex.mg_desc(degree, precision=precision_mg, singular=True,
                                                coarse_solver=coarse_solver))
    mesh = ex.mesh_desc(n_cells, periodic=(1, 1, 0), deformation=deformation,
                        mapping_degree=mapping_degree)
    bc = ex.bc_desc((ex.NEUMANN,) * 6)
    op = ex.LaplaceOperator(mesh, bc, degree)
    prec = ex.Preconditioner(op, ex.PRECOND_MULTIGRID,
                             ex.mg_desc(degree, precision=precision_mg, singular=True,
                                        coarse_solver=coarse_solver))
    return op, prec


def generate(ex, torch, op, prec, modes, n_steps, dt=0.01, rel_tol=1e-10, order=3, log=None):
    """Solve the whole sequence with the extrapolated guess. Returns FP64 solutions P and
    right-hand sides Bv (n_steps x n, device) and per-step iterations / times."""
    n = op.n
    P = torch.empty((n_steps, n), dtype=torch.float64, device="cuda")
    Bv = torch.empty((n_steps, n), dtype=torch.float64, device="cuda")
    its, times = [], []
    for s in range(n_steps):
        op.rhs_divergence_fourier(modes, s * dt, Bv[s])
        x = P[s]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        o = min(s, order)
        if o == 0:
            x.zero_()
        else:
            ex.extrapolate([P[s - 1 - i] for i in range(o)], x)
        info = ex.cg_solve(op, Bv[s], x, prec, rel_tol=rel_tol, rel_mode=1)
        ex.subtract_mean(x)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        its.append(info["iterations"])
        if log and (s % 50 == 0 or s == n_steps - 1):
            log(f"step {s}: {info['iterations']} iterations")
    return P, Bv, its, times


def solve_from(ex, torch, op, prec, b, x0, rel_tol=1e-10):
    x = x0.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info = ex.cg_solve(op, b, x, prec, rel_tol=rel_tol, rel_mode=1)
    torch.cuda.synchronize()
    return info, time.perf_counter() - t0, x


def ab_test(ex, torch, op, prec, net, P, Bv, s_begin, s_end, rel_tol=1e-10,
            arms=("zero", "extrapolation", "learned")):
    """Iterations and times of the guesses `arms` on steps [s_begin, s_end). Every arm starts from
    the same exact history; the time of an arm includes computing its guess."""
    H = 3 if net is None else net.history
    res = {a: [] for a in arms}
    fallbacks = 0
    x = torch.empty_like(P[0])
    for s in range(s_begin, s_end):
        hist = [P[s - 1 - i] for i in range(H)]
        b = Bv[s]
        bn = float(torch.linalg.norm(b))
        for arm in arms:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if arm == "zero":
                x.zero_()
            elif arm.startswith("extrapolation"):
                ex.extrapolate(hist, x)
            else:
                fallbacks += int(net.predict(hist, b, x))
            if arm.endswith("+smooth"):
                prec.smooth(op, b, x)
            info = ex.cg_solve(op, b, x, prec, rel_tol=rel_tol, rel_mode=1)
            torch.cuda.synchronize()
            res[arm].append((info["iterations"], time.perf_counter() - t0, info["r0"] / bn,
                             info["r_final"] / bn))
    out = {}
    for k, v in res.items():
        out[k] = {"mean_iterations": float(np.mean([r[0] for r in v])),
                  "total_time_s": float(sum(r[1] for r in v)),
                  "mean_initial_rel_residual": float(np.mean([r[2] for r in v])),
                  "max_final_rel_residual": float(max(r[3] for r in v))}
    for k in out:
        if k.startswith("learned"):
            out[k]["fallbacks"] = fallbacks
    return out


def zero_output_layer(net):
    """Set the output layer (W3, b3) to zero: the network's correction vanishes and the guess is
    exactly the extrapolation, i.e. a loss of 1 on every step."""
    p = net.params()
    p[-net.n_outputs * (net.width + 1):] = 0
    net.set_params(p)


def train(ex, net, torch, op, P, Bv, t_end, epochs, lr, loss=None, batch_cells=4096, seed=7,
          halve_every=5, from_extrapolation=False):
    """Train on steps [0, t_end) of the FP64 sequence; one call per epoch so that the per-epoch
    losses are returned. The learning rate halves every `halve_every` epochs, and the
    parameters after the epoch with the lowest training loss are kept: Adam on the residual
    loss can collapse late in training (the network output falling to zero, loss 1) after a
    good early epoch. Only training steps are looked at.

    With `from_extrapolation` the output layer starts at zero (the first guess is the
    extrapolation, loss 1, and the hidden layers keep their random start), and if the kept
    parameters do not beat the extrapolation on the training steps (loss >= 1, evaluated
    after training) the output layer is zeroed again, so the learned guess is never worse than
    the extrapolation on the data it has seen. The normalized losses make 1 the extrapolation's
    value for every loss kind."""
    if loss is None:
        loss = ex.GUESS_LOSS_ENERGY
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if from_extrapolation:
        zero_output_layer(net)
    losses, best, best_params = [], float("inf"), None
    for ep in range(epochs):
        rate = lr * 0.5 ** (ep // halve_every)
        losses.append(net.train(P, Bv, 0, t_end, 1, op=op, loss=loss, batch_cells=batch_cells,
                                lr=rate, seed=seed + ep))
        if losses[-1] < best:
            best, best_params = losses[-1], net.params()
    if best_params is not None and not losses[-1] <= best:
        net.set_params(best_params)
    if from_extrapolation:
        final = net.evaluate(P, Bv, 0, t_end, op=op, loss=loss)
        if not final < 1.0:
            zero_output_layer(net)
    torch.cuda.synchronize()
    return losses, time.perf_counter() - t0
