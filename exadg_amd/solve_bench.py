"""Time-to-solution of the Poisson solve (CG + hybrid p-multigrid) on the GPU.

Arms: FP64 outer CG with an FP64 V-cycle ("plain FP64 method") and with an FP32 V-cycle,
both to ||r|| <= 1e-10 ||b|| from a zero initial guess, manufactured sine right-hand side.
Used by bench.py; prints nothing itself.
"""
import time

import numpy as np


def run_solve(ex, torch, mesh, bc, degree, degrees, omega, coarse_solver, precision, reps=1):
    op = ex.LaplaceOperator(mesh, bc, degree)
    t0 = time.perf_counter()
    prec = ex.Preconditioner(op, ex.PRECOND_MULTIGRID,
                             ex.mg_desc(degree, degrees=degrees, precision=precision,
                                        coarse_solver=coarse_solver))
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    b = torch.zeros(op.n, dtype=torch.float64, device="cuda")
    op.rhs_sine(omega, b)
    x = torch.zeros_like(b)
    ex.cg_solve(op, b, x, prec, rel_tol=1e-10, rel_mode=1)  # warm-up
    times, info = [], None
    for _ in range(reps):
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = ex.cg_solve(op, b, x, prec, rel_tol=1e-10, rel_mode=1)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    return dict(n_dofs=op.n, iterations=info["iterations"], time_s=min(times), setup_s=t_setup,
                rel_residual=info["r_final"] / info["b_norm"]), x, op


def compare_arms(ex, torch, mesh, bc, degree, degrees, omega, coarse_solver, reps=1):
    r64, x64, op = run_solve(ex, torch, mesh, bc, degree, degrees, omega, coarse_solver, ex.FP64, reps)
    r32, x32, _ = run_solve(ex, torch, mesh, bc, degree, degrees, omega, coarse_solver, ex.FP32, reps)
    r32["rel_diff_vs_fp64_solution"] = float(torch.linalg.norm(x32 - x64) / torch.linalg.norm(x64))
    return {"fp64_mg": r64, "fp32_mg": r32,
            "speedup_fp32_mg": round(r64["time_s"] / r32["time_s"], 3)}
