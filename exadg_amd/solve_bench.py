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
                rel_residual=info["r_final"] / info["b_norm"],
                true_rel_residual=info["r_true"] / info["b_norm"]), x, op


def compare_arms(ex, torch, mesh, bc, degree, degrees, omega, coarse_solver, reps=1):
    r64, x64, op = run_solve(ex, torch, mesh, bc, degree, degrees, omega, coarse_solver, ex.FP64, reps)
    r32, x32, _ = run_solve(ex, torch, mesh, bc, degree, degrees, omega, coarse_solver, ex.FP32, reps)
    r32["rel_diff_vs_fp64_solution"] = float(torch.linalg.norm(x32 - x64) / torch.linalg.norm(x64))
    return {"fp64_mg": r64, "fp32_mg": r32,
            "speedup_fp32_mg": round(r64["time_s"] / r32["time_s"], 3)}


def mg_timing(ex, torch, n_cells=32, degree=4, reps=20, solves=5):
    """V-cycle and solve wall times on the pressure-sequence box (periodic x/y, Neumann z) for
    FP32/FP64 V-cycles with the CG and the Chebyshev coarse solver; the right-hand side is the
    Fourier divergence field at t = 0.5. Returns a list of dicts."""
    import time as _t
    mesh = ex.mesh_desc(n_cells, periodic=(1, 1, 0))
    bc = ex.bc_desc((ex.NEUMANN,) * 6)
    op = ex.LaplaceOperator(mesh, bc, degree)
    modes = ex.fourier_modes(5, 1, 8)
    b = torch.empty(op.n, dtype=torch.float64, device="cuda")
    op.rhs_divergence_fourier(modes, 0.5, b)
    x, y = torch.zeros_like(b), torch.zeros_like(b)

    def timeit(f, n):
        f()
        torch.cuda.synchronize()
        t0 = _t.perf_counter()
        for _ in range(n):
            f()
        torch.cuda.synchronize()
        return (_t.perf_counter() - t0) / n * 1e3

    out = []
    for prec in (ex.FP32, ex.FP64):
        for cs in (ex.COARSE_CG, ex.COARSE_CHEBYSHEV):
            p = ex.Preconditioner(op, ex.PRECOND_MULTIGRID,
                                  ex.mg_desc(degree, precision=prec, coarse_solver=cs, singular=True))
            tv = timeit(lambda: p.vmult(y, b), reps)

            def solve():
                x.zero_()
                return ex.cg_solve(op, b, x, p, rel_tol=1e-10, rel_mode=1)
            ts = timeit(solve, solves)
            out.append({"vcycle": "fp32" if prec == ex.FP32 else "fp64",
                        "coarse": "cg" if cs == ex.COARSE_CG else "chebyshev",
                        "vcycle_ms": round(tv, 3), "solve_ms": round(ts, 2),
                        "iterations": solve()["iterations"],
                        "coarse_iterations": p.mg_info()["last_coarse_iterations"]})
    return out


if __name__ == "__main__":
    import argparse
    import json
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import torch

    import exadg_amd
    ap = argparse.ArgumentParser(description="multigrid V-cycle / solve timings (one JSON line)")
    ap.add_argument("--cells", type=int, default=32)
    ap.add_argument("--degree", type=int, default=4)
    a = ap.parse_args()
    print(json.dumps(mg_timing(exadg_amd, torch, a.cells, a.degree)), flush=True)
