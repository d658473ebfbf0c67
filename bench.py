"""Benchmark of both halves of the metric on one MI355X; prints ONE JSON line.

1. `value`: FP64 SIPG Laplace vmult throughput (GDoF/s). Affine Cartesian [0,1]^3, 64^3 cells,
   degree 4 (32,768,000 DoF), homogeneous Dirichlet operator, input seeded uniform[-1,1]
   (SplitMix64), already in HBM. One step = one vmult over the whole vector.
2. `solve`: time to an FP64 relative residual of 1e-10 on the curved degree-5 mesh (48^3 cells),
   outer FP64 CG with the FP32 V-cycle against the FP64 V-cycle, with the final true residual.
3. `combined`: the curved degree-5 pressure sequence (50^3 cells, 27M DoF): FP32 V-cycle +
   learned guess against the plain FP64 method (FP64 V-cycle + extrapolated guess); the default
   is a bounded 40-step sequence (train on 30, compare on 10).
4. `cpu_baseline`: the oracle vmult on the same 64^3 workload, on the host cores.

    python bench.py [--gpus 1] [--steps 100] [--warmup 10] [--no-cpu] [--no-solve] [--no-combined]
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DG Laplace vmult GDoF/s; time-to-1e-10 solve, learned+FP32 MG vs plain"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E vendor peak
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic_sipg_cartesian_k4_64c.json")


def algorithmic_bytes(n_dofs, n_cells, s=8):
    """src read + dst write + per-cell data (J^{-1}, JxW, tau): 2 s N + 32 N_c bytes."""
    return 2 * s * n_dofs + 32 * n_cells


def physical_cores():
    """Sockets x cores per socket from lscpu (None if unavailable)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except Exception:
        return None
    f = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            f[k.strip()] = v.strip()
    try:
        return int(f["Socket(s)"]) * int(f["Core(s) per socket"])
    except (KeyError, ValueError):
        return None


def cpu_baseline(degree, n_cells_1d, target_s=12.0):
    """Oracle FP64 vmult (C++/OpenMP, one thread per core of this process's CPU share) on the
    same workload as the GPU line, built on this host with -march=native; the median of the
    calls that fit in about target_s seconds (at least 3) after one warm-up call."""
    import tempfile

    import oracle
    tmp = tempfile.mkdtemp(prefix="oracle_native_")
    native = oracle.use_native_build(tmp) is not None
    op = oracle.LaplaceOperator(oracle.make_mesh(n_cells_1d), degree)
    x = oracle.random_vector(op.n)
    op.vmult(x)
    times = []
    t_end = time.perf_counter() + target_s
    while time.perf_counter() < t_end or len(times) < 3:
        t0 = time.perf_counter()
        op.vmult(x)
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    threads = oracle.num_threads()
    phys = physical_cores()
    return {"value": round(op.n / t / 1e9, 6), "unit": "GDoF/s", "cores": threads, "kind": "port",
            "sample": f"oracle vmult (face-centric C++/OpenMP restatement, "
                      f"{'-march=native' if native else 'x86-64-v3'}), {n_cells_1d}^3 cells "
                      f"k={degree} ({op.n} DoF) FP64, median of {len(times)} calls, {threads} "
                      f"OpenMP threads; host: {phys} physical cores, {os.cpu_count()} logical"}


def curved_bytes(k, n_cells_1d, s=8):
    """Algorithmic bytes of the curved vmult: src + dst, the per-point metric (6 per Gauss
    point), the per-face-point data (7 per point) and tau per face."""
    n = k + 1
    nc = n_cells_1d ** 3
    n_faces = 3 * n_cells_1d ** 2 * (n_cells_1d + 1)
    return 2 * s * nc * n ** 3 + 6 * s * n ** 3 * nc + 7 * s * n ** 2 * n_faces + s * n_faces


def run_sweep(ex, torch, args, fp32=False):
    """Curved-mesh operator sweep: [0,1]^3 deformed by x_d += 0.1 prod sin(pi x_e), MappingQ(k),
    k = 2, 3, 5, 7 at ~23.9M DoF each, FP64 (or FP32) vmult, seeded uniform[-1,1] input, median
    of HIP-event timings."""
    res = []
    dtype, s = (torch.float32, 4) if fp32 else (torch.float64, 8)
    for k, m in ((2, 96), (3, 72), (5, 48), (7, 36)):
        op = ex.LaplaceOperator(ex.mesh_desc(m, deformation=0.1, frequency=1, mapping_degree=k),
                                ex.bc_desc(), k, ex.FP32 if fp32 else ex.FP64)
        gen = torch.Generator(device="cuda").manual_seed(k)
        x = torch.rand(op.n, dtype=torch.float64, device="cuda", generator=gen).to(dtype) * 2 - 1
        y = torch.empty_like(x)
        for _ in range(3):
            op.vmult(y, x)
        stream = torch.cuda.current_stream()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(20)]
        for a, b in ev:
            a.record(stream)
            op.vmult(y, x)
            b.record(stream)
        torch.cuda.synchronize()
        t = statistics.median(a.elapsed_time(b) for a, b in ev) * 1e-3
        nb = curved_bytes(k, m, s)
        res.append({"degree": k, "cells": f"{m}^3", "n_dofs": op.n,
                    "gdofs_per_s": round(op.n / t / 1e9, 2), "ms": round(t * 1e3, 4),
                    "algorithmic_GBps": round(nb / t / 1e9, 1),
                    "frac_of_hbm_peak": round(nb / t / 1e9 / HBM_PEAK_GBS, 4)})
        del op, x, y
        torch.cuda.empty_cache()
    return res


def run_guess(ex, torch, args):
    """Pressure sequence on a periodic-x/y walled-z box (32^3 cells, k=4, 4.1M DoF by default);
    train on the first steps,
    A/B on the rest. FP32-MG arms: zero / extrapolated / learned guess; FP64-MG arms (the plain
    method): zero / extrapolated guess."""
    from exadg_amd import pressure_sequence as ps

    def log(msg):
        print(f"[guess] {msg}", file=sys.stderr, flush=True)

    k, nc = args.guess_degree, args.guess_cells
    op, prec32 = ps.setup_box(ex, nc, k, ex.FP32)
    _, prec64 = ps.setup_box(ex, nc, k, ex.FP64)
    modes = ex.fourier_modes(2026, 1, args.guess_kmax)
    log(f"{op.n} DoF, {len(modes)} Fourier modes; generating {args.guess_steps} steps")
    t0 = time.perf_counter()
    P, Bv, its, times = ps.generate(ex, torch, op, prec32, modes, args.guess_steps, log=log)
    t_gen = time.perf_counter() - t0
    net = ex.GuessNet(k, history=3, width=128, seed=1)
    log("training")
    losses, t_train = ps.train(ex, net, torch, op, P, Bv, args.guess_train, args.guess_epochs,
                               lr=args.guess_lr, loss=ex.GUESS_LOSS_RESIDUAL)
    held = net.evaluate(P, Bv, args.guess_train - 3, args.guess_steps, op=op,
                        loss=ex.GUESS_LOSS_RESIDUAL)
    log(f"train loss {losses[0]:.4g} -> {losses[-1]:.4g}, held-out {held:.4g}; A/B FP32 MG")
    ab32 = ps.ab_test(ex, torch, op, prec32, net, P, Bv, args.guess_train, args.guess_steps)
    log("A/B FP64 MG")
    ab64 = ps.ab_test(ex, torch, op, prec64, None, P, Bv, args.guess_train, args.guess_steps,
                      arms=("zero", "extrapolation"))
    plain = ab64["extrapolation"]["total_time_s"]
    return {
        "workload": f"pressure sequence, [0,1]^3 periodic x/y, Neumann walls z, {nc}^3 cells, "
                    f"degree {k}, {op.n} DoF, dt 0.01, {len(modes)} modes |n| <= {args.guess_kmax}",
        "steps": args.guess_steps, "train_steps": args.guess_train,
        "generation": {"time_s": round(t_gen, 3), "mean_iterations": float(np.mean(its[3:]))},
        "training": {"epochs": args.guess_epochs, "time_s": round(t_train, 3),
                     "loss_first_epoch": losses[0], "loss_last_epoch": losses[-1],
                     "held_out_loss": held,
                     "loss": "residual |A e|^2 of the guess per step, divided by the "
                             "extrapolation's: 1 = no better than the extrapolation"},
        "fp32_mg": ab32, "fp64_mg": ab64,
        "speedup_learned_fp32mg_vs_plain_fp64mg_extrapolated": round(
            plain / ab32["learned"]["total_time_s"], 3),
        "speedup_learned_fp32mg_vs_fp64mg_zero_guess": round(
            ab64["zero"]["total_time_s"] / ab32["learned"]["total_time_s"], 3),
    }


def run_combined(ex, torch, args):
    """The pressure sequence on a curved mesh (50^3 cells, k=5, deformation 0.1, MappingQ(5),
    27M DoF), periodic x/y, Neumann walls z. Train on the first steps,
    then A/B on the rest: FP32 V-cycle with the learned / extrapolated guess against the plain
    FP64 method (FP64 V-cycle, extrapolated or zero guess)."""
    from exadg_amd import pressure_sequence as ps

    def log(msg):
        print(f"[combined] {msg}", file=sys.stderr, flush=True)

    k, nc = 5, args.combined_cells
    t0 = time.perf_counter()
    op, prec32 = ps.setup_box(ex, nc, k, ex.FP32, deformation=0.1, mapping_degree=k)
    _, prec64 = ps.setup_box(ex, nc, k, ex.FP64, op=op)
    t_setup = time.perf_counter() - t0
    modes = ex.fourier_modes(2026, 1, 8)
    log(f"{op.n} DoF, setup {t_setup:.1f} s; generating {args.combined_steps} steps")
    t0 = time.perf_counter()
    P, Bv, its, _ = ps.generate(ex, torch, op, prec32, modes, args.combined_steps, log=log)
    t_gen = time.perf_counter() - t0
    net = ex.GuessNet(k, history=3, width=128, seed=1)
    log("training")
    # the network starts as the extrapolation (zero output layer) and is reset to it if the
    # trained parameters do not beat the extrapolation on the training steps
    losses, t_train = ps.train(ex, net, torch, op, P, Bv, args.combined_train, args.combined_epochs,
                               lr=1e-3, loss=ex.GUESS_LOSS_RESIDUAL, from_extrapolation=True)
    kept = net.evaluate(P, Bv, 0, args.combined_train, op=op, loss=ex.GUESS_LOSS_RESIDUAL)
    log(f"train loss {losses[0]:.4g} -> {losses[-1]:.4g}, kept {kept:.4g}; A/B")
    s0, s1 = args.combined_train, args.combined_steps
    ab32 = ps.ab_test(ex, torch, op, prec32, net, P, Bv, s0, s1, arms=("extrapolation", "learned"))
    ab64 = ps.ab_test(ex, torch, op, prec64, None, P, Bv, s0, s1, arms=("zero", "extrapolation"))
    n_ab = s1 - s0
    per_step = lambda r: round(r["total_time_s"] / n_ab * 1e3, 2)
    return {
        "workload": f"pressure sequence on the curved mesh: [0,1]^3 deformed by 0.1 prod sin(pi x), "
                    f"MappingQ(5), periodic x/y, Neumann walls z, {nc}^3 cells, degree 5, {op.n} DoF, "
                    f"dt 0.01, {len(modes)} modes |n| <= 8",
        "steps": args.combined_steps, "train_steps": s0, "setup_s": round(t_setup, 1),
        "generation": {"time_s": round(t_gen, 2), "mean_iterations": float(np.mean(its[3:]))},
        "training": {"epochs": args.combined_epochs, "time_s": round(t_train, 2),
                     "loss_first_epoch": losses[0], "loss_last_epoch": losses[-1],
                     "loss_kept_parameters": kept,
                     "loss": "residual |A e|^2 of the guess per training step, divided by the "
                             "extrapolation's; training starts from the extrapolation (zero "
                             "output layer) and falls back to it unless the loss is below 1"},
        "ms_per_step": {"fp32_mg_learned": per_step(ab32["learned"]),
                        "fp32_mg_extrapolation": per_step(ab32["extrapolation"]),
                        "fp64_mg_extrapolation_plain": per_step(ab64["extrapolation"]),
                        "fp64_mg_zero_guess": per_step(ab64["zero"])},
        "fp32_mg": ab32, "fp64_mg": ab64,
        "speedup_learned_fp32mg_vs_plain_fp64mg_extrapolated": round(
            ab64["extrapolation"]["total_time_s"] / ab32["learned"]["total_time_s"], 3),
        "speedup_learned_fp32mg_vs_fp64mg_zero_guess": round(
            ab64["zero"]["total_time_s"] / ab32["learned"]["total_time_s"], 3),
    }


def measure_stream(torch, x, y, reps=20):
    """The box's practical HBM bandwidth on vectors of the benchmark's size, next to the vendor
    peak: a device-to-device copy (2 vectors moved) and a triad a = b + s c (3 vectors),
    medians of HIP-event timings."""
    z = torch.empty_like(x)
    stream = torch.cuda.current_stream()

    def median_ms(f):
        f()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(reps)]
        for a, b in ev:
            a.record(stream)
            f()
            b.record(stream)
        torch.cuda.synchronize()
        return statistics.median(a.elapsed_time(b) for a, b in ev)

    nbytes = x.numel() * x.element_size()
    t_copy = median_ms(lambda: y.copy_(x))
    t_triad = median_ms(lambda: torch.add(x, y, alpha=0.5, out=z))
    del z
    return {"copy_GBps": round(2 * nbytes / t_copy / 1e6, 1),
            "triad_GBps": round(3 * nbytes / t_triad / 1e6, 1),
            "bytes_per_vector": nbytes}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--degree", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true",
                    help="skip the mixed-precision solve (curved k=5 48^3, FP32 vs FP64 V-cycle)")
    ap.add_argument("--solve-small", action="store_true",
                    help="also time the solve on the Cartesian k=3 8^3 box (launch-bound)")
    ap.add_argument("--sweep", action="store_true",
                    help="also time the curved-mesh vmult at degrees 2, 3, 5, 7 (~24M DoF each)")
    ap.add_argument("--sweep-fp32", action="store_true",
                    help="with --sweep: the same sweep in FP32 as well (curved_sweep_fp32)")
    ap.add_argument("--guess", action="store_true",
                    help="also run the learned-initial-guess pressure sequence (32^3, k=4)")
    ap.add_argument("--guess-cells", type=int, default=32)
    ap.add_argument("--guess-degree", type=int, default=4)
    ap.add_argument("--guess-steps", type=int, default=400)
    ap.add_argument("--guess-train", type=int, default=300)
    ap.add_argument("--guess-epochs", type=int, default=20)
    ap.add_argument("--guess-kmax", type=int, default=8)
    ap.add_argument("--guess-lr", type=float, default=1e-3)
    ap.add_argument("--no-combined", action="store_true",
                    help="skip the combined experiment (curved k=5 pressure sequence, 50^3)")
    ap.add_argument("--combined-cells", type=int, default=50)
    # the default run is a bounded sequence (40 steps: train on 30, compare on 10); the full
    # sequence is --combined-steps 200 --combined-train 150
    ap.add_argument("--combined-steps", type=int, default=40)
    ap.add_argument("--combined-train", type=int, default=30)
    ap.add_argument("--combined-epochs", type=int, default=10)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run plain device copies of the same vector (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import exadg_amd
    torch.cuda.set_device(0)
    mesh = exadg_amd.mesh_desc(args.cells)
    op = exadg_amd.LaplaceOperator(mesh, exadg_amd.bc_desc(), args.degree, exadg_amd.FP64)
    n = op.n
    # seeded uniform[-1,1] input (SplitMix64 on the DoF index), generated in HBM
    x = exadg_amd.random_vector(n)
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        op.vmult(y, x)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        op.vmult(y, x)
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s.elapsed_time(e) for s, e in ev]
    if args.copy_probe:
        for _ in range(args.steps):
            y.copy_(x)
        torch.cuda.synchronize()
    stream_bw = measure_stream(torch, x, y)
    ms_step = wall * 1e3 / args.steps
    # achieved bandwidth: algorithmic bytes over the average kernel duration (HIP events on the
    # launch stream, each bracketing one launch)
    t_avg = statistics.mean(kernel_ms) * 1e-3
    ab = algorithmic_bytes(n, op.n_cells)
    achieved = ab / t_avg / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        try:
            traffic = json.load(open(PROFILE_TRAFFIC)).get("bytes_per_launch")
        except Exception:
            traffic = None
    out = {
        "metric": METRIC,
        "value": round(n / (ms_step * 1e-3) / 1e9, 3),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"SIPG vmult, affine Cartesian [0,1]^3, {args.cells}^3 cells, "
                               f"degree {args.degree}, FP64, homogeneous Dirichlet",
                   "n_dofs": n, "n_cells": op.n_cells, "degree": args.degree,
                   "kernel": "sipg_cartesian_kernel",
                   "kernel_ms_mean": round(statistics.mean(kernel_ms), 5),
                   "kernel_ms_median": round(statistics.median(kernel_ms), 5),
                   "kernel_ms_min": round(min(kernel_ms), 5)},
        # the roofline of a streaming operator is HBM; what holds this kernel below it is the
        # latency of each workgroup's chain (loads, LDS staging, five dependent LDS phases,
        # store), overlapped only partly across the resident workgroups: the VALU is busy ~40%
        # of the time, half of each wave's life is spent in waits, and probes removing VALU
        # work, barriers or LDS traffic each gain only a few percent
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "limiter": "latency of the per-workgroup load -> LDS phases -> store chain, "
                                "not HBM bytes (traffic ~1.2x algorithmic) nor VALU/LDS "
                                "throughput"},
        "stream": stream_bw,
    }
    if not args.no_solve:
        from exadg_amd import solve_bench
        x = y = None
        torch.cuda.empty_cache()
        log("solve: curved k=5 48^3, FP64 vs FP32 V-cycle")
        r = solve_bench.compare_arms(
            exadg_amd, torch, exadg_amd.mesh_desc(48, deformation=0.1, mapping_degree=5),
            exadg_amd.bc_desc(), 5, [5, 3, 2, 1], 3.141592653589793, exadg_amd.COARSE_CHEBYSHEV)
        r["workload"] = ("time to ||r|| <= 1e-10 ||b|| (outer FP64 CG, zero guess, sine RHS), "
                         "curved [0,1]^3 (x += 0.1 prod sin(pi x)), MappingQ(5), 48^3 cells, "
                         "degree 5, V-cycle DG5->3->2->1->CG-Q1 (Chebyshev coarse), FP32 vs FP64")
        out["solve"] = r
        if args.solve_small:
            out["solve_cart_k3_8c"] = solve_bench.compare_arms(
                exadg_amd, torch, exadg_amd.mesh_desc(8), exadg_amd.bc_desc(), 3, [3, 1],
                3.141592653589793, exadg_amd.COARSE_CG, reps=3)
    if args.sweep:
        x = y = None
        torch.cuda.empty_cache()
        out["curved_sweep"] = run_sweep(exadg_amd, torch, args)
        if args.sweep_fp32:
            out["curved_sweep_fp32"] = run_sweep(exadg_amd, torch, args, fp32=True)
    if args.guess:
        x = y = None
        torch.cuda.empty_cache()
        out["learned_guess"] = run_guess(exadg_amd, torch, args)
    if not args.no_combined:
        x = y = None
        torch.cuda.empty_cache()
        out["combined"] = run_combined(exadg_amd, torch, args)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(args.degree, args.cells)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
