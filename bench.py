"""Headline benchmark: Krylov solve time and HBM GB/s of the 2.1M-cell SPD
pressure-correction system (128^3 hex mesh, cells randomly permuted from seed
1, lognormal(0, 0.5) face densities, b ~ U[-1,1] from seed 2, x0 = 0), CG to
1e-8 relative residual, FP32 inner iterations with FP64 refinement (MIXED)
next to the plain FP64 solve.

One "step" is one complete solve from x0 = 0 with the matrix, b and x0
already resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(a, b, iters_full, budget_s=12.0):
    """Oracle CG (the reference's loop structure, its compile flags) on the host
    cores: time a bounded number of iterations on the same system and scale to
    the iteration count of a full solve."""
    import oracle

    threads = max(1, min(16, os.cpu_count() or 1))
    oracle.set_threads(threads)
    t0 = time.perf_counter()
    oracle.cg_tol(a, b, tol=0.0, maxit=5, fast=True)
    per_it = (time.perf_counter() - t0) / 5.0
    k = int(max(10, min(2000, budget_s / max(per_it, 1e-6))))
    t0 = time.perf_counter()
    oracle.cg_tol(a, b, tol=0.0, maxit=k, fast=True)
    per_it = (time.perf_counter() - t0) / k
    return {
        "value": per_it * iters_full * 1e3,
        "unit": "ms",
        "cores": threads,
        "kind": "port",
        "sample": f"{k} FP64 CG iterations of the same 128^3 system on {threads} host threads "
                  f"({per_it * 1e3:.2f} ms/iteration), scaled to the {iters_full} iterations of a full solve",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the solve is single-device; see the sharding note in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    from opencslsolver_amd import gen
    from opencslsolver_amd.solver import Solver

    n = args.n
    _log(f"bench: generating {n}^3 system")
    a, b, _ = gen.poisson3d(n, seed_mesh=1, seed_rhs=2, sigma=0.5, leak=1e-6)
    dev = torch.device("cuda:0")
    val_d = torch.from_numpy(a.val).to(dev)
    b_d = torch.from_numpy(b).to(dev)
    zero_d = torch.zeros(a.n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    s = Solver(a.row_ptr, a.col)
    s.set_values(val_d)
    s.set_rhs(b_d)

    def solve(prec):
        s.set_x(zero_d)
        return s.solve("cg", "tolerance", prec, tol=1e-8, max_iter=20000)

    _log("bench: warmup")
    for _ in range(args.warmup):
        solve("fp64")
        solve("mixed")

    def timed(prec):
        reps = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            reps.append(solve(prec))
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps, reps

    _log("bench: timing FP64")
    ms64, reps64 = timed("fp64")
    x64 = s.get_x()
    _log("bench: timing MIXED")
    msm, repsm = timed("mixed")
    xm = s.get_x()
    rel64 = float(np.linalg.norm(b - a.matvec(x64)) / np.linalg.norm(b))
    relm = float(np.linalg.norm(b - a.matvec(xm)) / np.linalg.norm(b))
    err = float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))

    # dominant kernel: the fused CG SpMV (q = A p, p.q), timed with HIP events
    # on the handle's stream
    spmv_ms_m, spmv_bytes_m = s.time_cg_spmv("mixed", reps=50)
    spmv_ms_64, spmv_bytes_64 = s.time_cg_spmv("fp64", reps=50)
    ach_m = spmv_bytes_m / (spmv_ms_m * 1e-3) / 1e9
    ach_64 = spmv_bytes_64 / (spmv_ms_64 * 1e-3) / 1e9

    rm, r64 = repsm[-1], reps64[-1]
    gbs_m = rm["bytes"] / (rm["solve_ms"] * 1e-3) / 1e9
    gbs_64 = r64["bytes"] / (r64["solve_ms"] * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        _log("bench: CPU baseline")
        cpu = cpu_baseline(a, b, r64["iterations"])

    out = {
        "metric": "Krylov solve ms + HBM GB/s at 2.1M cells, FP32-refined vs FP64, to 1e-8 rel res",
        "value": round(msm, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(msm, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded generator, regenerated on the box)",
        "config": {"workload": f"SPD pressure-correction CG, {n}^3 hex mesh permuted (seed 1), "
                               f"N={a.n}, nnz={a.nnz}, tol=1e-8, x0=0",
                   "n_rows": a.n, "nnz": a.nnz},
        "mixed": {"ms": round(msm, 3), "device_ms": round(rm["solve_ms"], 3),
                  "iterations": rm["iterations"], "replacements": rm["replacements"],
                  "algorithmic_GBps": round(gbs_m, 1), "fp64_relres": relm,
                  "error_vs_fp64": err},
        "fp64": {"ms": round(ms64, 3), "device_ms": round(r64["solve_ms"], 3),
                 "iterations": r64["iterations"], "algorithmic_GBps": round(gbs_64, 1),
                 "fp64_relres": rel64},
        "speedup_mixed_vs_fp64": round(ms64 / msm, 3),
        "solve_GBps": round(gbs_m, 1),
        "roofline": {"bound": "hbm", "kernel": "k_cg_spmv<float,float>",
                     "achieved": round(ach_m, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach_m / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_launch_us": round(spmv_ms_m * 1e3, 2),
                     "bytes_per_launch": spmv_bytes_m,
                     "fp64_kernel": {"achieved": round(ach_64, 1), "avg_launch_us": round(spmv_ms_64 * 1e3, 2),
                                     "bytes_per_launch": spmv_bytes_64}},
        "cpu_baseline": cpu,
    }
    s.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
