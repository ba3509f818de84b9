"""Headline benchmark: Krylov solve time and HBM GB/s of the 2.1M-cell SPD
pressure-correction system (128^3 hex mesh, cells randomly permuted from seed
1, lognormal(0, 0.5) face densities, b ~ U[-1,1] from seed 2, x0 = 0), CG to
1e-8 relative residual, FP32 inner iterations with FP64 refinement (MIXED)
next to the plain FP64 solve.

One "step" is one complete solve from x0 = 0 with the matrix, b and x0
already resident in HBM.  The JSON line also carries the side workloads:

* multi_rhs -- 96^3 Taylor-Green convection-diffusion system (seed 3), three
  right-hand sides (seeds 4-6) sharing A, ILU(0)-BiCGSTAB to 1e-10, MIXED and
  FP64, with the error of each MIXED solution against the FP64 one;
* batch -- 64 independent pressure systems of n^3 cells, n log-uniform in
  [22, 80] from seed 7, MIXED CG to 1e-8 in shared launches: total time,
  per-system latency p50/p99, and the same systems solved one by one;
* adjoint -- forward CG solve plus the adjoint (transpose CG solve and the
  sampled outer product dL/dA on the CSR pattern) on the 128^3 system;
* assembly -- the GPU coefficient assembly of each equation on a 128^3 hex
  mesh, beside the oracle's host restatement of the reference loops;
* tri -- exact triangular solves of ILU(0) factors in each of their forms;
* flow -- SIMPLE iterations of the flow solver on a 64^3 hex mesh, support
  loops on the device (state resident) and on the host, and on a 32^3 SU2
  box beside the reference's own run of the same file;
* sod -- the 1000-cell shock tube, 200 steps, beside the reference's own run;
* structure -- the CSR pattern of the 128^3 mesh built on the device;
* mesh -- SU2 reading, connectivity and geometry of a 64^3 box on the host.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--sections all]
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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads():
    """Host threads for the CPU baselines: the CPUs this process may run on,
    at most 16 (the host CPU share that comes with one GPU on this machine;
    more threads than that share only time-slice against each other)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_model():
    """The host CPU's model name and logical CPU count, from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            names = [ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")]
        return f"{names[0]} ({len(names)} logical CPUs)" if names else None
    except OSError:
        return None


def _ref_cg_child(n, threads, calls):
    """Child-process leg of the CPU baseline: the reference's own CG_solve
    (built from its sources for the build host's instruction set, so it runs
    apart from this process) on the same seeded system; prints JSON."""
    import oracle
    from opencslsolver_amd import gen

    a, b, _ = gen.poisson3d(n, seed_mesh=1, seed_rhs=2, sigma=0.5, leak=1e-6)
    call, setup = oracle.ref_cg_timing(a, b, np.zeros(a.n), threads, calls)
    print(json.dumps({"call_s": call, "setup_s": setup}))


def cpu_baseline(a, b, n, iters_full, budget_s=12.0):
    """The headline solve on the host cores, two ways.

    reference: the reference's own CG_solve (cmatrix_solver.cpp:635-680, built
    from its sources in oracle/_ref with its own flags and OpenMP), which stops
    after 4 iterations by design: the median call and its set-up are timed and
    one iteration = (call - set-up) / 4; its iteration computes A p twice and
    five dot products.  port: the oracle's textbook FP64 CG (one A p per
    iteration) with the reference's flags, the median of 5 chunks of
    iterations.  Both per-iteration costs are scaled to the iterations of the
    full FP64 solve: a full solve is not run, and the extrapolation is stated
    in the line."""
    import subprocess

    import oracle

    threads = host_threads()
    oracle.set_threads(threads)
    t0 = time.perf_counter()
    oracle.cg_tol(a, b, tol=0.0, maxit=5, fast=True)
    per_it = (time.perf_counter() - t0) / 5.0
    k = int(max(10, min(400, budget_s / 5.0 / max(per_it, 1e-6))))
    chunks = []
    for _ in range(5):
        t0 = time.perf_counter()
        oracle.cg_tol(a, b, tol=0.0, maxit=k, fast=True)
        chunks.append((time.perf_counter() - t0) / k)
    port_it = float(np.median(chunks))
    port = {"value": round(port_it * iters_full * 1e3, 1), "ms_per_iteration": round(port_it * 1e3, 3),
            "sample": f"median of 5 chunks of {k} FP64 CG iterations of the same system"}
    ref_it = None
    if oracle.ref_available():
        code = "import sys; sys.path.insert(0, sys.argv[1]); import bench; bench._ref_cg_child(*map(int, sys.argv[2:]))"
        r = subprocess.run([sys.executable, "-c", code, ROOT, str(n), str(threads), "7"],
                           capture_output=True, text=True, timeout=600)
        if r.returncode == 0:
            d = json.loads(r.stdout.strip().splitlines()[-1])
            ref_it = (d["call_s"] - d["setup_s"]) / 4.0
            ref_call = d
        else:
            _log(f"bench: reference CG unavailable here (exit {r.returncode}): {r.stderr[-300:]}")
    common = {"unit": "ms", "cores": threads, "cpu_model": cpu_model(),
              "iterations_scaled_to": iters_full}
    if ref_it is None:
        return {"value": port["value"], "kind": "port", **common,
                "sample": f"{port['sample']} on {threads} host threads ({port['ms_per_iteration']} ms/iteration), "
                          f"scaled to the {iters_full} iterations of the full FP64 solve"}
    return {"value": round(ref_it * iters_full * 1e3, 1), "kind": "reference", **common,
            "sample": f"the reference's CG_solve on {threads} OpenMP threads, median of 7 calls of its fixed 4 "
                      f"iterations ({ref_call['call_s'] * 1e3:.2f} ms per call, {ref_call['setup_s'] * 1e3:.2f} ms "
                      f"of it set-up): {ref_it * 1e3:.3f} ms/iteration, scaled to the {iters_full} iterations "
                      f"of the full FP64 solve",
            "ms_per_iteration": round(ref_it * 1e3, 3),
            "port": port, "reference_over_port_per_iteration": round(ref_it / port_it, 3)}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel factors)."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
        return d[kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def measured_traffic_like(fragment):
    """HBM bytes per launch of the first committed PMC entry whose name holds
    `fragment`, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    for k, v in d.items():
        if fragment in k:
            return v.get("bytes_per_launch")
    return None


def headline(args, out):
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

    # launch autotuning (setup, like building the layouts): 64 iterations of
    # every candidate launch per precision; the handle keeps the fastest
    _log("bench: autotuning the CG launch")
    from opencslsolver_amd import tuning

    tuned = {prec: tuning.tuned_cg(s, prec, iterations=64, retune=not args.reuse_tuning)
             for prec in ("fp64", "mixed")}

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

    # dominant kernel, timed with HIP events on the handle's stream: with one
    # kernel per iteration, the odd iteration kernel (the even one, which
    # leaves x alone, is reported beside it); with the split form, the CG
    # SpMV (q = A p and its dot products)
    def dominant(prec):
        t = tuned[prec]
        rpl = t["rows_per_lane"]
        ty = "float, float" if prec == "mixed" else "double, double"
        if t["cg_kernels"] == 1:
            ms, by = s.time_cg_iteration(prec, pairs=50)
            return {"kernel": f"k_cgf_iter<{ty}, {rpl}, 1>", "ms": ms[1], "bytes": by[1],
                    "even": {"kernel": f"k_cgf_iter<{ty}, {rpl}, 0>", "avg_launch_us": round(ms[0] * 1e3, 2),
                             "bytes_per_launch": by[0],
                             "achieved": round(by[0] / (ms[0] * 1e-3) / 1e9, 1)}}
        ms, by = s.time_cg_spmv(prec, reps=50)
        name = f"k_cgm_spmv<{ty}, {rpl}>" if prec == "mixed" else f"k_cg_spmv<{ty}, {rpl}>"
        return {"kernel": name, "ms": ms, "bytes": by}

    dom_m, dom_64 = dominant("mixed"), dominant("fp64")
    ach_m = dom_m["bytes"] / (dom_m["ms"] * 1e-3) / 1e9
    ach_64 = dom_64["bytes"] / (dom_64["ms"] * 1e-3) / 1e9

    rm, r64 = repsm[-1], reps64[-1]
    gbs_m = rm["bytes"] / (rm["solve_ms"] * 1e-3) / 1e9
    gbs_64 = r64["bytes"] / (r64["solve_ms"] * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        _log("bench: CPU baseline")
        cpu = cpu_baseline(a, b, n, r64["iterations"])

    kname, kname64 = dom_m["kernel"], dom_64["kernel"]
    triad = stream_triad()
    out.update({
        "value": round(msm, 3),
        "ms_per_step": round(msm, 3),
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
        "tuned_launch": tuned,
        "solve_GBps": round(gbs_m, 1),
        "roofline": {"bound": "hbm", "kernel": kname,
                     "achieved": round(ach_m, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach_m / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kname),
                     "avg_launch_us": round(dom_m["ms"] * 1e3, 2),
                     "bytes_per_launch": dom_m["bytes"],
                     **({"even_kernel": dom_m["even"]} if "even" in dom_m else {}),
                     "fp64_kernel": {"kernel": kname64, "achieved": round(ach_64, 1),
                                     "avg_launch_us": round(dom_64["ms"] * 1e3, 2),
                                     "bytes_per_launch": dom_64["bytes"],
                                     "traffic": measured_traffic(kname64)},
                     "stream_triad": triad,
                     "frac_of_triad": round(ach_m / triad["hbm_GBps"], 4)},
        "cpu_baseline": cpu,
    })

    if "adjoint" in args.sections:
        _log("bench: adjoint")
        xt_d = torch.from_numpy(gen.normal(9, a.n)).to(dev)

        def fwd_adj():
            s.set_x(zero_d)
            rf = s.solve("cg", "tolerance", "fp64", tol=1e-8)
            x = torch.empty(a.n, dtype=torch.float64, device=dev)
            s.get_x(out=x)
            _, _, ra = s.adjoint(x - xt_d, "cg", "fp64", tol=1e-8)
            return rf, ra

        fwd_adj()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rf, ra = fwd_adj()
        torch.cuda.synchronize()
        ms_pair = (time.perf_counter() - t0) * 1e3 / args.steps
        o_ms, o_bytes = s.time_sampled_outer(reps=50)
        out["adjoint"] = {"workload": f"forward + adjoint FP64 CG to 1e-8 on the {n}^3 system, "
                                      "L = 0.5||x - x_t||^2, x_t ~ N(0,1) seed 9",
                          "ms_forward_plus_adjoint": round(ms_pair, 3),
                          "forward_device_ms": round(rf["solve_ms"], 3),
                          "transpose_solve_device_ms": round(ra["solve_ms"], 3),
                          "transpose_iterations": ra["iterations"],
                          "outer_product": {"kernel": "k_sampled_outer<false>", "avg_launch_us": round(o_ms * 1e3, 2),
                                            "bytes_per_launch": o_bytes,
                                            "achieved_GBps": round(o_bytes / (o_ms * 1e-3) / 1e9, 1)}}
    s.close()


def multi_rhs(args, out):
    import torch

    from opencslsolver_amd import gen
    from opencslsolver_amd.solver import Solver

    _log("bench: 3-RHS convection-diffusion")
    a, _ = gen.convdiff3d(96, seed_perm=3)
    B = np.stack([gen.uniform(sd, a.n) for sd in (4, 5, 6)])
    dev = torch.device("cuda:0")
    B_d = torch.from_numpy(B).to(dev)
    with Solver(a.row_ptr, a.col) as s:
        s.set_values(torch.from_numpy(a.val).to(dev))
        res = {}
        # input order: the factor of the caller's numbering (ILU_decompose's);
        # colour order: the renumbered matrix, one level per colour
        for order, tag in ((0, ""), (1, "_colour")):
            for prec in ("fp64", "mixed"):
                s.solve_multi(B_d, precision=prec, tol=1e-10, bicgstab_order=order)  # warm (setup, graphs)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                X, reps = s.solve_multi(B_d, precision=prec, tol=1e-10, bicgstab_order=order)
                torch.cuda.synchronize()
                res[prec + tag] = ((time.perf_counter() - t0) * 1e3, X.cpu().numpy(), reps)
        n_colours = s.colour_order()[1]
    X64, Xm = res["fp64"][1], res["mixed"][1]
    out["multi_rhs"] = {
        "workload": f"ILU(0)-BiCGSTAB, 96^3 Taylor-Green convection-diffusion (seed 3), N={a.n}, nnz={a.nnz}, "
                    "3 RHS (seeds 4-6), tol 1e-10",
        "mixed_ms": round(res["mixed"][0], 3), "fp64_ms": round(res["fp64"][0], 3),
        "mixed_iterations": [r["iterations"] for r in res["mixed"][2]],
        "fp64_iterations": [r["iterations"] for r in res["fp64"][2]],
        "mixed_fp64_relres": [float(np.linalg.norm(B[k] - a.matvec(Xm[k])) / np.linalg.norm(B[k]))
                              for k in range(3)],
        "error_vs_fp64": [float(np.linalg.norm(Xm[k] - X64[k]) / np.linalg.norm(X64[k])) for k in range(3)],
        "colour_order": {
            "colours": n_colours,
            "mixed_ms": round(res["mixed_colour"][0], 3), "fp64_ms": round(res["fp64_colour"][0], 3),
            "mixed_iterations": [r["iterations"] for r in res["mixed_colour"][2]],
            "fp64_iterations": [r["iterations"] for r in res["fp64_colour"][2]],
            "mixed_fp64_relres": [float(np.linalg.norm(B[k] - a.matvec(res["mixed_colour"][1][k]))
                                        / np.linalg.norm(B[k])) for k in range(3)],
            "mixed_error_vs_input_order_fp64": [float(np.linalg.norm(res["mixed_colour"][1][k] - X64[k])
                                                      / np.linalg.norm(X64[k])) for k in range(3)],
        },
    }


def batch(args, out):
    from opencslsolver_amd import gen
    from opencslsolver_amd.solver import Batch, Solver

    _log("bench: batch of 64 systems")
    ns, sm, sr = gen.batch_plan(7, 64, 22, 80)
    systems = [gen.poisson3d(int(ns[k]), seed_mesh=int(sm[k]), seed_rhs=int(sr[k]))[:2] for k in range(64)]
    cells = int(sum(a.n for a, _ in systems))
    with Batch([(a.row_ptr, a.col) for a, _ in systems]) as bt:
        bt.set_values([a.val for a, _ in systems])
        bt.set_rhs([b for _, b in systems])
        # launch shape (rows per lane x workgroup size) from the autotuner
        shape, us_it, table = bt.tune("mixed", iterations=32)
        bt.solve("mixed", tol=1e-8)  # warm (graph capture)
        # solutions in host memory: the call copies 64 guesses in and 64
        # solutions out over PCIe
        t0 = time.perf_counter()
        xs, reps, lat = bt.solve("mixed", tol=1e-8)
        total_host_io = (time.perf_counter() - t0) * 1e3
        # guesses and solutions resident in device memory (torch tensors)
        import torch
        xd = [torch.zeros(a.n, dtype=torch.float64, device="cuda") for a, _ in systems]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, reps_d, lat = bt.solve("mixed", x0=xd, tol=1e-8)
        torch.cuda.synchronize()
        total = (time.perf_counter() - t0) * 1e3
        dev_match = all(np.array_equal(x, t.cpu().numpy()) for x, t in zip(xs, xd))
        del xd
    rel = [float(np.linalg.norm(b - a.matvec(x)) / np.linalg.norm(b)) for (a, b), x in zip(systems, xs)]
    seq = 0.0
    for a, b in ([] if args.skip_batch_sequential else systems):
        with Solver(a.row_ptr, a.col) as s:
            s.set_values(a.val)
            s.set_rhs(b)
            s.set_x(np.zeros(a.n))
            s.solve("cg", "tolerance", "mixed", tol=1e-8)
            s.set_x(np.zeros(a.n))
            seq += s.solve("cg", "tolerance", "mixed", tol=1e-8)["solve_ms"]
    out["batch"] = {
        "workload": f"64 SPD pressure systems, n^3 cells with n log-uniform in [22, 80] (seed 7), "
                    f"{cells} cells in total, MIXED CG to 1e-8",
        "total_ms": round(total, 3), "device_ms": round(reps_d[0]["solve_ms"], 3),
        "total_ms_host_io": round(total_host_io, 3), "device_resident_equals_host_io": dev_match,
        "latency_p50_ms": round(float(np.percentile(lat, 50)), 3),
        "latency_p99_ms": round(float(np.percentile(lat, 99)), 3),
        "per_system_ms": [round(float(v), 3) for v in lat],
        "iterations": [r["iterations"] for r in reps],
        "max_fp64_relres": max(rel),
        "sequential_sum_ms": None if args.skip_batch_sequential else round(seq, 3),
        # algorithmic bytes of every system's iterations and replacements
        # (the single-system split-CG count with x updated every iteration)
        # over the wall time of the iteration loop
        "algorithmic_GB": round(sum(r["bytes"] for r in reps) / 1e9, 3),
        "algorithmic_GBps": round(sum(r["bytes"] for r in reps) / (reps[0]["solve_ms"] * 1e-3) / 1e9, 1),
        "launch": dict(shape, us_per_iteration=round(us_it, 2)),
        "tuning": [[t["rows_per_lane"], t["block_size"], round(t["us_per_iteration"], 2)] for t in table],
    }


def profile_kernels(args):
    """Launch each measured kernel alone (for rocprofv3 kernel traces and PMC
    passes): a 256 MiB device copy (the counters' calibration), the CG SpMV
    and the one-kernel CG iteration (even and odd) in MIXED and FP64, and the
    sampled outer product, on the headline system."""
    import torch

    from opencslsolver_amd import gen
    from opencslsolver_amd.solver import Solver

    dev = torch.device("cuda:0")
    src = torch.ones(32 << 20, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    a, b, _ = gen.poisson3d(args.n, seed_mesh=1, seed_rhs=2, sigma=0.5, leak=1e-6)
    res = {"copy_bytes": 2 * src.numel() * 8}
    from opencslsolver_amd import tuning

    with Solver(a.row_ptr, a.col) as s:
        s.set_values(a.val)
        s.set_rhs(b)
        for prec in ("mixed", "fp64"):  # the launch the headline run tuned (cached), else tuned here
            res[f"tuned_{prec}"] = tuning.tuned_cg(s, prec, iterations=64)
        for prec in ("mixed", "fp64"):
            ms, by = s.time_cg_spmv(prec, reps=args.steps)
            res[f"cg_spmv_{prec}"] = {"avg_us": ms * 1e3, "bytes": by, "GBps": by / (ms * 1e-3) / 1e9}
            for rpl in (1, 2, 4):  # every layout, so that whichever a later tuning picks has a traffic entry
                ms1, by1 = s.time_cg_spmv(prec, reps=args.steps, rows_per_lane=rpl)
                res[f"cg_spmv_{prec}_r{rpl}"] = {"avg_us": ms1 * 1e3, "bytes": by1,
                                                 "GBps": by1 / (ms1 * 1e-3) / 1e9}
            ms2, by2 = s.time_cg_iteration(prec, pairs=args.steps)
            for par, name in ((0, "even"), (1, "odd")):
                res[f"cg_iter_{prec}_{name}"] = {"avg_us": ms2[par] * 1e3, "bytes": by2[par],
                                                 "GBps": by2[par] / (ms2[par] * 1e-3) / 1e9}
        ms, by = s.time_sampled_outer(reps=args.steps)
        res["sampled_outer"] = {"avg_us": ms * 1e3, "bytes": by, "GBps": by / (ms * 1e-3) / 1e9}
    print(json.dumps(res), flush=True)


def assembly(args, out):
    """Coefficient assembly (the face, boundary-face and cell loops of each
    equation) on an n^3 hex mesh with the headline's permuted cell numbering,
    every boundary kind on its sides, viscous, blended, transient; a seeded
    random flow state.  GPU kernel time per equation (HIP events) beside the
    oracle's OpenMP scatter loops (the reference's structure) on the host."""
    import oracle
    from opencslsolver_amd import flow, gen

    n = args.n
    _log(f"bench: assembly on a {n}^3 hex mesh")
    _, _, perm = gen.poisson3d(n, seed_mesh=1, seed_rhs=2)
    mesh, rp, col = gen.hex_mesh(n, perm=perm)
    bcs = [dict(type=flow.VELOCITY_INLET, velocity=(3.0, 0.5, -0.2), temperature=300.0),
           dict(type=flow.PRESSURE_OUTLET, pressure=1e5), dict(type=flow.WALL, no_slip=1),
           dict(type=flow.WALL, no_slip=0), dict(type=flow.SYMMETRY),
           dict(type=flow.SUPERSONIC_INLET, velocity=(400.0, 0.0, 0.0), pressure=1e5, temperature=300.0)]
    prm = flow.default_params(viscous=1, blended=1, transient=1, bcs=bcs)
    N, F, B = mesh["cell_volume"].shape[0], mesh["face_left"].shape[0], mesh["bc_left"].shape[0]
    rng = np.random.default_rng(21)
    st = dict(u=rng.normal(size=(N, 3)), u_old=rng.normal(size=(N, 3)), rho=1.0 + rng.random(N),
              rho_old=1.0 + rng.random(N), p=1e5 + 1e3 * rng.random(N), p_old=1e5 + 1e3 * rng.random(N),
              p_corr=rng.normal(size=N), grad_p=rng.normal(size=(N, 3)), T=280.0 + 40.0 * rng.random(N),
              h=3e5 + 1e4 * rng.random(N), h_old=3e5 + 1e4 * rng.random(N), dt=np.full(N, 1e-4),
              ap=np.zeros(N), ap_vec=np.zeros((N, 3)), f_mdot=rng.normal(size=F) * 1e-3,
              f_rho=1.0 + rng.random(F), f_qy=rng.random(F) * 1e-3, f_stress=rng.normal(size=(F, 3)) * 1e-5,
              b_T=280.0 + 40.0 * rng.random(B), b_p=np.full(B, 1e5), b_rho=1.0 + rng.random(B),
              b_mdot=rng.normal(size=B) * 1e-3, b_h=np.full(B, 3e5), b_qy=rng.random(B) * 1e-3,
              b_u=rng.normal(size=(B, 3)), b_supersonic=np.zeros(B, np.int32))
    names = ["momentum_x", "momentum_y", "momentum_z", "pressure", "enthalpy"]

    order_names = {flow.CELL_ORDER_CURVE: "Hilbert curve through the cell centres",
                   flow.CELL_ORDER_LOCALITY: "reverse Cuthill-McKee"}

    def timed(m, r, c, label, cell_order=flow.CELL_ORDER_AUTO):
        _log(f"bench: assembly, {label}")
        res = {}
        prm.cell_order = cell_order
        with flow.Flow(m, r, c, prm) as fl:
            order = fl.cell_order()
            res["device_cell_order"] = ("caller" if np.array_equal(order, np.arange(order.size))
                                        else order_names.get(cell_order, "Hilbert curve through the cell centres"))
            fl.set_state(**st)
            for eq, name in enumerate(names):
                ms, by = fl.time_assemble(eq, reps=10)
                res[name] = {"avg_us": round(ms * 1e3, 2), "bytes_per_launch": by,
                             "achieved_GBps": round(by / (ms * 1e-3) / 1e9, 1),
                             "frac": round(by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                if n == 128:  # the committed PMC passes were taken on the 128^3 mesh
                    tr = measured_traffic_like(f" {name} 128^3 {label}")
                    if tr:
                        res[name]["traffic"] = tr
                        res[name]["traffic_over_algorithmic"] = round(tr / by, 2)
        return res

    # AUTO keeps the permuted mesh's cells along the Hilbert curve
    res = timed(mesh, rp, col, "permuted cells, device curve order")
    # the same mesh with the cells kept in reverse Cuthill-McKee order
    res_rcm = timed(mesh, rp, col, "permuted cells, device locality order", flow.CELL_ORDER_LOCALITY)
    # the same loops with the generator's lexicographic cell numbering, where
    # a cell's neighbours sit near it in memory (AUTO keeps it)
    mesh_n, rp_n, col_n = gen.hex_mesh(n)
    res_nat = timed(mesh_n, rp_n, col_n, "lexicographic cells")
    del mesh_n, rp_n, col_n
    cpu = None
    if not args.no_cpu_baseline:
        threads = host_threads()
        oracle.set_threads(threads)
        t0 = time.perf_counter()
        for eq in range(5):
            oracle.assemble(eq, mesh, rp, col, prm, st, fast=True)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        cpu = {"value": round(cpu_ms, 1), "unit": "ms for the five equations", "cores": threads, "kind": "port",
               "sample": "one assembly of each equation on the same mesh and state (the oracle's restatement of "
                         "the reference loops, OpenMP with atomic adds into linearly searched CSR slots)"}
    def short(r):
        return {k: (v if not isinstance(v, dict) else
                    {kk: v[kk] for kk in ("avg_us", "frac", "traffic_over_algorithmic") if kk in v})
                for k, v in r.items()}

    def total_us(r):
        return round(sum(v["avg_us"] for v in r.values() if isinstance(v, dict)), 2)

    out["assembly"] = {"workload": f"face/boundary/cell coefficient loops, {n}^3 hex mesh (N={N}, faces={F}, "
                                   f"boundary faces={B}), viscous, blended, transient, permuted cells",
                       "kernel": "k_assemble", "gpu": res,
                       "gpu_total_us": total_us(res),
                       "gpu_rcm_order": short(res_rcm),
                       "gpu_rcm_order_total_us": total_us(res_rcm),
                       "gpu_natural_order": short(res_nat),
                       "gpu_natural_order_total_us": total_us(res_nat),
                       "cpu_baseline": cpu}


def stream_triad():
    """Measured STREAM-triad bandwidth (a = b + 3 c, FP64, 16-byte accesses):
    three 512 MiB arrays, far beyond the 256 MB last-level cache (the HBM
    rate), and three 32 MiB ones, which the cache holds across launches."""
    import ctypes

    from opencslsolver_amd._native import hip_lib

    lib = hip_lib()
    res = {}
    for key, n, reps in (("hbm_GBps", 64 << 20, 20), ("cache_resident_GBps", 4 << 20, 200)):
        ms, by = ctypes.c_double(), ctypes.c_double()
        st = lib.ocsl_time_stream_triad(n, reps, ctypes.byref(ms), ctypes.byref(by))
        if st != 0:
            raise RuntimeError(f"ocsl_time_stream_triad failed ({st})")
        res[key] = round(by.value / (ms.value * 1e-3) / 1e9, 1)
    return res


def mesh_read(args, out):
    """SU2 mesh set-up on the host (libocsl_mesh.so): parsing, cell/face
    connectivity, geometry and the finite-volume face factors of a 64^3
    hexahedral box in random node and cell numbering (seed 31), beside the
    reference's own Read_SU2_mesh + CFVM::create_FVM_from_geometry (the
    reference build in oracle/_ref, when present) on a 32^3 sample."""
    import tempfile

    from opencslsolver_amd import mesh

    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        for n in (32, 64):
            _log(f"bench: SU2 mesh {n}^3")
            path = os.path.join(tmp, f"box{n}.su2")
            with open(path, "w") as f:
                f.write(mesh.box_su2(n, "hex", h=1.0 / n, seed=31))
            t0 = time.perf_counter()
            m = mesh.read_su2(path)
            res[n] = (time.perf_counter() - t0, m["cell_volume"].shape[0], m["face_left"].shape[0],
                      m["bc_left"].shape[0])
        cpu = None
        import subprocess

        import oracle

        ref_s = None
        if not args.no_cpu_baseline and oracle.ref_available():
            # a child process: the reference build targets the build host's
            # instruction set, and a failure there must not end this run
            code = ("import sys, json; sys.path.insert(0, sys.argv[1]); import oracle; "
                    "print(json.dumps(oracle.ref_su2_stage_times(sys.argv[2])))")
            r = subprocess.run([sys.executable, "-c", code, ROOT, os.path.join(tmp, "box32.su2")],
                               capture_output=True, text=True, timeout=300)
            if r.returncode == 0:
                ref_s = json.loads(r.stdout.strip().splitlines()[-1])
            else:
                _log(f"bench: reference mesh reader unavailable here (exit {r.returncode})")
        if ref_s is not None:
            # like for like: the port's read_su2 parses, connects and computes
            # the geometry and face factors, as Read_SU2_mesh and
            # create_FVM_from_geometry do; the matrix structure is a separate
            # step on both sides (the port's is the device structure builder)
            like = ref_s[0] + ref_s[1]
            cpu = {"value": round(like * 1e3, 1), "unit": "ms for the 32^3 file", "cores": 1, "kind": "reference",
                   "sample": "the reference's Read_SU2_mesh + create_FVM_from_geometry (built from its sources, "
                             "timed inside the library, its progress output included) on the 32^3 file",
                   "stages_ms": {"Read_SU2_mesh": round(ref_s[0] * 1e3, 1),
                                 "create_FVM_from_geometry": round(ref_s[1] * 1e3, 1),
                                 "define_matrix_structure": round(ref_s[2] * 1e3, 1)},
                   "speedup_32": round(like / res[32][0], 1)}
    t64 = res[64]
    out["mesh"] = {"workload": f"SU2 hexahedral box, 64^3 cells ({t64[1]} cells, {t64[2]} interior faces, "
                               f"{t64[3]} boundary faces), random numbering (seed 31): read, connect, geometry, "
                               "face factors",
                   "ms_64": round(t64[0] * 1e3, 1), "cells_per_s": round(t64[1] / t64[0]),
                   "ms_32": round(res[32][0] * 1e3, 1), "cpu_baseline": cpu}


def structure(args, out):
    """Matrix structure of the n^3 hex mesh (CFVM::define_matrix_structure) built
    on the device: from host arrays (PCIe included) and from device arrays."""
    import ctypes

    import torch

    from opencslsolver_amd import gen
    from opencslsolver_amd._native import hip_lib
    from opencslsolver_amd.solver import build_structure

    n = args.n
    _log(f"bench: structure of the {n}^3 mesh")
    mesh, rp_g, col_g = gen.hex_mesh(n)
    fl, fr, bl = mesh["face_left"], mesh["face_right"], mesh["bc_left"]
    build_structure(n ** 3, fl, fr, bl)  # warm (module load)
    t0 = time.perf_counter()
    rp, col, fs, ds = build_structure(n ** 3, fl, fr, bl, slots=True)
    host_ms = (time.perf_counter() - t0) * 1e3
    ok = bool(np.array_equal(rp, rp_g) and np.array_equal(col, col_g))
    dev = torch.device("cuda:0")
    d_fl, d_fr, d_bl = (torch.from_numpy(a).to(dev) for a in (fl, fr, bl))
    d_rp = torch.empty(n ** 3 + 1, dtype=torch.int64, device=dev)
    d_col = torch.empty(col.shape[0], dtype=torch.int64, device=dev)
    d_fs = torch.empty(2 * fl.shape[0], dtype=torch.int64, device=dev)
    d_ds = torch.empty(n ** 3, dtype=torch.int64, device=dev)
    lib = hip_lib()
    nnz = ctypes.c_int64()
    vp = ctypes.c_void_p
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        st = lib.ocsl_build_structure(n ** 3, fl.shape[0], vp(d_fl.data_ptr()), vp(d_fr.data_ptr()), bl.shape[0],
                                      vp(d_bl.data_ptr()), 1, vp(d_rp.data_ptr()), vp(d_col.data_ptr()),
                                      col.shape[0], ctypes.byref(nnz), vp(d_fs.data_ptr()), vp(d_ds.data_ptr()))
        assert st == 0
    torch.cuda.synchronize()
    dev_ms = (time.perf_counter() - t0) * 1e3 / reps
    out["structure"] = {
        "workload": f"CSR structure + face and diagonal slot maps of the {n}^3 hex mesh: {n ** 3} cells, "
                    f"{fl.shape[0]} interior faces, {bl.shape[0]} boundary faces, {col.shape[0]} entries",
        "device_arrays_ms": round(dev_ms, 3),
        "host_arrays_ms": round(host_ms, 3),
        "equals_generator": ok,
    }


def tri_solves(args, out):
    """Exact ILU(0) triangular solves (ILU_forsub / ILU_backsub of the 3-sweep
    factor, level-scheduled) on randomly numbered convection-diffusion
    systems (about 20 level sets: one launch per level) and a lexicographic
    one (one level per diagonal plane: one workgroup for the whole solve)."""
    from opencslsolver_amd import gen
    from opencslsolver_amd.solver import Solver

    res = {}
    cases = [("1000-row chain (one row per level)", gen.chain1d(1000, 5))]
    for n, permute in ((32, True), (48, True), (64, True), (96, True), (96, False)):
        cases.append((f"{n}^3 {'permuted' if permute else 'lexicographic'}",
                      gen.convdiff3d(n, seed_perm=3, permute=permute)[0]))
    for key, a in cases:
        _log(f"bench: triangular solves, {key}")
        with Solver(a.row_ptr, a.col) as s:
            s.set_values(a.val)
            res[key] = {"lower_us": round(s.time_tri_solve(True, reps=20) * 1e3, 2),
                        "upper_us": round(s.time_tri_solve(False, reps=20) * 1e3, 2),
                        # the forms the solvers choose between
                        "lower_us_one_workgroup": round(s.time_tri_solve(True, reps=20, form=1) * 1e3, 2),
                        "lower_us_launch_per_level": round(s.time_tri_solve(True, reps=20, form=2) * 1e3, 2)}
            if a.n <= 1000:
                res[key]["lower_us_serial"] = round(s.time_tri_solve(True, reps=20, form=3) * 1e3, 2)
    out["tri_solves"] = {"workload": "exact L and U solves of the 3-sweep ILU(0) factor of the convection-diffusion "
                                     "system (HIP events, 20 solves each)", **res}


# SIMPLE run on an SU2 box, the same settings for this package and for the
# reference's own run of the file: viscous, blended, transient dt 1e-4, CG
# for the pressure correction and ILU(0)-BiCGSTAB for the rest, one boundary
# condition per marker of mesh.box_su2 (x-: velocity inlet, x+: pressure
# outlet, y-: no-slip wall, y+: slip wall, z-/z+: symmetry)
FLOW_SU2_STEPS = 3
FLOW_SU2_SETTINGS = dict(steps=FLOW_SU2_STEPS, inner=1, dt=1e-4, steady=0, implicit_underrelax=0.4,
                         implicit_underrelax_max=0.8, ramp_iterations=100, solver_type=1, solver_mode=1, conv_type=1,
                         visc_type=0, init_shock_tube=0, p0=1e5, T0=300.0, u0=(2.0, 0.0, 0.0), p2=0.0, T2=0.0)
FLOW_SU2_BCS = dict(bc_type=[3, 5, 2, 2, 1, 1], bc_no_slip=[0, 0, 1, 0, 0, 0], bc_const_T=[0] * 6,
                    bc_T=[300.0, 0.0, 0.0, 0.0, 0.0, 0.0], bc_p=[0.0, 1e5, 0.0, 0.0, 0.0, 0.0],
                    bc_u=[[3.0, 0.5, -0.2]] + [[0.0, 0.0, 0.0]] * 5)


def _flow_su2_text(n):
    from opencslsolver_amd import mesh

    return mesh.box_su2(n, "hex", h=1.0 / n, seed=31)


def _ref_flow_child(n, threads, out_path):
    """Child-process leg of the flow-step CPU baseline: the reference's own
    SIMPLE run of the SU2 box (built from its sources for the build host's
    instruction set), timed by its own per-iteration clock; writes the final
    fields to out_path and prints JSON."""
    import tempfile

    import oracle

    text = _flow_su2_text(n)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "box.su2")
        with open(path, "w") as f:
            f.write(text)
        os.chdir(tmp)  # the reference writes its VTK files into the working directory
        sys.stdout.flush()
        saved = os.dup(1)
        os.dup2(2, 1)  # and its progress lines to stdout
        try:
            r = oracle.ref_flow_time_file(path, FLOW_SU2_SETTINGS, FLOW_SU2_BCS, threads, n ** 3)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
    np.savez(out_path, **{k: r[k] for k in ("p", "u", "T", "rho", "h")})
    print(json.dumps({k: r[k] for k in ("iter_us", "total_s", "solves")}))


def flow_su2_vs_reference(args, n=32):
    """The SIMPLE step on an n^3 hexahedral SU2 box in random numbering, read
    by this package's reader (matrix structure from the device builder), run
    for FLOW_SU2_STEPS iterations from the initial state; then timed over
    further iterations.  Beside it, the reference's own run of the same file
    on the host cores, timed by its own per-iteration clock, and the largest
    relative difference of the final fields between the two runs."""
    import subprocess
    import tempfile

    import torch

    import oracle
    from opencslsolver_amd import flow, mesh
    from opencslsolver_amd.solver import build_structure

    _log(f"bench: SIMPLE iterations on a {n}^3 SU2 box, beside the reference")
    m = mesh.parse_su2(_flow_su2_text(n))
    N = m["cell_volume"].shape[0]
    rp, col = build_structure(N, m["face_left"], m["face_right"], m["bc_left"])
    b = FLOW_SU2_BCS
    bcs = [dict(type=b["bc_type"][k], no_slip=b["bc_no_slip"][k], const_temperature=b["bc_const_T"][k],
                temperature=b["bc_T"][k], pressure=b["bc_p"][k], velocity=tuple(b["bc_u"][k]))
           for k in range(len(b["bc_type"]))]
    st = FLOW_SU2_SETTINGS
    prm = flow.default_params(viscous=1, blended=1, transient=1, dt=st["dt"],
                              implicit_underrelax=st["implicit_underrelax"],
                              implicit_underrelax_max=st["implicit_underrelax_max"],
                              ramp_iterations=st["ramp_iterations"], p0=st["p0"], t0=st["T0"], u0=st["u0"], bcs=bcs,
                              method_momentum=1, method_pressure=0, method_enthalpy=1)
    with flow.Flow(m, rp, col, prm) as fl:
        fl.initialize()
        fl.run(FLOW_SU2_STEPS, 1)
        gpu = fl.get_state()
        iters = max(3, args.steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fl.run(iters, 1)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / iters
    res = {"workload": f"SIMPLE iteration, {n}^3 hexahedral SU2 box in random numbering (seed 31, N={N}), inlet/"
                       "outlet/no-slip and slip walls/symmetry, viscous, blended, transient dt 1e-4, REFERENCE-mode "
                       "solves (BiCGSTAB momentum and enthalpy, CG pressure)",
           "gpu_ms_per_iteration": round(ms, 3)}
    if args.no_cpu_baseline or not oracle.ref_available():
        return res
    threads = host_threads()
    with tempfile.TemporaryDirectory() as tmp:
        out_path = os.path.join(tmp, "ref_fields.npz")
        code = ("import sys; sys.path.insert(0, sys.argv[1]); import bench; "
                "bench._ref_flow_child(int(sys.argv[2]), int(sys.argv[3]), sys.argv[4])")
        r = subprocess.run([sys.executable, "-c", code, ROOT, str(n), str(threads), out_path],
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            res["cpu_baseline"] = {"error": r.stderr.strip().splitlines()[-1:]}
            return res
        d = json.loads(r.stdout.strip().splitlines()[-1])
        ref = np.load(out_path)
        diff = {f: float(np.max(np.abs(gpu[f] - ref[f])) / max(np.max(np.abs(ref[f])), 1e-300))
                for f in ("p", "u", "T", "rho", "h")}
    ref_ms = d["iter_us"] * 1e-3
    res["cpu_baseline"] = {"value": round(ref_ms, 2), "unit": "ms per SIMPLE iteration", "cores": threads,
                           "kind": "reference", "cpu_model": cpu_model(),
                           "sample": f"the reference's own run of the same SU2 file ({FLOW_SU2_STEPS} iterations "
                                     f"from the initial state, {d['solves']} solves), timed by its per-iteration "
                                     f"clock (Croot::start_solver); whole run incl. reading and set-up "
                                     f"{d['total_s']:.2f} s"}
    res["speedup_gpu_vs_reference"] = round(ref_ms / ms, 1)
    res["fields_max_rel_diff_after_3_iterations"] = diff
    return res


def _ref_sod_child(steps, threads, out_path):
    """Child-process leg of the shock-tube CPU baseline: the reference's own
    run (built from its sources for the build host's instruction set) timed
    by its own per-iteration clock; writes the final fields to out_path and
    prints JSON."""
    import tempfile

    import oracle

    with tempfile.TemporaryDirectory() as tmp:
        os.chdir(tmp)  # the reference writes its VTK files into the working directory
        sys.stdout.flush()
        saved = os.dup(1)
        os.dup2(2, 1)  # and its progress lines to stdout
        try:
            r = oracle.ref_sod_time(steps, threads)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
    np.savez(out_path, **{k: r[k] for k in ("p", "u", "T", "rho", "h")})
    print(json.dumps({k: r[k] for k in ("iter_us", "total_s", "solves")}))


def sod(args, out):
    """The reference's CPU-runnable case: the 1D shock tube, 1000 cells over
    1 m (the reference's own 1D mesh, recorded in the fixture
    flow_sod1000_mesh.npz), slip walls, empty side faces, inviscid, blended
    central convection, dt 1e-6, 200 steps of 10 inner iterations, CG for the
    pressure correction and REFERENCE ILU(0)-BiCGSTAB for momentum and
    enthalpy, the state resident on the device.  Beside it the reference's
    own run of the same case on the host, timed by its per-iteration clock,
    and the largest relative difference of the final fields."""
    import subprocess
    import tempfile

    import torch

    import oracle
    from opencslsolver_amd import flow

    steps, inner = args.sod_steps, 10
    _log(f"bench: shock tube, {steps} steps x {inner} inner iterations")
    mz = np.load(os.path.join(ROOT, "tests", "golden", "flow_sod1000_mesh.npz"))
    bcs = [dict(type=flow.WALL, no_slip=0), dict(type=flow.EMPTY), dict(type=flow.WALL, no_slip=0)]
    prm = flow.default_params(viscous=0, blended=1, transient=1, dt=1e-6, implicit_underrelax=0.4,
                              implicit_underrelax_max=0.8, ramp_iterations=100, shock_tube=1, p0=1e5, t0=348.4,
                              u0=(0.0, 0.0, 0.0), p2=1e4, t2=278.7, bcs=bcs, method_momentum=1, method_pressure=0,
                              method_enthalpy=1)
    with flow.Flow(flow.mesh_from_npz(mz), mz["row_ptr"], mz["col"], prm) as fl:
        fl.initialize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fl.run(steps, inner)
        torch.cuda.synchronize()
        gpu_s = time.perf_counter() - t0
        gpu = fl.get_state()
    res = {"workload": f"1D shock tube, 1000 cells, {steps} steps x {inner} inner iterations (dt 1e-6), CG for the "
                       "pressure correction, REFERENCE ILU(0)-BiCGSTAB for momentum and enthalpy",
           "gpu_s": round(gpu_s, 3), "gpu_ms_per_iteration": round(gpu_s * 1e3 / (steps * inner), 3)}
    if args.no_cpu_baseline or not oracle.ref_available():
        out["sod"] = res
        return
    legs = {}
    with tempfile.TemporaryDirectory() as tmp:
        for threads in sorted({1, host_threads()}):
            out_path = os.path.join(tmp, f"ref_{threads}.npz")
            code = ("import sys; sys.path.insert(0, sys.argv[1]); import bench; "
                    "bench._ref_sod_child(int(sys.argv[2]), int(sys.argv[3]), sys.argv[4])")
            r = subprocess.run([sys.executable, "-c", code, ROOT, str(steps), str(threads), out_path],
                               capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                legs[threads] = {"error": r.stderr.strip().splitlines()[-1:]}
                continue
            d = json.loads(r.stdout.strip().splitlines()[-1])
            ref = np.load(out_path)
            d["fields_max_rel_diff"] = {f: float(np.max(np.abs(gpu[f] - ref[f])) / max(np.max(np.abs(ref[f])), 1e-300))
                                        for f in ("p", "u", "T", "rho", "h")}
            legs[threads] = d
    ok = {t: d for t, d in legs.items() if "iter_us" in d}
    if ok:
        best = min(ok, key=lambda t: ok[t]["iter_us"])
        d = ok[best]
        res["cpu_baseline"] = {"value": round(d["iter_us"] * 1e-3, 3), "unit": "ms per SIMPLE iteration",
                               "cores": best, "kind": "reference", "cpu_model": cpu_model(),
                               "sample": f"the reference's own run of the same case ({d['solves']} solves), timed "
                                         f"by its per-iteration clock (Croot::start_solver); the faster of 1 and "
                                         f"{host_threads()} OpenMP threads; whole run {d['total_s']:.2f} s",
                               "by_threads_ms_per_iteration": {str(t): round(v["iter_us"] * 1e-3, 3)
                                                               for t, v in ok.items()}}
        res["speedup_gpu_vs_reference"] = round(d["iter_us"] * 1e-3 / res["gpu_ms_per_iteration"], 2)
        res["fields_max_rel_diff_vs_reference_1_thread"] = ok.get(1, d)["fields_max_rel_diff"]
    out["sod"] = res


def flow_step(args, out):
    """One SIMPLE iteration of the flow solver (three momentum solves with
    ILU(0)-BiCGSTAB, the pressure correction with CG, enthalpy; REFERENCE-mode
    solves, the drop-in semantics) on a permuted n^3 hex mesh with an inlet,
    an outlet, walls and symmetry planes, viscous, blended, transient: wall
    time per iteration with the support loops as device kernels over the
    resident state, and with their host restatement (per-equation uploads)."""
    import torch

    from opencslsolver_amd import flow, gen

    n = args.flow_n
    _log(f"bench: SIMPLE iterations on a {n}^3 hex mesh")
    rng = np.random.default_rng(31)
    perm = rng.permutation(n ** 3)
    mesh, rp, col = gen.hex_mesh(n, perm=perm)
    bcs = [dict(type=flow.VELOCITY_INLET, velocity=(3.0, 0.5, -0.2), temperature=300.0),
           dict(type=flow.PRESSURE_OUTLET, pressure=1e5), dict(type=flow.WALL, no_slip=1),
           dict(type=flow.WALL, no_slip=0), dict(type=flow.SYMMETRY), dict(type=flow.SYMMETRY)]
    res = {}
    for name, support in (("device", flow.SUPPORT_DEVICE), ("host_support", flow.SUPPORT_HOST)):
        prm = flow.default_params(viscous=1, blended=1, transient=1, dt=1e-4, p0=1e5, t0=300.0,
                                  u0=(2.0, 0.0, 0.0), bcs=bcs, method_pressure=0, support=support)
        with flow.Flow(mesh, rp, col, prm) as fl:
            fl.initialize()
            fl.run(1, 1)  # warm: layouts, level schedules, graphs
            torch.cuda.synchronize()
            iters = max(1, args.steps)
            t0 = time.perf_counter()
            fl.run(iters, 1)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / iters
            st = fl.get_state()
        res[name] = {"ms_per_iteration": round(ms, 3), "p_mean": float(np.mean(st["p"])),
                     "T_mean": float(np.mean(st["T"]))}
    out["flow_step"] = {"workload": f"SIMPLE iteration, {n}^3 permuted hex mesh (N={n ** 3}), inlet/outlet/walls/"
                                    "symmetry, viscous, blended, transient dt 1e-4, REFERENCE-mode solves "
                                    "(BiCGSTAB momentum and enthalpy, CG pressure)",
                        **res,
                        "speedup_device_vs_host_support": round(res["host_support"]["ms_per_iteration"]
                                                                / res["device"]["ms_per_iteration"], 2),
                        "su2_box": flow_su2_vs_reference(args)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--sections", default="adjoint,multi,batch,assembly,tri,flow,sod,structure,mesh",
                    help="side workloads to add to the headline: adjoint, multi, batch, assembly, tri, flow, sod, "
                         "structure, mesh (or none)")
    ap.add_argument("--flow-n", type=int, default=64, help="hex mesh size of the flow-step section")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--reuse-tuning", action="store_true",
                    help="apply the cached tuned launch when present instead of tuning again")
    ap.add_argument("--skip-headline", action="store_true",
                    help="run only the side sections (for profiling them); the JSON line then has no value")
    ap.add_argument("--sod-steps", type=int, default=200,
                    help="time steps of the shock tube (sod section); fewer for a profiled run")
    ap.add_argument("--skip-batch-sequential", action="store_true",
                    help="batch section: leave out the same systems solved one by one")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only launch the measured kernels, --steps times each (for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the solvers are single-device)", file=sys.stderr)
        sys.exit(2)
    args.sections = set(x.strip() for x in args.sections.split(",")) if args.sections != "none" else set()
    if args.profile_kernels:
        profile_kernels(args)
        return

    out = {
        "metric": "Krylov solve ms + HBM GB/s at 2.1M cells, FP32-refined vs FP64, to 1e-8 rel res",
        "value": None,
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded generator, regenerated on the box)",
    }
    if not args.skip_headline:
        headline(args, out)
    if "multi" in args.sections:
        multi_rhs(args, out)
    if "batch" in args.sections:
        batch(args, out)
    if "assembly" in args.sections:
        assembly(args, out)
    if "tri" in args.sections:
        tri_solves(args, out)
    if "flow" in args.sections:
        flow_step(args, out)
    if "sod" in args.sections:
        sod(args, out)
    if "structure" in args.sections:
        structure(args, out)
    if "mesh" in args.sections:
        mesh_read(args, out)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
