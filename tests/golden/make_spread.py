"""Measures how far rounding alone moves each result the parity tests compare
with the reference, and writes the figures to tests/golden/spread.json.

The golden fixtures come from the reference built with its own flags
(-O3 -ffast-math -march=native: contracted multiply-adds, reassociated
reductions).  This script runs the same workloads through the same sources
built without contraction or fast-math (make -C oracle ref-ieee) and records,
per quantity, the distance between the two builds, measured exactly as the
GPU test measures its own distance from the fixture.  A parity bar that has
to be looser than 1e-10 is set from these figures.

    make -C oracle ref ref-ieee && python tests/golden/make_spread.py
"""
from __future__ import annotations

import functools
import glob
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from opencslsolver_amd.gen import CSR  # noqa: E402

import make_golden  # noqa: E402
import parity  # noqa: E402


rel, res_rel, hist_rel = parity.rel, parity.res_rel, parity.hist_rel


def with_ieee(fn, *args, **kw):
    """Runs a make_golden generator against the IEEE reference build."""
    fast = oracle.ref
    oracle.ref = functools.partial(fast, "ieee")
    try:
        return fn(*args, **kw)
    finally:
        oracle.ref = fast


def solver_spread():
    out = {}
    for path in sorted(glob.glob(os.path.join(HERE, "solver_*.npz"))):
        d = np.load(path)
        a = CSR(d["row_ptr"], d["col"], d["val"])
        r_bi, x_bi = oracle.ref_ilu_bicgstab_solve(a, d["b"], d["x0"], variant="ieee")
        r_cg, x_cg = oracle.ref_cg_solve(a, d["b"], d["x0"], variant="ieee")
        out[os.path.basename(path)[:-4]] = dict(
            bicg_res=res_rel(r_bi, d["bicg_res"]), bicg_x=rel(x_bi, d["bicg_x"]),
            cg_res=res_rel(r_cg, d["cg_res"]), cg_x=rel(x_cg, d["cg_x"]))
    return out


def sod_system_spread():
    """The recorded shock-tube systems solved alone (REFERENCE BiCGSTAB)."""
    out = {}
    for tag in ("bicgstab", "split"):
        d = np.load(os.path.join(HERE, f"sod_{tag}_2steps.npz"))
        res, xs = [], []
        for k in range(d["dump_val"].shape[0]):
            if d["dump_method"][k] != 1:
                continue
            a = CSR(d["row_ptr"], d["col"], d["dump_val"][k])
            r, x = oracle.ref_ilu_bicgstab_solve(a, d["dump_b"][k], d["dump_x0"][k], variant="ieee")
            res.append(res_rel(r, d["res_hist"][k]))
            xs.append(rel(x, d["dump_x"][k]))
        out[tag] = dict(res=max(res), x=max(xs))
    return out


FIELDS = ("p", "T", "rho", "h", "u")


def run_spread(r, d):
    e = {k: rel(r[k], d[k]) for k in FIELDS}
    e["hist"] = hist_rel(r["res_hist"], d["res_hist"])
    return e


def sod_run_spread():
    out = {}
    for tag, st, sm in (("bicgstab", 1, 0), ("split", 0, 1)):
        for steps in (2, 200):
            d = np.load(os.path.join(HERE, f"sod_{tag}_{steps}steps.npz"))
            r = with_ieee(make_golden.sod_run, steps, st, sm, 0)
            out[f"{tag}_{steps}"] = run_spread(r, d)
    return out


def flow_run_spread():
    out = {}
    for tag, cfg in make_golden.FLOW_CASES.items():
        d = np.load(os.path.join(HERE, f"flow_{tag}.npz"))
        r = with_ieee(make_golden.flow_run, cfg, 0, 0)
        out[tag] = run_spread(r, d)
    return out


def main():
    if not (oracle.ref_available() and oracle.ref_available("ieee")):
        sys.exit("both reference builds are needed: make -C oracle ref ref-ieee (needs the reference sources)")
    res = dict(
        note=("distance of the reference built with -O2 -ffp-contract=off from the fixtures "
              "(the reference's -O3 -ffast-math -march=native build), per quantity"),
        solver=solver_spread(), sod_systems=sod_system_spread(), sod_runs=sod_run_spread(),
        flow_runs=flow_run_spread())
    path = os.path.join(HERE, "spread.json")
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
