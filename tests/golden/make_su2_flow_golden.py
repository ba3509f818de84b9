"""Writes tests/golden/flow_su2_{hex,prism,tet}.npz: the reference's own
SIMPLE runs on 3D boxes read from SU2 files (Emesh_source::from_file,
croot.cpp:72-76; Read_SU2_mesh, cgeometry_supply.h:201-433), six markers with
six boundary conditions (velocity inlet, pressure outlet, no-slip wall, slip
wall, two symmetry planes); viscous, blended, transient, CG for the pressure
correction and ILU(0)-BiCGSTAB for the rest.  Each fixture holds the SU2 text,
the run settings, the per-solve residual history, the reference's matrix
structure and the final fields, from oracle/_ref/libocsl_ref.so
(ref_sod_driver.cpp's ref_flow_run_file).  The same runs through the build
without contraction or fast-math (oracle/_ref/ieee) give the
reference-vs-reference spread of every compared quantity, written to
tests/golden/spread.json under "su2_runs".

    make -C oracle ref ref-ieee && python tests/golden/make_su2_flow_golden.py

(after make_spread.py, which rewrites spread.json without this section)
"""
from __future__ import annotations

import ctypes
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from opencslsolver_amd import mesh  # noqa: E402

import parity  # noqa: E402
from make_golden import PRESSURE_OUTLET, SYMMETRY, VELOCITY_INLET, WALL, RefFlowConfig  # noqa: E402

CASES = {"hex": (6, 51), "prism": (5, 52), "tet": (4, 53)}
SETTINGS = dict(steps=2, inner=3, dt=1e-4, steady=0, implicit_underrelax=0.4, implicit_underrelax_max=0.8,
                ramp_iterations=100, solver_type=1, solver_mode=1, conv_type=1, visc_type=0, init_shock_tube=0,
                p0=1e5, T0=300.0, u0=(2.0, 0.0, 0.0), p2=0.0, T2=0.0)
# one per marker of mesh.box_su2 (xmin, xmax, ymin, ymax, zmin, zmax)
BCS = dict(bc_type=[VELOCITY_INLET, PRESSURE_OUTLET, WALL, WALL, SYMMETRY, SYMMETRY],
           bc_no_slip=[0, 0, 1, 0, 0, 0], bc_const_T=[0, 0, 0, 0, 0, 0],
           bc_T=[300.0, 0.0, 0.0, 0.0, 0.0, 0.0], bc_p=[0.0, 1e5, 0.0, 0.0, 0.0, 0.0],
           bc_u=[[3.0, 0.5, -0.2]] + [[0.0, 0.0, 0.0]] * 5)


def run(variant: str, su2_text: str, n_cells: int) -> dict:
    c = RefFlowConfig()
    for k, v in SETTINGS.items():
        if k == "u0":
            for d in range(3):
                c.u0[d] = v[d]
        else:
            setattr(c, k, v)
    c.n_cells = n_cells
    nb = len(BCS["bc_type"])
    i32 = lambda a: (ctypes.c_int * nb)(*a)  # noqa: E731
    f64 = lambda a: (ctypes.c_double * len(a))(*a)  # noqa: E731
    hist_cap = SETTINGS["steps"] * SETTINGS["inner"] * 5
    hist = np.zeros(hist_cap)
    nnz_cap = 16 * n_cells
    rp, col = np.zeros(n_cells + 1, np.uint64), np.zeros(nnz_cap, np.uint64)
    nnz = ctypes.c_int64()
    p, T, rho, h = (np.zeros(n_cells) for _ in range(4))
    u = np.zeros(3 * n_cells)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    r = oracle.ref(variant)
    r.ref_flow_run_file.restype = ctypes.c_int64
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "box.su2")
        with open(path, "w") as f:
            f.write(su2_text)
        os.chdir(tmp)  # the reference writes its VTK files into the working directory
        sys.stdout.flush()
        saved = os.dup(1)
        os.dup2(2, 1)  # and its progress lines to stdout
        try:
            cnt = r.ref_flow_run_file(
                ctypes.byref(c), path.encode(), nb, i32(BCS["bc_type"]), i32(BCS["bc_no_slip"]),
                i32(BCS["bc_const_T"]), f64(BCS["bc_T"]), f64(BCS["bc_p"]),
                f64([v for u3 in BCS["bc_u"] for v in u3]), ctypes.c_int64(n_cells), ctypes.c_int64(hist_cap),
                P(hist), ctypes.c_int64(nnz_cap), P(rp), P(col), ctypes.byref(nnz), P(p), P(u), P(T), P(rho), P(h))
        finally:
            os.dup2(saved, 1)
            os.close(saved)
            os.chdir(cwd)
    assert cnt > 0, (variant, cnt)
    k = int(nnz.value)
    return dict(count=cnt, res_hist=hist[:cnt], row_ptr=rp, col=col[:k], p=p, u=u.reshape(n_cells, 3), T=T,
                rho=rho, h=h)


def main():
    spread_path = os.path.join(HERE, "spread.json")
    spread = json.load(open(spread_path))
    spread["su2_runs"] = {}
    for kind, (n, seed) in CASES.items():
        text = mesh.box_su2(n, kind, h=0.01, seed=seed)
        n_cells = mesh.parse_su2(text)["cell_volume"].shape[0]
        out = run("", text, n_cells)
        ieee = run("ieee", text, n_cells)
        meta = dict(SETTINGS, **BCS, kind=kind, n=n, seed=seed, n_cells=n_cells)
        np.savez_compressed(os.path.join(HERE, f"flow_su2_{kind}.npz"), su2_text=np.frombuffer(text.encode(), np.uint8),
                            meta=np.frombuffer(json.dumps(meta).encode(), np.uint8), **out)
        sp = {f: parity.rel(ieee[f], out[f]) for f in ("p", "T", "rho", "h", "u")}
        sp["hist"] = parity.hist_rel(ieee["res_hist"], out["res_hist"])
        spread["su2_runs"][kind] = sp
        print(kind, n_cells, "cells,", out["count"], "solves; spread", sp, flush=True)
    with open(spread_path, "w") as f:
        json.dump(spread, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
