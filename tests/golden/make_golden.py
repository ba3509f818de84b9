"""Writes the golden vectors under tests/golden/ from the reference solver.

Runs oracle/_ref/libocsl_ref.so -- the reference compiled from its own
sources by ``make -C oracle ref`` in a container that holds them -- on seeded
inputs and stores inputs and outputs as .npz data files (no code).  The GPU
box never runs this script; the tests read the .npz files.

    python tests/golden/make_golden.py
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

import oracle  # noqa: E402
from opencslsolver_amd import gen  # noqa: E402

_f = ctypes.POINTER(ctypes.c_double)
_u = ctypes.POINTER(ctypes.c_uint64)
_i64 = ctypes.POINTER(ctypes.c_int64)
_ip = ctypes.POINTER(ctypes.c_int)


def _save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


def solver_cases():
    """Reference CG / ILU-BiCGSTAB results on small seeded systems."""
    cases = {}
    a, b, x0 = gen.tridiag_unit_test_system()
    cases["unit10"] = (a, b, x0)
    a, b, _ = gen.poisson3d(6, seed_mesh=1, seed_rhs=2, sentinels=True)
    cases["poisson6s"] = (a, b, gen.uniform(11, a.n))
    a, b, _ = gen.poisson3d(8, seed_mesh=5, seed_rhs=6, sentinels=False)
    cases["poisson8"] = (a, b, np.zeros(a.n))
    ac, _ = gen.convdiff3d(6, seed_perm=3, sentinels=True)
    cases["convdiff6s"] = (ac, gen.uniform(4, ac.n), gen.uniform(12, ac.n))
    for name, (a, b, x0) in cases.items():
        res_cg, x_cg = oracle.ref_cg_solve(a, b, x0)
        res_bi, x_bi = oracle.ref_ilu_bicgstab_solve(a, b, x0)
        m = oracle.ref_ilu_decompose(a)
        y = oracle.ref_forsub(a, m, b)
        z = oracle.ref_backsub(a, m, y)
        g = oracle.ref_gemv(a, x0, -1.0, 1.0, b)
        _save(f"solver_{name}", row_ptr=a.row_ptr, col=a.col, val=a.val, b=b, x0=x0,
              cg_res=np.array(res_cg), cg_x=x_cg, bicg_res=np.array(res_bi), bicg_x=x_bi,
              ilu_m=m, forsub_y=y, backsub_x=z, gemv_r=g,
              norm_bb=np.array(oracle.ref().ref_norm(ctypes.c_int64(a.n), b.ctypes.data_as(_f),
                                                      b.ctypes.data_as(_f))),
              abssum_b=np.array(oracle.ref().ref_abssum(ctypes.c_int64(a.n), b.ctypes.data_as(_f))))


SOD = dict(n_cells=1000, lx=1.0, ly=0.1, lz=0.1, dt=1e-6, p_left=1e5, t_left=348.4, p_right=1e4,
           t_right=278.7, conv_type=1, visc_type=1, inner=10)


def sod_run(steps, solver_type, solver_mode, max_dump):
    c = SOD
    n = c["n_cells"]
    nnz_cap = 5 * n
    hist_cap = steps * c["inner"] * 5
    hist = np.zeros(hist_cap)
    rp = np.zeros(n + 1, np.uint64)
    col = np.zeros(nnz_cap, np.uint64)
    nnz = ctypes.c_int64()
    dv = np.zeros((max(max_dump, 1), nnz_cap))
    db = np.zeros((max(max_dump, 1), n))
    dx0 = np.zeros((max(max_dump, 1), n))
    dx = np.zeros((max(max_dump, 1), n))
    dm = np.zeros(max(max_dump, 1), np.int32)
    p, T, rho, h = (np.zeros(n) for _ in range(4))
    u = np.zeros(3 * n)
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as tmp:
        os.chdir(tmp)  # the reference writes solution.vtk into the working directory
        try:
            cnt = oracle.ref().ref_sod_run(
                n, ctypes.c_double(c["lx"]), ctypes.c_double(c["ly"]), ctypes.c_double(c["lz"]),
                steps, c["inner"], ctypes.c_double(c["dt"]), solver_type, solver_mode,
                c["conv_type"], c["visc_type"], ctypes.c_double(c["p_left"]),
                ctypes.c_double(c["t_left"]), ctypes.c_double(c["p_right"]),
                ctypes.c_double(c["t_right"]), ctypes.c_int64(hist_cap), hist.ctypes.data_as(_f),
                max_dump, ctypes.c_int64(nnz_cap), rp.ctypes.data_as(_u), col.ctypes.data_as(_u),
                ctypes.byref(nnz), dv.ctypes.data_as(_f), db.ctypes.data_as(_f),
                dx0.ctypes.data_as(_f), dx.ctypes.data_as(_f), dm.ctypes.data_as(_ip),
                p.ctypes.data_as(_f), u.ctypes.data_as(_f), T.ctypes.data_as(_f),
                rho.ctypes.data_as(_f), h.ctypes.data_as(_f))
        finally:
            os.chdir(cwd)
    assert cnt > 0, cnt
    k = int(nnz.value)
    return dict(count=cnt, res_hist=hist[:cnt], row_ptr=rp, col=col[:k], dump_val=dv[:max_dump, :k],
                dump_b=db[:max_dump], dump_x0=dx0[:max_dump], dump_x=dx[:max_dump],
                dump_method=dm[:max_dump], p=p, u=u.reshape(n, 3), T=T, rho=rho, h=h)


def sod_cases():
    meta = dict(SOD)
    meta["note"] = ("1D shock tube, 1000 cells over 1 m, Ly = Lz = 0.1, slip walls at both ends, "
                    "empty side faces, inviscid, blended central convection, transient, "
                    "10 inner iterations per step; R = 287.05 as the reference's air constants")
    # solver_mode 0 = the configured method for every equation (the reference);
    # 1 = CG for the pressure correction, ILU-BiCGSTAB otherwise (patched dispatch).
    for tag, st, sm, steps, dump in (("cg", 0, 0, 2, 10), ("bicgstab", 1, 0, 2, 10),
                                     ("split", 0, 1, 2, 10)):
        r = sod_run(steps, st, sm, dump)
        _save(f"sod_{tag}_2steps", **r)
    # the all-CG configuration diverges to NaN within the first steps (CG on the
    # nonsymmetric momentum systems); its 2-step fixture above records that.
    for tag, st, sm in (("bicgstab", 1, 0), ("split", 0, 1)):
        r = sod_run(200, st, sm, 0)
        _save(f"sod_{tag}_200steps", **{k: v for k, v in r.items() if not k.startswith("dump")})
    with open(os.path.join(HERE, "sod_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libocsl_ref.so missing: make -C oracle ref (needs the reference sources)")
    solver_cases()
    sod_cases()
