"""Writes tests/golden/vtk_sod40.npz: the VTK file the reference itself writes
at the end of a run (Croot::start_solver, croot.cpp:239-240, through
Cpost_processing::write_VTK_file / write_VTK_data, cpost_processing.cpp:26-127,
260-333) -- a 40-cell Sod tube, 2 steps -- with the mesh and cell fields it
was written from (ref_su2_driver.cpp's ref_last_run_dump), all from
oracle/_ref/libocsl_ref.so, which oracle/Makefile builds from the reference's
sources.

    python tests/golden/make_vtk_golden.py
"""
import ctypes
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

import oracle  # noqa: E402
from make_golden import SOD  # noqa: E402

N, STEPS = 40, 2


def main():
    c = dict(SOD)
    r = oracle.ref()
    f64 = ctypes.POINTER(ctypes.c_double)
    u64 = ctypes.POINTER(ctypes.c_uint64)
    i32 = ctypes.POINTER(ctypes.c_int32)
    nnz_cap, hist_cap = 5 * N, STEPS * c["inner"] * 5
    hist, rp, col = np.zeros(hist_cap), np.zeros(N + 1, np.uint64), np.zeros(nnz_cap, np.uint64)
    nnz = ctypes.c_int64()
    dv, db, dx0, dx = np.zeros((1, nnz_cap)), np.zeros((1, N)), np.zeros((1, N)), np.zeros((1, N))
    dm = np.zeros(1, np.int32)
    p, T, rho, h, u = np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(3 * N)
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as tmp:
        os.chdir(tmp)  # the reference writes solution.vtk into the working directory
        try:
            cnt = r.ref_sod_run(
                N, ctypes.c_double(c["lx"]), ctypes.c_double(c["ly"]), ctypes.c_double(c["lz"]), STEPS, c["inner"],
                ctypes.c_double(c["dt"]), 0, 0, c["conv_type"], c["visc_type"], ctypes.c_double(c["p_left"]),
                ctypes.c_double(c["t_left"]), ctypes.c_double(c["p_right"]), ctypes.c_double(c["t_right"]),
                ctypes.c_int64(hist_cap), hist.ctypes.data_as(f64), 0, ctypes.c_int64(nnz_cap),
                rp.ctypes.data_as(u64), col.ctypes.data_as(u64), ctypes.byref(nnz), dv.ctypes.data_as(f64),
                db.ctypes.data_as(f64), dx0.ctypes.data_as(f64), dx.ctypes.data_as(f64), dm.ctypes.data_as(i32),
                p.ctypes.data_as(f64), u.ctypes.data_as(f64), T.ctypes.data_as(f64), rho.ctypes.data_as(f64),
                h.ctypes.data_as(f64))
            assert cnt > 0, cnt
            text = open(os.path.join(tmp, "solution.vtk"), "rb").read()
        finally:
            os.chdir(cwd)
    nn, nc = ctypes.c_int64(), ctypes.c_int64()
    r.ref_last_run_sizes(ctypes.byref(nn), ctypes.byref(nc))
    n_nodes, n_cells = nn.value, nc.value
    out = dict(nodes=np.zeros((n_nodes, 3)), cell_type=np.zeros(n_cells, np.int32),
               cell_nodes=np.zeros((n_cells, 8), np.int32), density=np.zeros(n_cells), pressure=np.zeros(n_cells),
               temperature=np.zeros(n_cells), viscosity=np.zeros(n_cells), mach=np.zeros(n_cells),
               velocity=np.zeros((n_cells, 3)), pressure_grad=np.zeros((n_cells, 3)))
    r.ref_last_run_dump(*(out[k].ctypes.data_as(ctypes.c_void_p) for k in (
        "nodes", "cell_type", "cell_nodes", "density", "pressure", "temperature", "viscosity", "mach", "velocity",
        "pressure_grad")))
    np.savez_compressed(os.path.join(HERE, "vtk_sod40.npz"), vtk_text=np.frombuffer(text, np.uint8), **out)
    print({k: v.shape for k, v in out.items()}, len(text), "bytes of VTK")


if __name__ == "__main__":
    main()
