"""Writes tests/golden/vtk_sod40.npz and vtk_duct64.npz: the two VTK files the
reference itself writes at the end of a run (Croot::start_solver,
croot.cpp:235-240): the cell file (Cpost_processing::write_VTK_file /
write_VTK_data, cpost_processing.cpp:26-127, 260-333) and the boundary-face
file (write_boundary_VTK_file / write_boundary_VTK_data, :129-193, 335-401),
with the mesh, cell fields and boundary-face data they were written from
(ref_su2_driver.cpp's ref_last_run_dump and ref_last_run_bc_dump); a
40-cell Sod tube (inviscid, slip walls: no wall stress) and the duct64 flow
case (viscous, velocity inlet, no-slip walls, outlet); all from
oracle/_ref/libocsl_ref.so, which oracle/Makefile builds from the
reference's sources with its own flags, and the duct case once more from
oracle/_ref/ieee/libocsl_ref.so (no contraction, no fast-math:
vtk_duct64_ieee.npz), whose geometry rounds as IEEE arithmetic in the
written order does.

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

import make_golden  # noqa: E402
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
            bc_text = open(os.path.join(tmp, "solution_BC.vtk"), "rb").read()
        finally:
            os.chdir(cwd)
    out = last_run_arrays(r)
    np.savez_compressed(os.path.join(HERE, "vtk_sod40.npz"), vtk_text=np.frombuffer(text, np.uint8),
                        vtk_bc_text=np.frombuffer(bc_text, np.uint8), **out)
    print({k: v.shape for k, v in out.items()}, len(text), len(bc_text), "bytes of VTK")
    # a viscous run with no-slip walls, an inlet and an outlet, from the
    # reference's own build and from the build without contraction
    for variant, name in (("", "vtk_duct64.npz"), ("ieee", "vtk_duct64_ieee.npz")):
        fr = make_golden.flow_run(make_golden.FLOW_CASES["duct64"], 0, 0, keep_vtk=True, variant=variant)
        out = last_run_arrays(oracle.ref(variant))
        np.savez_compressed(os.path.join(HERE, name),
                            vtk_text=np.frombuffer(fr["vtk"]["solution.vtk"], np.uint8),
                            vtk_bc_text=np.frombuffer(fr["vtk"]["solution_BC.vtk"], np.uint8), **out)
        print({k: v.shape for k, v in out.items()}, name)


def last_run_arrays(r):
    """Mesh, cell fields and boundary-face data of the reference's last run."""
    f64 = ctypes.POINTER(ctypes.c_double)
    i32 = ctypes.POINTER(ctypes.c_int32)
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
    nb = ctypes.c_int64()
    r.ref_last_run_bc_size(ctypes.byref(nb))
    B = nb.value
    bc = dict(bc_nn=np.zeros(B, np.int32), bc_nodes=np.zeros((B, 4), np.int32), bc_marker=np.zeros(B, np.int32),
              bc_mach=np.zeros(B), bc_stress=np.zeros((B, 3)))
    r.ref_last_run_bc_dump(bc["bc_nn"].ctypes.data_as(i32), bc["bc_nodes"].ctypes.data_as(i32),
                           bc["bc_marker"].ctypes.data_as(i32), bc["bc_mach"].ctypes.data_as(f64),
                           bc["bc_stress"].ctypes.data_as(f64))
    out.update(bc)
    return out


if __name__ == "__main__":
    main()
