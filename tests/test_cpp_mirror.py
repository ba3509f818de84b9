"""The C++ drop-in for CMatrix_solver (include/ocsl_cmatrix_solver.hpp).

tests/cpp/cmatrix_solver_dropin.cpp is a caller written in the reference's
own call forms (CFVM::define_matrix_structure, the assembly loops'
insert/add/get calls, CMatrix_solver::ABSsum(b_vector), solve(), and every
other public member of cmatrix_solver.h:27-95).  It must compile and link
against the header unchanged; on the CPU its structure must equal the
oracle's restatement of define_matrix_structure; on the GPU every result must
match the oracle on the same system."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
import parity
from opencslsolver_amd.gen import CSR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "cmatrix_solver_dropin.cpp")
LIBDIR = os.path.join(ROOT, "opencslsolver_amd")
UINT = ("row_ptr", "col")
INT = ("face_left", "face_right", "bc_left", "diag_index")


def _build(tmp_path):
    exe = os.path.join(str(tmp_path), "cmatrix_solver_dropin")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"), SRC,
           "-o", exe, "-L" + LIBDIR, "-locsl_hip", "-Wl,-rpath," + LIBDIR, "-L/opt/rocm/lib",
           "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def _run(exe, *args):
    out = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    d = {}
    for line in out.stdout.splitlines():
        name, cnt, *vals = line.split()
        if name == "n_rows":
            d[name] = tuple(int(v) for v in [cnt, *vals])
            continue
        assert int(cnt) == len(vals), name
        if name in UINT:
            d[name] = np.array([int(v) for v in vals], dtype=np.uint64)
        elif name in INT:
            d[name] = np.array([int(v) for v in vals], dtype=np.int64)
        else:
            d[name] = np.array([float(v) for v in vals])
    return d


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_dropin_builds_the_reference_structure(tmp_path):
    """cfvm.cpp:84-97 through the header's CCSR::insert_element: the pattern
    (sorted rows, the diagonal, each neighbour once, one sentinel column per
    boundary row), the reserved n_nnz and extract_diag_indexes()."""
    d = _run(_build(tmp_path), "structure")
    n = len(d["x0"])
    rp, col = oracle.define_matrix_structure(n, d["face_left"], d["face_right"], d["bc_left"])
    assert np.array_equal(d["row_ptr"], rp)
    assert np.array_equal(d["col"], col)
    assert d["n_rows"] == (n, n, 6 * n)  # each cell lists its 6 faces' neighbours
    for i in range(n):
        assert col[d["diag_index"][i]] == i
    a = CSR(d["row_ptr"], d["col"], d["val"])
    sentinel = d["col"] == np.uint64(oracle.SENTINEL_COL)
    assert np.count_nonzero(sentinel) == len(np.unique(d["bc_left"]))
    assert np.all(d["val"][sentinel] == 0.0)
    assert a.nnz == len(d["val"])


@pytest.mark.gpu
def test_dropin_matches_oracle_on_gpu(tmp_path):
    d = _run(_build(tmp_path))
    a = CSR(d["row_ptr"], d["col"], d["val"])
    b, x0 = d["b"], d["x0"]
    oracle.set_threads(1)
    parity.check("dropin ABSsum", parity.res_rel(d["abssum_b"][0], oracle.abssum(b), floor=0), 1e-13)
    parity.check("dropin sum", abs(d["sum_b"][0] - np.sum(b)) / np.sum(np.abs(b)), 1e-13)
    parity.check("dropin norm", parity.res_rel(d["norm_bx"][0], oracle.norm(b, x0), floor=0), 1e-13)
    parity.check("dropin gemv", parity.rel(d["gemv_ax"], oracle.gemv(a, x0)), 1e-14)
    parity.check("dropin gemv residual", parity.rel(d["gemv_r"], oracle.gemv(a, x0, -1.0, 1.0, b)), 1e-14)
    # elementwise: each product rounded, then the sum (exact IEEE)
    assert np.array_equal(d["axpy1"], 2.5 * b)
    assert np.array_equal(d["axpy2"], 0.5 * b - x0)
    assert np.array_equal(d["product"], b * x0)
    m = oracle.ilu_decompose(a, 3)
    parity.check("dropin ILU_decompose", parity.rel(d["ilu_m"], m), 1e-12)
    y = oracle.forsub(a, m, b)
    parity.check("dropin ILU_forsub", parity.rel(d["forsub_y"], y), 1e-12)
    parity.check("dropin ILU_backsub", parity.rel(d["backsub_x"], oracle.backsub(a, m, y)), 1e-12)
    r_cg, x_cg, _ = oracle.cg_reference(a, b, x0)
    parity.check("dropin CG_solve residual", parity.res_rel(d["cg_res"][0], r_cg, floor=0), 1e-10)
    parity.check("dropin CG_solve x", parity.rel(d["cg_x"], x_cg), 1e-10)
    r_bi, x_bi, _ = oracle.ilu_bicgstab_reference(a, b, x0)
    parity.check("dropin ILU_BICGSTAB_solve residual", parity.res_rel(d["bicg_res"][0], r_bi), 1e-10)
    parity.check("dropin ILU_BICGSTAB_solve x", parity.rel(d["bicg_x"], x_bi), 1e-10)
    parity.check("dropin factor left in M", parity.rel(d["bicg_M"], m), 1e-12)
    # solve() on the re-assembled system with method CG: CG_solve's result
    assert np.array_equal(d["solve_cg_x"], d["cg_x"])
    assert d["tol_status"][0] == 0
    relres = np.linalg.norm(b - a.matvec(d["tol_x"])) / np.linalg.norm(b)
    parity.check("dropin TOLERANCE solve host relres", relres, 1e-12)
