"""The C++ CMatrix_solver mirror (include/ocsl_cmatrix_solver.hpp) compiles
and links against the HIP library here; on a GPU it reproduces the
reference's unit-test results (oracle)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from opencslsolver_amd import gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "cmatrix_solver_example.cpp")
LIBDIR = os.path.join(ROOT, "opencslsolver_amd")


def _build(tmp_path):
    exe = os.path.join(str(tmp_path), "cmatrix_solver_example")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), SRC, "-o", exe,
           "-L" + LIBDIR, "-locsl_hip", "-Wl,-rpath," + LIBDIR, "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_mirror_compiles_and_links(tmp_path):
    assert os.path.exists(_build(tmp_path))


@pytest.mark.gpu
def test_mirror_reproduces_reference_unit_test(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout.split("\n")
    vals = {}
    xs = []
    for line in out:
        if not line:
            continue
        k, v = line.split()
        if k == "x":
            xs.append(float(v))
        else:
            vals[k] = float(v)
    a, b, x0 = gen.tridiag_unit_test_system()
    r_o, x_o, _ = oracle.cg_reference(a, b, x0)
    assert vals["cg"] == pytest.approx(r_o, rel=1e-10)
    assert np.allclose(xs[:10], x_o, rtol=1e-10, atol=0)
    r_b, x_b, _ = oracle.ilu_bicgstab_reference(a, b, x0)
    assert vals["bicgstab"] == pytest.approx(r_b, rel=1e-8, abs=1e-28)
    assert np.allclose(xs[10:], x_b, rtol=1e-9, atol=1e-14)
    assert vals["abssum"] == 20.0 and vals["norm"] == 20.0
