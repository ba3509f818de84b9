"""The flow handle's Hilbert-curve cell order (opencslsolver_amd/csrc/hip/
ocsl_order.h) checked on the host: tests/cpp/order_check.cpp compiles the
header the library is built from and checks the curve's defining property
(a bijection of the 2^b cube whose consecutive indices are face-adjacent
points, b = 1..6), the cell centres rebuilt from the faces of a randomly
numbered box, and the order of whole boxes (every stored cell a face
neighbour of the one before it on a 32^3 box)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "order_check.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_cell_order_curve(tmp_path):
    exe = os.path.join(str(tmp_path), "order_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror",
                    "-I" + os.path.join(ROOT, "opencslsolver_amd", "csrc", "hip"), SRC, "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "order ok" in r.stdout
