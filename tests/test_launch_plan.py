"""The tolerance-CG launch planning (opencslsolver_amd/csrc/hip/ocsl_launch.h)
checked on the host: tests/cpp/launch_plan_check.cpp compiles the header the
library is built from and sweeps it over every system size up to 4096 rows,
every n^3 system up to 130^3 and every class of launch option, checking the
kernels' index coverage against the allocations, the partial-sum slots, the
iteration chunks and the uniqueness of the graph-cache keys."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "launch_plan_check.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_launch_plan(tmp_path):
    exe = os.path.join(str(tmp_path), "launch_plan_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "opencslsolver_amd", "csrc", "hip"), SRC, "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "launch plan ok" in r.stdout
