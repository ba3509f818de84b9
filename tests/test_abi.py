"""The C-ABI libraries load and export every symbol their headers declare
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "opencslsolver_amd")


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ocsl_[A-Za-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("ocsl.h", "libocsl_hip.so"), ("ocsl_flow.h", "libocsl_hip.so"),
                                        ("ocsl_gen.h", "libocsl_gen.so"), ("ocsl_mesh.h", "libocsl_mesh.so")])
def test_exports(header, lib):
    path = os.path.join(PKG, lib)
    assert os.path.exists(path), f"{lib} not built"
    so = ctypes.CDLL(path)
    names = _declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    from opencslsolver_amd import _abi

    assert sorted(_abi.SIGNATURES) == _declared("ocsl.h")


def test_flow_signatures_cover_header():
    from opencslsolver_amd import flow

    assert sorted(flow.SIGNATURES) == _declared("ocsl_flow.h")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no host compiler")
def test_flow_struct_layouts(tmp_path):
    """ctypes mirrors of the ABI structs have the C compiler's sizes and field
    offsets (a C program over the headers prints them)."""
    import ctypes

    from opencslsolver_amd import _abi, flow

    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "ocsl_flow.h"\n'
        'int main(void) {\n'
        '  printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(ocsl_bc), sizeof(ocsl_mesh), sizeof(ocsl_flow_state),\n'
        '         sizeof(ocsl_flow_params), offsetof(ocsl_flow_params, solve_options),\n'
        '         offsetof(ocsl_flow_params, cell_order), sizeof(ocsl_options), sizeof(ocsl_report));\n'
        '  return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == [ctypes.sizeof(flow.BC), ctypes.sizeof(flow.Mesh), ctypes.sizeof(flow.State),
                   ctypes.sizeof(flow.Params), flow.Params.solve_options.offset, flow.Params.cell_order.offset,
                   ctypes.sizeof(_abi.Options), ctypes.sizeof(_abi.Report)]


def test_no_device_is_reported_not_fatal():
    from opencslsolver_amd import solver

    assert solver.device_count() >= 0
