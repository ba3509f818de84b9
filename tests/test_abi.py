"""The C-ABI libraries load and export every symbol their headers declare
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

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


def test_flow_struct_layouts():
    """ctypes mirrors of the flow structs have the C sizes (checked against
    offsets computed from the header's field list)."""
    import ctypes

    from opencslsolver_amd import _abi, flow

    assert ctypes.sizeof(flow.BC) == 3 * 4 + 4 + 5 * 8
    assert ctypes.sizeof(flow.Mesh) == 3 * 8 + 19 * 8
    assert ctypes.sizeof(flow.State) == 29 * 8
    assert flow.Params.solve_options.offset % 8 == 0
    assert ctypes.sizeof(flow.Params) == flow.Params.solve_options.offset + ctypes.sizeof(_abi.Options)


def test_no_device_is_reported_not_fatal():
    from opencslsolver_amd import solver

    assert solver.device_count() >= 0
