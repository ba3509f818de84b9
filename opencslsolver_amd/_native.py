"""Loading of the package's in-tree native libraries.

The HIP library is required: there is no CPU fallback anywhere in the product
path, so a missing or unloadable ``libocsl_hip.so`` raises immediately.
"""
from __future__ import annotations

import ctypes
import functools
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
GEN_LIB_PATH = os.path.join(_HERE, "libocsl_gen.so")
HIP_LIB_PATH = os.path.join(_HERE, "libocsl_hip.so")
MESH_LIB_PATH = os.path.join(_HERE, "libocsl_mesh.so")


class NativeLibraryMissing(RuntimeError):
    pass


def _load(path: str) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run __graft_entry__.build() or make -C opencslsolver_amd/csrc")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


@functools.lru_cache(maxsize=None)
def gen_lib() -> ctypes.CDLL:
    lib = _load(GEN_LIB_PATH)
    lib.ocsl_gen_hash.restype = ctypes.c_uint64
    return lib


@functools.lru_cache(maxsize=None)
def hip_lib() -> ctypes.CDLL:
    from . import _abi

    lib = _load(HIP_LIB_PATH)
    _abi.declare(lib)
    return lib


@functools.lru_cache(maxsize=None)
def mesh_lib() -> ctypes.CDLL:
    lib = _load(MESH_LIB_PATH)
    c = ctypes
    p = c.c_void_p
    lib.ocsl_mesh_read_su2.argtypes = [c.c_char_p, c.POINTER(p)]
    lib.ocsl_mesh_parse_su2.argtypes = [c.c_char_p, c.c_int64, c.POINTER(p)]
    lib.ocsl_mesh_destroy.argtypes = [p]
    lib.ocsl_mesh_destroy.restype = None
    lib.ocsl_mesh_last_error.restype = c.c_char_p
    lib.ocsl_mesh_sizes.argtypes = [p] + [c.POINTER(c.c_int64)] * 4 + [c.POINTER(c.c_int32)]
    lib.ocsl_mesh_marker_tag.argtypes = [p, c.c_int32, c.c_char_p, c.c_int64]
    lib.ocsl_mesh_cells.argtypes = [p, p, p]
    lib.ocsl_mesh_faces.argtypes = [p] + [p] * 9
    lib.ocsl_mesh_bc_faces.argtypes = [p] + [p] * 9
    lib.ocsl_mesh_write_vtk.argtypes = [p, c.c_char_p, p, c.c_int]
    lib.ocsl_mesh_write_boundary_vtk.argtypes = [p, c.c_char_p, p, c.c_int]
    return lib
