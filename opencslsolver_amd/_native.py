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
