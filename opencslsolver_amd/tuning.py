"""Persistent cache of the CG launch autotuner (ocsl_tune_cg).

A tuned launch (kernels per iteration, rows per lane, threads per workgroup,
SpMV and vector workgroups) depends on the system's size and pattern and on the precision; it
is cached under a key of (rows, stored entries, pattern hash, precision) in a
JSON file, so later processes skip the tuning runs.
"""
from __future__ import annotations

import json
import os
import zlib

import numpy as np

from .solver import Solver

DEFAULT_PATH = os.path.join(os.path.expanduser("~"), ".cache", "opencslsolver_amd", "cg_tuning.json")


def _key(solver: Solver, precision: str) -> str:
    pat = zlib.crc32(np.ascontiguousarray(solver.row_ptr).tobytes())
    pat = zlib.crc32(np.ascontiguousarray(solver.col).tobytes(), pat)
    return f"v2|n={solver.n}|nnz={solver.nnz}|pattern={pat:08x}|{precision}"


def tuned_cg(solver: Solver, precision: str = "mixed", iterations: int = 64, path: str | None = None,
             retune: bool = False) -> dict:
    """Applies the cached launch for this system, tuning (and caching) it first
    when absent.  Returns the launch fields with the measured us/iteration."""
    path = path or os.environ.get("OCSL_TUNE_CACHE", DEFAULT_PATH)
    cache = {}
    if os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    key = _key(solver, precision)
    if key in cache and not retune:
        entry = cache[key]
        solver.set_tuned_cg(precision, **{k: entry[k] for k in ("cg_kernels", "rows_per_lane", "block_size",
                                                               "grid_blocks", "vec_blocks")})
        return entry
    best, us, _ = solver.tune_cg(precision, iterations)
    entry = dict(best, us_per_iteration=us)
    cache[key] = entry
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1)
    os.replace(tmp, path)
    return entry
