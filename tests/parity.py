"""Parity metrics and bars shared by the tests that compare with the
reference's golden vectors.

Every bar is 1e-10 relative unless rounding alone moves the quantity further:
tests/golden/spread.json holds, per quantity, the distance between two builds
of the reference itself (its own -O3 -ffast-math -march=native build, which
wrote the fixtures, and -O2 -ffp-contract=off), measured with the metrics
below; the bar is then five times that distance.  Each check prints the
achieved value and its bar, and appends them as a JSON line to the file named
by OCSL_PARITY_LOG when that is set.
"""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FLOOR_BAR = 1e-10
SPREAD_FACTOR = 5.0
# ||r||^2 / ||b||^2 below this is rounding noise (relative residual 1e-12)
RES_FLOOR = 1e-24
# per-solve residual histories: entries below this are rounding noise
HIST_FLOOR = 1e-14

_SPREAD = None


def spread(*keys) -> float:
    global _SPREAD
    if _SPREAD is None:
        with open(os.path.join(GOLDEN, "spread.json")) as f:
            _SPREAD = json.load(f)
    v = _SPREAD
    for k in keys:
        v = v[k]
    return float(v)


def bar(*keys) -> float:
    """1e-10, or five times the measured reference-vs-reference spread."""
    return max(FLOOR_BAR, SPREAD_FACTOR * spread(*keys))


def rel(a, b) -> float:
    """max |a - b| / max |b| (fields, solutions, factors)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def res_rel(a, ref, floor=RES_FLOOR) -> float:
    """A returned residual against the reference's: relative above the noise
    floor; below it both must be noise (0 if so, inf if not)."""
    a, ref = float(a), float(ref)
    if abs(ref) > floor:
        return abs(a - ref) / abs(ref)
    return 0.0 if abs(a) <= floor else float("inf")


def hist_rel(h, ref, floor=HIST_FLOOR) -> float:
    """Per-solve residual histories: the largest relative difference over the
    reference entries above the floor; entries at or below it must be noise on
    both sides (NaN where the reference has its 0/0 residual), else inf."""
    h, ref = np.asarray(h, float), np.asarray(ref, float)
    if h.shape != ref.shape:
        return float("inf")
    nan = np.isnan(h)
    if not np.all(np.isnan(ref[nan]) | (np.abs(ref[nan]) <= floor)):
        return float("inf")
    big = ~nan & ~np.isnan(ref) & (np.abs(ref) > floor)
    small = ~nan & ~big
    if np.any(np.abs(h[small]) >= 100.0 * floor):
        return float("inf")
    return float(np.max(np.abs(h[big] - ref[big]) / np.abs(ref[big]))) if big.any() else 0.0


def check(name: str, achieved: float, limit: float) -> None:
    print(f"parity {name}: {achieved:.3e} (bar {limit:.1e})", flush=True)
    path = os.environ.get("OCSL_PARITY_LOG")
    if path:
        with open(path, "a") as f:
            f.write(json.dumps({"check": name, "achieved": achieved, "bar": limit}) + "\n")
    assert achieved <= limit, f"{name}: {achieved:.3e} > bar {limit:.1e}"
