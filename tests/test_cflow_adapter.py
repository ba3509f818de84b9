"""The CFlow_solver-side adapter (include/ocsl_cflow_adapter.hpp).

tests/cpp/cflow_adapter_check.cpp builds objects shaped like the reference's
CFVM and CFlow_solver from a reference run's recorded mesh and flow states
(the golden flow fixtures), hands them to the adapter, and prints what comes
back.  On the CPU the AoS -> flat arrays -> AoS round trip must be exact; on
the GPU, FlowAssembly::assemble on each recorded state must give the
reference's own assembled system (values, right-hand side, initial guess, and
ap / ap_vec for momentum) to the assembly bar of the direct tests, 1e-12."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import parity
from flowcases import CASES, FLOW_META, GOLDEN, params_for, records, state_at

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "cflow_adapter_check.cpp")
LIBDIR = os.path.join(ROOT, "opencslsolver_amd")
MESH = ("cell_volume", "face_left", "face_right", "face_area_vector", "face_unit", "face_alpha", "face_area",
        "face_dist", "face_ksi", "face_geo_aver", "bc_left", "bc_marker", "bc_area_vector", "bc_unit",
        "bc_alpha", "bc_area", "bc_dist", "bc_ksi", "bc_perp")
STATE = ("u", "u_old", "rho", "rho_old", "p", "p_old", "p_corr", "grad_p", "T", "h", "h_old", "ap", "ap_vec",
         "dt", "f_mdot", "f_rho", "f_qy", "f_stress", "b_T", "b_p", "b_rho", "b_mdot", "b_h", "b_qy", "b_u",
         "b_supersonic")


def _build(tmp_path):
    exe = os.path.join(str(tmp_path), "cflow_adapter_check")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), SRC, "-o", exe,
           "-L" + LIBDIR, "-locsl_hip", "-Wl,-rpath," + LIBDIR, "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def _line(f, name, a):
    a = np.asarray(a, dtype=np.float64).ravel()
    f.write(f"{name} {a.size} " + " ".join(repr(float(v)) for v in a) + "\n")


def _write_input(path, case):
    d = np.load(os.path.join(GOLDEN, f"flow_{case}.npz"))
    n = d["mesh_cell_volume"].size
    k_max = records(case, d)
    with open(path, "w") as f:
        for k in MESH:
            _line(f, "mesh_" + k, d["mesh_" + k])
        _line(f, "row_ptr", d["row_ptr"].astype(np.float64))
        col = d["col"].astype(np.float64)
        col[d["col"] >= np.uint64(n)] = -1.0  # the sentinel column
        _line(f, "col", col)
        _line(f, "records", [k_max])
        f.write(f"params_hex {len(bytes(params_for(FLOW_META[case])))} "
                f"{bytes(params_for(FLOW_META[case])).hex()}\n")
        for k in range(k_max):
            _, st = state_at(d, k, n)
            for name in STATE:
                key = "b_super" if name == "b_supersonic" else name
                _line(f, f"state{k}_{key}", st[name])
    return d, k_max


def _run(exe, path, mode):
    r = subprocess.run([exe, path, mode], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    out = {}
    for line in r.stdout.splitlines():
        name, cnt, *vals = line.split()
        assert int(cnt) == len(vals), name
        out[name] = np.array([float(v) for v in vals])
    return out


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
@pytest.mark.parametrize("case", CASES)
def test_adapter_round_trip_is_exact(tmp_path, case):
    exe = _build(tmp_path)
    path = os.path.join(str(tmp_path), "in.txt")
    d, k_max = _write_input(path, case)
    out = _run(exe, path, "roundtrip")
    n = d["mesh_cell_volume"].size
    assert np.array_equal(out["rt_face_unit"], d["mesh_face_unit"].ravel())
    assert np.array_equal(out["rt_bc_perp"], d["mesh_bc_perp"])
    assert np.array_equal(out["rt_bc_marker"], d["mesh_bc_marker"])
    for k in range(k_max):
        _, st = state_at(d, k, n)
        for name in ("u", "h", "ap_vec", "f_stress", "b_u"):
            assert np.array_equal(out[f"rt{k}_{name}"], np.asarray(st[name], float).ravel()), (k, name)
        assert np.array_equal(out[f"rt{k}_b_super"], np.asarray(st["b_supersonic"], float))


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_adapter_assembles_the_reference_systems(tmp_path, case):
    exe = _build(tmp_path)
    path = os.path.join(str(tmp_path), "in.txt")
    d, k_max = _write_input(path, case)
    out = _run(exe, path, "assemble")
    worst = {"val": 0.0, "b": 0.0, "ap": 0.0}
    for k in range(k_max):
        worst["val"] = max(worst["val"], parity.rel(out[f"k{k}_val"], d["dump_val"][k]))
        worst["b"] = max(worst["b"], parity.rel(out[f"k{k}_b"], d["dump_b"][k]))
        assert np.array_equal(out[f"k{k}_x0"], d["dump_x0"][k]), k
        eq = k % 5
        if eq <= 2:
            worst["ap"] = max(worst["ap"], parity.rel(out[f"k{k}_ap"], d["state_ap"][k]))
            apv = out[f"k{k}_ap_vec"].reshape(-1, 3)[:, eq]
            worst["ap"] = max(worst["ap"], parity.rel(apv, d["state_ap_vec"][k][:, eq]))
    for name, v in worst.items():
        parity.check(f"adapter {case} {name}", v, 1e-12)
