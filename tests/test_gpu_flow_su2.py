"""SIMPLE runs on 3D meshes read from SU2 files against the reference's own
runs on the same files (tests/golden/flow_su2_*.npz, written by
tests/golden/make_su2_flow_golden.py from the reference build): hexahedral,
prism and tetrahedral boxes, six markers (velocity inlet, pressure outlet,
no-slip wall, slip wall, two symmetry planes), viscous, blended, transient,
CG for the pressure correction and ILU(0)-BiCGSTAB for the rest, 2 steps of 3
inner iterations.  The mesh comes through libocsl_mesh.so, the matrix
structure from the device builder (equal to the reference's), the run through
the device-resident flow handle; fields and the per-solve residual history
must match to 1e-10 (the two reference builds differ by at most 6e-12 here)."""
import json
import os

import numpy as np
import pytest

import parity
from opencslsolver_amd import flow, mesh
from opencslsolver_amd.solver import build_structure

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def su2_params(meta):
    """ocsl_flow_params of a fixture's run (one boundary condition per marker;
    CG for the pressure correction, ILU(0)-BiCGSTAB for the rest)."""
    bcs = [dict(type=meta["bc_type"][m], no_slip=meta["bc_no_slip"][m], const_temperature=meta["bc_const_T"][m],
                temperature=meta["bc_T"][m], pressure=meta["bc_p"][m], velocity=tuple(meta["bc_u"][m]))
           for m in range(len(meta["bc_type"]))]
    return flow.default_params(
        viscous=0 if meta["visc_type"] else 1, blended=meta["conv_type"], transient=0 if meta["steady"] else 1,
        dt=meta["dt"], implicit_underrelax=meta["implicit_underrelax"],
        implicit_underrelax_max=meta["implicit_underrelax_max"], ramp_iterations=meta["ramp_iterations"],
        shock_tube=meta["init_shock_tube"], p0=meta["p0"], t0=meta["T0"], u0=tuple(meta["u0"]), p2=meta["p2"],
        t2=meta["T2"], bcs=bcs, method_momentum=1, method_pressure=0, method_enthalpy=1)


@pytest.mark.parametrize("kind", ["hex", "prism", "tet"])
def test_su2_flow_run_matches_reference(kind):
    d = np.load(os.path.join(GOLDEN, f"flow_su2_{kind}.npz"))
    meta = json.loads(d["meta"].tobytes())
    m = mesh.parse_su2(d["su2_text"].tobytes().decode())
    n = m["cell_volume"].shape[0]
    print(f"su2 {kind}: {n} cells", flush=True)
    rp, col = build_structure(n, m["face_left"], m["face_right"], m["bc_left"])
    assert np.array_equal(rp, d["row_ptr"]) and np.array_equal(col, d["col"])
    with flow.Flow(m, rp, col, su2_params(meta)) as fl:
        fl.initialize()
        fl.run(meta["steps"], meta["inner"])
        st = fl.get_state()
        hist = fl.solve_history()
    for f in ("p", "T", "rho", "h", "u"):
        parity.check(f"su2 {kind} {f}", parity.rel(st[f], d[f]), parity.bar("su2_runs", kind, f))
    assert len(hist) == int(d["count"])
    parity.check(f"su2 {kind} residual history", parity.hist_rel(hist, d["res_hist"]),
                 parity.bar("su2_runs", kind, "hist"))
