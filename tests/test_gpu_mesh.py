"""SU2-read meshes on the device path: the matrix pattern from the device
structure builder against the oracle's restatement of
CFVM::define_matrix_structure, and the GPU assembly of every equation on
tetrahedral, prism and hexahedral meshes against the oracle's scatter loops."""
import numpy as np
import pytest

import oracle
from opencslsolver_amd import flow, mesh
from opencslsolver_amd.solver import build_structure

from test_gpu_flow import _close, _random_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,n,seed", [("tet", 3, 21), ("prism", 4, 22), ("hex", 5, 23)])
def test_assembly_on_su2_mesh_matches_oracle(kind, n, seed):
    oracle.set_threads(1)
    m = mesh.parse_su2(mesh.box_su2(n, kind, h=0.02, seed=seed))
    N = m["cell_volume"].shape[0]
    print(f"su2 {kind} n={n} seed={seed}: {N} cells", flush=True)
    rp, col = build_structure(N, m["face_left"], m["face_right"], m["bc_left"])
    rp_o, col_o = oracle.define_matrix_structure(N, m["face_left"], m["face_right"], m["bc_left"])
    assert np.array_equal(rp, rp_o) and np.array_equal(col, col_o)
    bcs = [dict(type=flow.VELOCITY_INLET, velocity=(3.0, 0.5, -0.2), temperature=300.0),
           dict(type=flow.PRESSURE_OUTLET, pressure=1e5),
           dict(type=flow.WALL, no_slip=1), dict(type=flow.WALL, no_slip=0),
           dict(type=flow.SYMMETRY), dict(type=flow.SUPERSONIC_INLET, velocity=(400.0, 0.0, 0.0), pressure=1e5,
                                          temperature=300.0)]
    prm = flow.default_params(viscous=1, blended=1, transient=1, bcs=bcs)
    rng = np.random.default_rng(seed)
    with flow.Flow(m, rp, col, prm) as fl:
        st = _random_state(rng, fl.n, fl.F, fl.B)
        for eq in range(5):
            fl.set_state(**st)
            val, b, x0 = fl.assemble(eq)
            v_o, b_o, x_o, _, _ = oracle.assemble(eq, m, rp, col, prm, st)
            assert _close(val, v_o, 1e-12), (kind, eq)
            assert _close(b, b_o, 1e-12), (kind, eq)
            assert np.array_equal(x0, x_o), (kind, eq)


def test_simple_steps_on_su2_mesh_and_vtk_output(tmp_path):
    """A prism box read from SU2 text runs SIMPLE iterations (inlet, outlet,
    walls, symmetry) with device-built structure; the fields written to VTK
    at full precision read back as the flow state."""
    from test_mesh import VTK_NAMES, _read_vtk

    text = mesh.box_su2(6, "prism", h=0.01, seed=41)
    m = mesh.parse_su2(text)
    N = m["cell_volume"].shape[0]
    print(f"su2 prism box: {N} cells", flush=True)
    rp, col = build_structure(N, m["face_left"], m["face_right"], m["bc_left"])
    bcs = [dict(type=flow.VELOCITY_INLET, velocity=(3.0, 0.0, 0.0), temperature=300.0),
           dict(type=flow.PRESSURE_OUTLET, pressure=1e5),
           dict(type=flow.WALL, no_slip=1), dict(type=flow.WALL, no_slip=1),
           dict(type=flow.SYMMETRY), dict(type=flow.SYMMETRY)]
    prm = flow.default_params(viscous=1, blended=1, transient=1, dt=1e-4, p0=1e5, t0=300.0, u0=(2.0, 0.0, 0.0),
                              bcs=bcs, method_pressure=0)
    with flow.Flow(m, rp, col, prm) as fl:
        fl.initialize()
        fl.run(3, 2)
        st = fl.get_state()
    assert all(np.all(np.isfinite(st[k])) for k in ("p", "T", "rho", "u"))
    assert np.all(st["rho"] > 0) and np.abs(st["u"][:, 0]).max() > 0
    path = str(tmp_path / "flow.vtk")
    fields = mesh.vtk_fields(st, list(prm.sutherland))
    mesh.write_vtk(text, path, fields, precision=17)
    _, _, _, data = _read_vtk(path)
    assert list(data) == VTK_NAMES
    assert np.array_equal(data["Pressure[Pa]"], st["p"]) and np.array_equal(data["Velocity[m/s]"], st["u"])
    assert np.array_equal(data["Temperature[K]"], st["T"]) and np.array_equal(data["MACH"], st["mach"])
