"""Python side of include/ocsl_flow.h: the pressure-based flow solver whose
coefficient assembly runs on the GPU (CFlow_solver's face/cell loops,
cflow_solver_cell_face_loops.cpp), whose linear systems go to the HIP
solvers of the same library, and whose remaining SIMPLE-step loops run as
device kernels over a device-resident state (or, with support=SUPPORT_HOST,
as a host restatement kept as a cross-check)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from ._native import hip_lib
from .solver import _check

_i32 = ctypes.c_int32
_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)

# ocsl_bc_type (enums.h:38 order)
EMPTY, SYMMETRY, WALL, VELOCITY_INLET, SUPERSONIC_INLET, PRESSURE_OUTLET = range(6)
MOMENTUM_X, MOMENTUM_Y, MOMENTUM_Z, PRESSURE, ENTHALPY = range(5)
# ocsl_flow_cell_order: how the device stores cells
CELL_ORDER_AUTO, CELL_ORDER_CALLER, CELL_ORDER_LOCALITY, CELL_ORDER_CURVE = range(4)
# ocsl_flow_support: where the SIMPLE step's support loops run
SUPPORT_DEVICE, SUPPORT_HOST = range(2)
MAX_BC = 16

MESH_FLOAT = ("face_area_vector", "face_unit", "face_alpha", "face_area", "face_dist", "face_ksi",
              "face_geo_aver")
MESH_BC_FLOAT = ("bc_area_vector", "bc_unit", "bc_alpha", "bc_area", "bc_dist", "bc_ksi", "bc_perp")


class Mesh(ctypes.Structure):
    _fields_ = ([("n_cells", ctypes.c_int64), ("n_faces", ctypes.c_int64), ("n_bc", ctypes.c_int64),
                 ("cell_volume", _f64p), ("face_left", _i32p), ("face_right", _i32p)]
                + [(k, _f64p) for k in MESH_FLOAT]
                + [("bc_left", _i32p), ("bc_marker", _i32p)]
                + [(k, _f64p) for k in MESH_BC_FLOAT])


class BC(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("no_slip", ctypes.c_int), ("const_temperature", ctypes.c_int),
                ("temperature", ctypes.c_double), ("pressure", ctypes.c_double),
                ("velocity", ctypes.c_double * 3)]


class Params(ctypes.Structure):
    _fields_ = [("viscous", ctypes.c_int), ("blended", ctypes.c_int), ("transient", ctypes.c_int),
                ("dt", ctypes.c_double), ("implicit_underrelax", ctypes.c_double),
                ("implicit_underrelax_max", ctypes.c_double), ("ramp_iterations", ctypes.c_int),
                ("gamma", ctypes.c_double), ("molar_mass", ctypes.c_double), ("t_ref", ctypes.c_double),
                ("h_formation", ctypes.c_double), ("sutherland", ctypes.c_double * 2),
                ("prandtl", ctypes.c_double), ("shock_tube", ctypes.c_int), ("p0", ctypes.c_double),
                ("t0", ctypes.c_double), ("u0", ctypes.c_double * 3), ("p2", ctypes.c_double),
                ("t2", ctypes.c_double), ("n_bc", ctypes.c_int), ("bc", BC * MAX_BC),
                ("method_momentum", ctypes.c_int), ("method_pressure", ctypes.c_int),
                ("method_enthalpy", ctypes.c_int), ("mode", ctypes.c_int), ("precision", ctypes.c_int),
                ("support", ctypes.c_int), ("solve_options", _abi.Options), ("cell_order", ctypes.c_int),
                ("momentum_together", ctypes.c_int)]


CELL_FIELDS = {"u": 3, "u_old": 3, "rho": 1, "rho_old": 1, "p": 1, "p_old": 1, "p_corr": 1, "grad_p": 3,
               "grad_pc": 3, "T": 1, "h": 1, "h_old": 1, "ap": 1, "ap_vec": 3, "dt": 1, "mach": 1}
FACE_FIELDS = {"f_mdot": 1, "f_rho": 1, "f_qy": 1, "f_stress": 3}
BC_FIELDS = {"b_T": 1, "b_p": 1, "b_p_out": 1, "b_rho": 1, "b_mdot": 1, "b_h": 1, "b_qy": 1, "b_u": 3}


class State(ctypes.Structure):
    _fields_ = ([(k, _f64p) for k in CELL_FIELDS] + [(k, _f64p) for k in FACE_FIELDS]
                + [(k, _f64p) for k in BC_FIELDS] + [("b_supersonic", _i32p)])


SIGNATURES = {
    "ocsl_flow_default_params": (None, [ctypes.POINTER(Params)]),
    "ocsl_flow_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(Mesh), _u64p, _u64p,
                                        ctypes.POINTER(Params)]),
    "ocsl_flow_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ocsl_flow_solver": (ctypes.c_void_p, [ctypes.c_void_p]),
    "ocsl_flow_initialize": (ctypes.c_int, [ctypes.c_void_p]),
    "ocsl_flow_set_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(State)]),
    "ocsl_flow_get_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(State)]),
    "ocsl_flow_assemble": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int]),
    "ocsl_flow_viscous_stress": (ctypes.c_int, [ctypes.c_void_p]),
    "ocsl_flow_time_assemble": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _f64p, _f64p]),
    "ocsl_flow_iterate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ocsl_flow_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "ocsl_flow_solve_history": (ctypes.c_int, [ctypes.c_void_p, _f64p, ctypes.c_int64,
                                               ctypes.POINTER(ctypes.c_int64)]),
    "ocsl_flow_cell_order_get": (ctypes.c_int, [ctypes.c_void_p, _i32p]),
}


def _lib():
    lib = hip_lib()
    if not getattr(lib, "_ocsl_flow_declared", False):
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib._ocsl_flow_declared = True
    return lib


def default_params(**kw) -> Params:
    p = Params()
    _lib().ocsl_flow_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k == "bcs":
            p.n_bc = len(v)
            for m, bc in enumerate(v):
                b = p.bc[m]
                b.type = bc.get("type", EMPTY)
                b.no_slip = bc.get("no_slip", 0)
                b.const_temperature = bc.get("const_temperature", 0)
                b.temperature = bc.get("temperature", 0.0)
                b.pressure = bc.get("pressure", 0.0)
                for d in range(3):
                    b.velocity[d] = bc.get("velocity", (0.0, 0.0, 0.0))[d]
        elif k == "u0":
            for d in range(3):
                p.u0[d] = v[d]
        elif k == "solve_options":
            for ok, ov in v.items():
                setattr(p.solve_options, ok, ov)
        else:
            if not hasattr(p, k):
                raise TypeError(f"unknown flow parameter {k}")
            setattr(p, k, v)
    return p


def mesh_struct(mesh: dict):
    """(Mesh, arrays kept alive) of a mesh dict."""
    n = int(np.asarray(mesh["cell_volume"]).shape[0])
    F = int(np.asarray(mesh["face_left"]).shape[0])
    B = int(np.asarray(mesh["bc_left"]).shape[0])
    keep = {}
    m = Mesh(n_cells=n, n_faces=F, n_bc=B)
    for k in ("cell_volume",) + MESH_FLOAT + MESH_BC_FLOAT:
        a = np.ascontiguousarray(mesh[k], dtype=np.float64)
        keep[k] = a
        setattr(m, k, a.ctypes.data_as(_f64p))
    for k in ("face_left", "face_right", "bc_left", "bc_marker"):
        a = np.ascontiguousarray(mesh[k], dtype=np.int32)
        keep[k] = a
        setattr(m, k, a.ctypes.data_as(_i32p))
    return m, keep


def state_struct(fields: dict, n: int, F: int, B: int):
    """(State, arrays) of a dict of state fields; missing fields are NULL.
    The arrays are private float64/int32 copies the C side may write."""
    st = State()
    arrays = {}
    sizes = {**{k: n * c for k, c in CELL_FIELDS.items()}, **{k: F * c for k, c in FACE_FIELDS.items()},
             **{k: B * c for k, c in BC_FIELDS.items()}}
    for k, v in fields.items():
        if k == "b_supersonic":
            a = np.array(v, dtype=np.int32).reshape(-1)
            assert a.shape[0] == B
            st.b_supersonic = a.ctypes.data_as(_i32p)
        else:
            a = np.array(v, dtype=np.float64).reshape(-1)
            assert a.shape[0] == sizes[k], (k, a.shape, sizes[k])
            setattr(st, k, a.ctypes.data_as(_f64p))
        arrays[k] = a
    return st, arrays


class Flow:
    """CFlow_solver on one mesh: GPU assembly + GPU linear solves."""

    def __init__(self, mesh: dict, row_ptr, col, params: Params):
        self._lib = _lib()
        self.n = int(np.asarray(mesh["cell_volume"]).shape[0])
        self.F = int(np.asarray(mesh["face_left"]).shape[0])
        self.B = int(np.asarray(mesh["bc_left"]).shape[0])
        m, self._mesh_keep = mesh_struct(mesh)
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint64)
        self.col = np.ascontiguousarray(col, dtype=np.uint64)
        self.nnz = int(self.col.shape[0])
        h = ctypes.c_void_p()
        _check(self._lib.ocsl_flow_create(ctypes.byref(h), ctypes.byref(m), self.row_ptr.ctypes.data_as(_u64p),
                                          self.col.ctypes.data_as(_u64p), ctypes.byref(params)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ocsl_flow_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _sizes(self):
        out = {}
        for k, c in CELL_FIELDS.items():
            out[k] = (self.n, c)
        for k, c in FACE_FIELDS.items():
            out[k] = (self.F, c)
        for k, c in BC_FIELDS.items():
            out[k] = (self.B, c)
        return out

    def initialize(self):
        _check(self._lib.ocsl_flow_initialize(self._h))

    def set_state(self, **fields):
        st, _keep = state_struct(fields, self.n, self.F, self.B)
        _check(self._lib.ocsl_flow_set_state(self._h, ctypes.byref(st)))

    def get_state(self) -> dict:
        st = State()
        out = {}
        for k, (n, c) in self._sizes().items():
            a = np.zeros((n, c) if c > 1 else n)
            out[k] = a
            setattr(st, k, a.ctypes.data_as(_f64p))
        sup = np.zeros(self.B, np.int32)
        out["b_supersonic"] = sup
        st.b_supersonic = sup.ctypes.data_as(_i32p)
        _check(self._lib.ocsl_flow_get_state(self._h, ctypes.byref(st)))
        return out

    def assemble(self, equation: int):
        """(values in the caller's slot layout, b, x0) of one equation."""
        val, b, x0 = np.empty(self.nnz), np.empty(self.n), np.empty(self.n)
        _check(self._lib.ocsl_flow_assemble(self._h, equation, val.ctypes.data_as(ctypes.c_void_p),
                                            b.ctypes.data_as(ctypes.c_void_p), x0.ctypes.data_as(ctypes.c_void_p),
                                            _abi.HOST))
        return val, b, x0

    def time_assemble(self, equation: int, reps: int = 20):
        """(average ms, algorithmic bytes) of one assembly launch of an equation."""
        ms, by = ctypes.c_double(), ctypes.c_double()
        _check(self._lib.ocsl_flow_time_assemble(self._h, equation, reps, ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value

    def viscous_stress(self):
        _check(self._lib.ocsl_flow_viscous_stress(self._h))

    def iterate(self, end_of_step=False):
        _check(self._lib.ocsl_flow_iterate(self._h, int(end_of_step)))

    def run(self, steps: int, inner: int):
        _check(self._lib.ocsl_flow_run(self._h, steps, inner))

    def solve_history(self) -> np.ndarray:
        ln = ctypes.c_int64()
        _check(self._lib.ocsl_flow_solve_history(self._h, None, 0, ctypes.byref(ln)))
        out = np.empty(ln.value)
        _check(self._lib.ocsl_flow_solve_history(self._h, out.ctypes.data_as(_f64p), ln.value, ctypes.byref(ln)))
        return out

    def cell_order(self) -> np.ndarray:
        """The device's cell order: entry k is the caller's cell stored k-th."""
        out = np.empty(self.n, dtype=np.int32)
        _check(self._lib.ocsl_flow_cell_order_get(self._h, out.ctypes.data_as(_i32p)))
        return out


def mesh_from_npz(d, prefix="mesh_") -> dict:
    """Mesh arrays stored with a prefix (as in the golden fixtures)."""
    keys = ("cell_volume", "face_left", "face_right", "bc_left", "bc_marker") + MESH_FLOAT + MESH_BC_FLOAT
    return {k: d[prefix + k] for k in keys}
