"""CPU oracle for the OpenCSL sparse-Krylov hot path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()``
(as the checker) and ``bench.py``'s cpu_baseline leg (timed); the product
path never imports it.

``liboracle.so`` is the plain-C restatement (ocsl_oracle.c, IEEE, no
contraction) used as the checker; ``liboracle_fast.so`` is the same source
with the reference's compile flags, used as the CPU baseline.  ``ref()``
loads ``_ref/libocsl_ref.so``, the reference solver compiled from its own
sources (only where those sources exist; see oracle/Makefile).

Parity status of each function (how it is pinned) is listed in
tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes
import functools
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_d = ctypes.c_double
_i64 = ctypes.c_int64


def _p(a, t=_f64p):
    return None if a is None else a.ctypes.data_as(t)


@functools.lru_cache(maxsize=None)
def _lib(name: str) -> ctypes.CDLL:
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run make -C oracle")
    lib = ctypes.CDLL(path)
    for fn in ("or_norm", "or_sum", "or_abssum", "or_cg_reference", "or_ilu_bicgstab_reference"):
        getattr(lib, fn).restype = ctypes.c_double
    lib.or_transpose.restype = ctypes.c_int64
    return lib


def lib(fast: bool = False) -> ctypes.CDLL:
    return _lib("liboracle_fast.so" if fast else "liboracle.so")


def _ref_path(variant: str = "") -> str:
    """The reference build: "" its own flags (make -C oracle ref), "ieee"
    without contraction or fast-math (make -C oracle ref-ieee)."""
    return os.path.join(_HERE, "_ref", variant, "libocsl_ref.so") if variant else \
        os.path.join(_HERE, "_ref", "libocsl_ref.so")


def ref_available(variant: str = "") -> bool:
    return os.path.exists(_ref_path(variant))


@functools.lru_cache(maxsize=None)
def ref(variant: str = "") -> ctypes.CDLL:
    path = _ref_path(variant)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run make -C oracle ref (needs the reference sources)")
    r = ctypes.CDLL(path)
    for fn in ("ref_cg_solve", "ref_ilu_bicgstab_solve", "ref_norm", "ref_abssum"):
        getattr(r, fn).restype = ctypes.c_double
    r.ref_sod_run.restype = ctypes.c_int64
    return r


class RefFlowConfig(ctypes.Structure):
    """Run settings of the reference drivers' flow runs (RefFlowConfig in
    ref_sod_driver.cpp)."""
    _fields_ = [("n_cells", ctypes.c_int), ("lx", ctypes.c_double), ("ly", ctypes.c_double),
                ("lz", ctypes.c_double), ("steps", ctypes.c_int), ("inner", ctypes.c_int),
                ("dt", ctypes.c_double), ("steady", ctypes.c_int), ("implicit_underrelax", ctypes.c_double),
                ("implicit_underrelax_max", ctypes.c_double), ("ramp_iterations", ctypes.c_int),
                ("solver_type", ctypes.c_int), ("solver_mode", ctypes.c_int), ("conv_type", ctypes.c_int),
                ("visc_type", ctypes.c_int), ("init_shock_tube", ctypes.c_int), ("p0", ctypes.c_double),
                ("T0", ctypes.c_double), ("u0", ctypes.c_double * 3), ("p2", ctypes.c_double),
                ("T2", ctypes.c_double), ("bc_type", ctypes.c_int * 3), ("bc_no_slip", ctypes.c_int * 3),
                ("bc_const_T", ctypes.c_int * 3), ("bc_T", ctypes.c_double * 3), ("bc_p", ctypes.c_double * 3),
                ("bc_u", (ctypes.c_double * 3) * 3)]


def ref_flow_config(settings: dict) -> RefFlowConfig:
    c = RefFlowConfig()
    for k, v in settings.items():
        if k == "u0":
            for d in range(3):
                c.u0[d] = v[d]
        else:
            setattr(c, k, v)
    return c


def ref_flow_time_file(path: str, settings: dict, bcs: dict, threads: int, n_cells: int) -> dict:
    """The reference's SIMPLE run on the SU2 file at `path` (one boundary
    condition per marker: bcs holds bc_type, bc_no_slip, bc_const_T, bc_T,
    bc_p, bc_u lists), on `threads` OpenMP threads, timed by the reference's
    own per-iteration clock (ref_sod_driver.cpp's ref_flow_time_file).  The
    reference writes its VTK files into the current directory.  Returns
    iter_us, total_s, the number of solves and the final fields."""
    c = ref_flow_config(settings)
    c.n_cells = n_cells
    nb = len(bcs["bc_type"])
    i32 = lambda a: (ctypes.c_int * nb)(*a)  # noqa: E731
    f64 = lambda a: (ctypes.c_double * len(a))(*a)  # noqa: E731
    p, T, rho, h = (np.zeros(n_cells) for _ in range(4))
    u = np.zeros(3 * n_cells)
    it_us, tot_s = ctypes.c_double(), ctypes.c_double()
    r = ref()
    r.ref_flow_time_file.restype = ctypes.c_int64
    cnt = r.ref_flow_time_file(ctypes.byref(c), path.encode(), nb, i32(bcs["bc_type"]), i32(bcs["bc_no_slip"]),
                               i32(bcs["bc_const_T"]), f64(bcs["bc_T"]), f64(bcs["bc_p"]),
                               f64([v for u3 in bcs["bc_u"] for v in u3]), int(threads), _i64(n_cells),
                               ctypes.byref(it_us), ctypes.byref(tot_s), _p(p), _p(u), _p(T), _p(rho), _p(h))
    if cnt <= 0:
        raise RuntimeError(f"ref_flow_time_file returned {cnt}")
    return dict(iter_us=it_us.value, total_s=tot_s.value, solves=int(cnt), p=p, u=u.reshape(n_cells, 3), T=T,
                rho=rho, h=h)


# The shock tube of the reference fixtures (ref_sod_run): 1000 cells over 1 m,
# Ly = Lz = 0.1, slip walls at both ends, empty side faces, inviscid, blended
# central convection, transient dt 1e-6, 10 inner iterations per step.
SOD_SETTINGS = dict(n_cells=1000, lx=1.0, ly=0.1, lz=0.1, dt=1e-6, p_left=1e5, t_left=348.4, p_right=1e4,
                    t_right=278.7, conv_type=1, visc_type=1, inner=10)


def ref_sod_time(steps: int, threads: int, solver_type: int = 0, solver_mode: int = 1) -> dict:
    """The reference's shock-tube run (SOD_SETTINGS; solver_mode 1: CG for the
    pressure correction, ILU(0)-BiCGSTAB otherwise) on `threads` OpenMP
    threads, timed by the reference's own per-iteration clock.  The reference
    writes its VTK files into the current directory.  Returns iter_us,
    total_s, the number of solves and the final fields."""
    import time

    c = SOD_SETTINGS
    n = c["n_cells"]
    nnz_cap = 5 * n
    hist = np.zeros(1)
    rp, col = np.zeros(n + 1, np.uint64), np.zeros(nnz_cap, np.uint64)
    nnz = ctypes.c_int64()
    dummy = np.zeros(1)
    dm = np.zeros(1, np.int32)
    p, T, rho, h = (np.zeros(n) for _ in range(4))
    u = np.zeros(3 * n)
    r = ref()
    r.ref_last_iteration_us.restype = ctypes.c_double
    r.ref_set_threads(int(threads))
    t0 = time.perf_counter()
    try:
        cnt = r.ref_sod_run(n, _d(c["lx"]), _d(c["ly"]), _d(c["lz"]), steps, c["inner"], _d(c["dt"]), solver_type,
                            solver_mode, c["conv_type"], c["visc_type"], _d(c["p_left"]), _d(c["t_left"]),
                            _d(c["p_right"]), _d(c["t_right"]), _i64(0), _p(hist), 0, _i64(nnz_cap),
                            _p(rp, _u64p), _p(col, _u64p), ctypes.byref(nnz), _p(dummy), _p(dummy), _p(dummy),
                            _p(dummy), _p(dm, _ip), _p(p), _p(u), _p(T), _p(rho), _p(h))
    finally:
        r.ref_set_threads(1)
    total = time.perf_counter() - t0
    if cnt <= 0:
        raise RuntimeError(f"ref_sod_run returned {cnt}")
    return dict(iter_us=r.ref_last_iteration_us(), total_s=total, solves=int(cnt), p=p, u=u.reshape(n, 3), T=T,
                rho=rho, h=h)


def set_threads(t: int) -> None:
    lib().or_set_threads(int(t))
    lib(True).or_set_threads(int(t))


def _csr(a):
    return _i64(a.n), _p(a.row_ptr, _u64p), _p(a.col, _u64p)


# --- BLAS / SpMV ---------------------------------------------------------

def norm(a, b, fast=False):
    return lib(fast).or_norm(_i64(a.size), _p(a), _p(b))


def abssum(a):
    return lib().or_abssum(_i64(a.size), _p(a))


def gemv(A, x, alpha=1.0, beta=0.0, z=None):
    out = np.empty(A.n)
    lib().or_gemv(*_csr(A), _p(A.val), _d(alpha), _p(x), _d(beta), _p(z), _p(out))
    return out


def diag_index(A):
    d = np.empty(A.n, np.int64)
    lib().or_diag_index(*_csr(A), _p(d, _i64p))
    return d


def ilu_decompose(A, sweeps=3):
    m = np.empty(A.nnz)
    lib().or_ilu_decompose(*_csr(A), _p(A.val), _p(diag_index(A), _i64p), sweeps, _p(m))
    return m


def ilu_exact(A):
    m = np.empty(A.nnz)
    lib().or_ilu_exact(*_csr(A), _p(A.val), _p(diag_index(A), _i64p), _p(m))
    return m


def forsub(A, m, b):
    y = np.empty(A.n)
    lib().or_forsub(*_csr(A), _p(m), _p(diag_index(A), _i64p), _p(b), _p(y))
    return y


def backsub(A, m, y):
    x = np.empty(A.n)
    lib().or_backsub(*_csr(A), _p(m), _p(diag_index(A), _i64p), _p(y), _p(x))
    return x


def forsub_jacobi(A, m, b, sweeps=3, w=0.9, y0=None):
    y = np.zeros(A.n) if y0 is None else np.array(y0, dtype=np.float64)
    lib().or_forsub_jacobi(*_csr(A), _p(m), _p(diag_index(A), _i64p), _p(b), sweeps, _d(w), _p(y))
    return y


def backsub_jacobi(A, m, y, sweeps=3, w=0.9, x0=None):
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64)
    lib().or_backsub_jacobi(*_csr(A), _p(m), _p(diag_index(A), _i64p), _p(y), sweeps, _d(w), _p(x))
    return x


# --- solvers ---------------------------------------------------------------

def cg_reference(A, b, x0):
    """CG_solve (cmatrix_solver.cpp:635-680); returns (res, x, history)."""
    x = np.array(x0, dtype=np.float64)
    h = np.zeros(8)
    nh = ctypes.c_int()
    res = lib().or_cg_reference(*_csr(A), _p(A.val), _p(b), _p(x), _p(h), ctypes.byref(nh))
    return res, x, h[: nh.value].copy()


def ilu_bicgstab_reference(A, b, x0):
    """ILU_BICGSTAB_solve (cmatrix_solver.cpp:682-762); returns (res, x, history)."""
    x = np.array(x0, dtype=np.float64)
    h = np.zeros(8)
    nh = ctypes.c_int()
    res = lib().or_ilu_bicgstab_reference(*_csr(A), _p(A.val), _p(b), _p(x), _p(h),
                                          ctypes.byref(nh))
    return res, x, h[: nh.value].copy()


def cg_tol(A, b, x0=None, tol=1e-8, maxit=100000, fast=False):
    """Textbook CG to ||r||/||b|| <= tol; returns (iters, x, history, true_relres)."""
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64)
    h = np.zeros(maxit + 1)
    tr = ctypes.c_double()
    it = lib(fast).or_cg_tol(*_csr(A), _p(A.val), _p(b), _p(x), _d(tol), maxit, _p(h),
                             ctypes.byref(tr))
    return it, x, h[: it + 1].copy(), tr.value


def bicgstab_tol(A, b, x0=None, tol=1e-10, maxit=10000, ilu_sweeps=3, tri_sweeps=0, tri_w=0.9,
                 fast=False):
    """Right-preconditioned ILU(0)-BiCGSTAB; returns (iters, x, history, true_relres)."""
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64)
    h = np.zeros(maxit + 1)
    tr = ctypes.c_double()
    it = lib(fast).or_bicgstab_tol(*_csr(A), _p(A.val), _p(b), _p(x), _d(tol), maxit, ilu_sweeps,
                                   tri_sweeps, _d(tri_w), _p(h), ctypes.byref(tr))
    return it, x, h[: it + 1].copy(), tr.value


def transpose(A):
    from opencslsolver_amd.gen import CSR

    nnz_real = int(np.count_nonzero(A.col < np.uint64(A.n)))
    rp = np.empty(A.n + 1, np.uint64)
    col = np.empty(nnz_real, np.uint64)
    val = np.empty(nnz_real)
    lib().or_transpose(*_csr(A), _p(A.val), _p(rp, _u64p), _p(col, _u64p), _p(val))
    return CSR(rp, col, val)


def sampled_outer(A, lam, x):
    out = np.empty(A.nnz)
    lib().or_sampled_outer(*_csr(A), _p(lam), _p(x), _p(out))
    return out


def assemble(eq, mesh, row_ptr, col, params, state, urf=0.5, fast=False):
    """The reference's face/boundary/cell loops for one equation (0..2
    momentum, 3 pressure correction, 4 enthalpy) on a mesh dict and state
    dict; returns (val, b, x0, ap, ap_vec) -- ap / ap_vec updated as the
    momentum cell loop does."""
    from opencslsolver_amd import flow

    m, _keep = flow.mesh_struct(mesh)
    n, F, B = m.n_cells, m.n_faces, m.n_bc
    fields = dict(state)
    fields.setdefault("ap", np.zeros(n))
    fields.setdefault("ap_vec", np.zeros((n, 3)))
    st, arrays = flow.state_struct(fields, n, F, B)
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint64)
    cl = np.ascontiguousarray(col, dtype=np.uint64)
    val = np.zeros(cl.shape[0])
    b = np.zeros(n)
    x0 = np.zeros(n)
    rc = lib(fast).or_assemble(int(eq), ctypes.byref(m), ctypes.byref(st), ctypes.byref(params), _d(urf),
                               _p(rp, _u64p), _p(cl, _u64p), _p(val), _p(b), _p(x0))
    if rc != 0:
        raise ValueError(f"or_assemble failed: {rc}")
    return val, b, x0, arrays["ap"], arrays["ap_vec"].reshape(n, 3)


# --- reference build (_ref) --------------------------------------------------

def ref_su2_arrays(path: str) -> dict:
    """The reference's mesh set-up on one SU2 file -- Read_SU2_mesh
    (cgeometry_supply.h:201-433), CFVM::create_FVM_from_geometry
    (cfvm.cpp:100-170), CFVM::define_matrix_structure (cfvm.cpp:74-98) --
    as arrays (ref_su2_driver.cpp).  Its progress lines go to stderr."""
    import sys

    r = ref()
    sz = (ctypes.c_int64 * 5)()
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        r.ref_su2_read(str(path).encode(), sz)
    finally:
        os.dup2(saved, 1)
        os.close(saved)
    N, F, B, nnz, _ = list(sz)
    spec = [("cell_volume", N, np.float64), ("cell_center", (N, 3), np.float64),
            ("face_left", F, np.int32), ("face_right", F, np.int32),
            ("face_area_vector", (F, 3), np.float64), ("face_unit", (F, 3), np.float64),
            ("face_dist", F, np.float64), ("face_alpha", (F, 3), np.float64), ("face_ksi", F, np.float64),
            ("face_geo_aver", F, np.float64), ("face_area", F, np.float64),
            ("bc_left", B, np.int32), ("bc_marker", B, np.int32), ("bc_area_vector", (B, 3), np.float64),
            ("bc_unit", (B, 3), np.float64), ("bc_dist", B, np.float64), ("bc_alpha", (B, 3), np.float64),
            ("bc_ksi", B, np.float64), ("bc_perp", B, np.float64), ("bc_area", B, np.float64),
            ("row_ptr", N + 1, np.uint64), ("col", nnz, np.uint64)]
    out = {k: np.empty(shape, dt) for k, shape, dt in spec}
    r.ref_su2_copy(*(out[k].ctypes.data_as(ctypes.c_void_p) for k, _, _ in spec))
    return out


def ref_su2_stage_times(path: str) -> tuple:
    """Seconds of the reference's three mesh set-up stages on one SU2 file,
    timed inside the reference library (so its loading is not counted):
    Read_SU2_mesh, create_FVM_from_geometry, define_matrix_structure.  Its
    progress lines go to stderr."""
    import sys

    r = ref()
    sz = (ctypes.c_int64 * 5)()
    st = (ctypes.c_double * 3)()
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        r.ref_su2_read_timed(str(path).encode(), sz, st)
    finally:
        os.dup2(saved, 1)
        os.close(saved)
    return tuple(st)


def ref_cg_timing(A, b, x0, threads: int, calls: int = 5) -> tuple:
    """Median seconds of one call of the reference's CG_solve (its fixed 4
    iterations, cmatrix_solver.cpp:635-680) and of that call's set-up alone,
    with `threads` OpenMP threads (ref_driver.cpp)."""
    call, setup = ctypes.c_double(), ctypes.c_double()
    ref().ref_cg_timing(*_csr(A), _p(A.val), _p(b), _p(np.ascontiguousarray(x0, np.float64)),
                        int(threads), int(calls), ctypes.byref(call), ctypes.byref(setup))
    return call.value, setup.value


def ref_cg_solve(A, b, x0, variant=""):
    x = np.array(x0, dtype=np.float64)
    res = ref(variant).ref_cg_solve(*_csr(A), _p(A.val), _p(b), _p(x))
    return res, x


def ref_ilu_bicgstab_solve(A, b, x0, variant=""):
    x = np.array(x0, dtype=np.float64)
    res = ref(variant).ref_ilu_bicgstab_solve(*_csr(A), _p(A.val), _p(b), _p(x))
    return res, x


def ref_ilu_decompose(A):
    m = np.empty(A.nnz)
    ref().ref_ilu_decompose(*_csr(A), _p(A.val), _p(m))
    return m


def ref_forsub(A, m, b):
    y = np.empty(A.n)
    ref().ref_forsub(*_csr(A), _p(m), _p(b), _p(y))
    return y


def ref_backsub(A, m, y):
    x = np.empty(A.n)
    ref().ref_backsub(*_csr(A), _p(m), _p(y), _p(x))
    return x


def ref_gemv(A, x, alpha=1.0, beta=0.0, z=None):
    out = np.empty(A.n)
    ref().ref_gemv(*_csr(A), _p(A.val), _d(alpha), _p(x), _d(beta), _p(z), _p(out))
    return out


SENTINEL_COL = (1 << 64) - 1


def define_matrix_structure(n_cells, face_left, face_right, bc_left):
    """CSR structure of CFVM::define_matrix_structure (cfvm.cpp:74-98).

    Every cell i inserts (i, i), then (i, j) for each entry j of its neighbour
    list as Cgeometry builds it (cgeometry.cpp:602-612: per face the cell
    across it, -1 for a boundary face; the -1 converts to the size_t column
    2^64 - 1).  Insertion follows CCSR::insert_element (cmatrix_solver.cpp:
    57-100): a column already in the row is only overwritten; otherwise it
    goes before the first larger column of the row, or at the row's end.  An
    insertion into row i changes no other row's columns (it only shifts later
    row pointers), so rows are built one at a time here.  Pure-Python loops:
    small meshes only.  Returns (row_ptr, col) as uint64 arrays."""
    n = int(n_cells)
    neigh = [[] for _ in range(n)]
    for l, r in zip(np.asarray(face_left).tolist(), np.asarray(face_right).tolist()):
        neigh[l].append(r)
        neigh[r].append(l)
    for c in np.asarray(bc_left).tolist():
        neigh[c].append(-1)
    row_ptr = [0]
    cols = []
    for i in range(n):
        row = []

        def insert(j):
            j &= SENTINEL_COL
            if j in row:
                return
            for k, c in enumerate(row):
                if j < c:
                    row.insert(k, j)
                    return
            row.append(j)

        insert(i)
        for j in neigh[i]:
            insert(j)
        cols.extend(row)
        row_ptr.append(len(cols))
    return np.array(row_ptr, np.uint64), np.array(cols, np.uint64)
