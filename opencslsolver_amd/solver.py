"""Python handle over the C ABI (include/ocsl.h).

``Solver`` owns one ``ocsl_handle``: a CSR structure in the reference's
layout (uint64 row_ptr / column_index, sentinel columns allowed) on the
current HIP device.  Data may be numpy arrays (host) or torch CUDA tensors
(device, read in place on the handle's stream after the caller's stream is
synchronised).  There is no CPU fallback: without a HIP device or without
``libocsl_hip.so`` every call raises.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from ._native import hip_lib

_f64p = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)

METHODS = {"cg": _abi.CG, "ilu_bicgstab": _abi.ILU_BICGSTAB, "bicgstab": _abi.ILU_BICGSTAB}
MODES = {"reference": _abi.MODE_REFERENCE, "tolerance": _abi.MODE_TOLERANCE}
PRECISIONS = {"fp64": _abi.FP64, "mixed": _abi.MIXED}


class OcslError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"ocsl status {status}: {what}")
        self.status = status


def _check(st: int, allow_warn: bool = False) -> int:
    if st < 0 or (st > 0 and not allow_warn):
        msg = hip_lib().ocsl_last_error()
        raise OcslError(st, msg.decode() if msg else "")
    return st


def _is_torch_cuda(a) -> bool:
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(a, torch.Tensor) and a.is_cuda


def _ptr(a, dtype=np.float64):
    """(pointer, where, keepalive) of a numpy array or a torch CUDA tensor."""
    if _is_torch_cuda(a):
        import torch

        if a.dtype != torch.float64 or not a.is_contiguous():
            raise TypeError("device arrays must be contiguous float64 tensors")
        torch.cuda.current_stream(a.device).synchronize()
        return ctypes.c_void_p(a.data_ptr()), _abi.DEVICE, a
    arr = np.ascontiguousarray(a, dtype=dtype)
    return arr.ctypes.data_as(ctypes.c_void_p), _abi.HOST, arr


def default_options(**kw) -> _abi.Options:
    o = _abi.Options()
    hip_lib().ocsl_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k}")
        setattr(o, k, v)
    return o


def device_count() -> int:
    return int(hip_lib().ocsl_device_count())


class Solver:
    """One CSR structure on the GPU with its solvers (CMatrix_solver's solve path)."""

    def __init__(self, row_ptr, column_index):
        self._lib = hip_lib()
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint64)
        self.col = np.ascontiguousarray(column_index, dtype=np.uint64)
        self.n = int(self.row_ptr.shape[0] - 1)
        self.nnz = int(self.col.shape[0])
        h = _abi.c_handle()
        _check(self._lib.ocsl_create(ctypes.byref(h), self.n, self.row_ptr.ctypes.data_as(_u64p),
                                     self.col.ctypes.data_as(_u64p)))
        self._h = h
        self._keep = {}

    # -- lifetime --------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.ocsl_destroy(self._h)
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

    @property
    def stream(self) -> int:
        return int(self._lib.ocsl_get_stream(self._h) or 0)

    # -- data --------------------------------------------------------------
    def set_values(self, val) -> None:
        p, where, keep = _ptr(val)
        self._keep["val"] = keep
        _check(self._lib.ocsl_set_values(self._h, p, where))

    def set_rhs(self, b) -> None:
        p, where, keep = _ptr(b)
        self._keep["b"] = keep
        _check(self._lib.ocsl_set_rhs(self._h, p, where))

    def set_x(self, x0) -> None:
        p, where, keep = _ptr(x0)
        _check(self._lib.ocsl_set_x(self._h, p, where))

    def get_x(self, out=None) -> np.ndarray:
        if out is not None and _is_torch_cuda(out):
            _check(self._lib.ocsl_get_x(self._h, ctypes.c_void_p(out.data_ptr()), _abi.DEVICE))
            return out
        x = np.empty(self.n)
        _check(self._lib.ocsl_get_x(self._h, x.ctypes.data_as(ctypes.c_void_p), _abi.HOST))
        return x

    # -- solves ------------------------------------------------------------
    def solve(self, method="cg", mode="tolerance", precision="fp64", allow_warn=True, **opts) -> dict:
        o = default_options(**opts)
        rep = _abi.Report()
        st = self._lib.ocsl_solve(self._h, METHODS[method], MODES[mode], PRECISIONS[precision],
                                  ctypes.byref(o), ctypes.byref(rep))
        _check(st, allow_warn=allow_warn)
        d = rep.as_dict()
        d["status"] = st
        return d

    def solve_multi(self, B, X0=None, method="ilu_bicgstab", precision="fp64", **opts):
        """k right-hand sides (rows of B, shape (k, n)) sharing the matrix.
        Returns (X of shape (k, n), list of k reports).  Torch CUDA tensors
        (float64, contiguous) are used in place on the device."""
        o = default_options(**opts)
        if _is_torch_cuda(B):
            import torch

            k = B.shape[0]
            X = torch.zeros_like(B) if X0 is None else X0
            _ptr(B)  # orders torch's stream before the call
            pB, pX, where = ctypes.c_void_p(B.data_ptr()), ctypes.c_void_p(X.data_ptr()), _abi.DEVICE
        else:
            B = np.ascontiguousarray(B, dtype=np.float64)
            k = B.shape[0]
            X = np.zeros_like(B) if X0 is None else np.array(X0, dtype=np.float64, order="C")
            pB, pX, where = B.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p), _abi.HOST
        reps = (_abi.Report * k)()
        st = self._lib.ocsl_solve_multi(self._h, METHODS[method], PRECISIONS[precision], k, pB, pX, where,
                                        ctypes.byref(o), reps)
        _check(st, allow_warn=True)
        return X, [r.as_dict() for r in reps]

    def set_reference_graphs(self, on: bool) -> None:
        """REFERENCE-mode solves replay one hipGraph per right-hand-side address
        after their first run (default), or issue every launch each time."""
        _check(self._lib.ocsl_set_reference_graphs(self._h, 1 if on else 0))

    def set_tri_form(self, form: int = 0):
        """Form of the exact triangular solves: 0 chosen from the level
        schedule, 1 one workgroup, 2 one launch per level, 3 serial over the
        rows staged in LDS; the results are bitwise the same."""
        _check(self._lib.ocsl_set_tri_form(self._h, int(form)))

    def time_tri_solve(self, lower=True, reps=20, form=0) -> float:
        """Average ms of one exact triangular solve of the 3-sweep ILU(0)
        factor (HIP events on the handle's stream); form 0 as the solvers
        choose it, 1 one workgroup, 2 one launch per level."""
        ms = ctypes.c_double()
        _check(self._lib.ocsl_time_tri_solve_form(self._h, 1 if lower else 0, int(form), reps, ctypes.byref(ms)))
        return ms.value

    def colour_order(self):
        """(order, n_colours) of colour-ordered BiCGSTAB solves: order[i] is the
        input row at position i."""
        nc = ctypes.c_int()
        out = np.empty(self.n, np.int32)
        _check(self._lib.ocsl_colour_order(self._h, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nc)))
        return out, nc.value

    def CG_solve(self) -> float:
        """CMatrix_solver::CG_solve (cmatrix_solver.cpp:635-680)."""
        res = ctypes.c_double()
        _check(self._lib.ocsl_CG_solve(self._h, ctypes.byref(res)), allow_warn=True)
        return res.value

    def ILU_BICGSTAB_solve(self) -> float:
        """CMatrix_solver::ILU_BICGSTAB_solve (cmatrix_solver.cpp:682-762)."""
        res = ctypes.c_double()
        _check(self._lib.ocsl_ILU_BICGSTAB_solve(self._h, ctypes.byref(res)), allow_warn=True)
        return res.value

    def ILU_BICGSTAB_solve_multi(self, B, X0):
        """ILU_BICGSTAB_solve for the k rows of B (shape (k, n)) at once, X0
        the k guesses: each row's solution and residual are bitwise those of
        ILU_BICGSTAB_solve with that right-hand side and guess.  Returns (X,
        residuals)."""
        B = np.ascontiguousarray(B, dtype=np.float64)
        k = B.shape[0]
        X = np.array(X0, dtype=np.float64, order="C")
        if X.shape != B.shape:
            raise ValueError("X0 must have the shape of B")
        res = np.zeros(k)
        st = self._lib.ocsl_ILU_BICGSTAB_solve_multi(self._h, k, B.ctypes.data_as(ctypes.c_void_p),
                                                     X.ctypes.data_as(ctypes.c_void_p), _abi.HOST,
                                                     res.ctypes.data_as(_f64p))
        _check(st, allow_warn=True)
        return X, res

    def history(self) -> np.ndarray:
        ln = ctypes.c_int()
        _check(self._lib.ocsl_get_history(self._h, None, 0, ctypes.byref(ln)))
        out = np.empty(ln.value)
        _check(self._lib.ocsl_get_history(self._h, out.ctypes.data_as(_f64p), ln.value,
                                          ctypes.byref(ln)))
        return out

    # -- reference helpers -------------------------------------------------
    def abs_sum(self, v) -> float:
        p, where, _k = _ptr(v)
        out = ctypes.c_double()
        _check(self._lib.ocsl_ABSsum(self._h, p, where, ctypes.byref(out)))
        return out.value

    def sum(self, v) -> float:
        p, where, _k = _ptr(v)
        out = ctypes.c_double()
        _check(self._lib.ocsl_sum(self._h, p, where, ctypes.byref(out)))
        return out.value

    def norm(self, a, b) -> float:
        pa, where, _ka = _ptr(a)
        pb, _w, _kb = _ptr(b)
        out = ctypes.c_double()
        _check(self._lib.ocsl_norm(self._h, pa, pb, where, ctypes.byref(out)))
        return out.value

    def gemv(self, a: float, x, b: float = 0.0, z=None) -> np.ndarray:
        px, _w, _kx = _ptr(x)
        pz, _w2, _kz = _ptr(z) if z is not None else (None, _abi.HOST, None)
        out = np.empty(self.n)
        _check(self._lib.ocsl_gemv(self._h, a, px, b, pz, out.ctypes.data_as(ctypes.c_void_p),
                                   _abi.HOST))
        return out

    def ilu_decompose(self, sweeps: int = 3) -> np.ndarray:
        out = np.empty(self.nnz)
        _check(self._lib.ocsl_ILU_decompose(self._h, sweeps, out.ctypes.data_as(ctypes.c_void_p),
                                            _abi.HOST))
        return out

    def forsub(self, m, b) -> np.ndarray:
        """CMatrix_solver::ILU_forsub with the factor m (slot order): y = L^-1 b."""
        m = np.ascontiguousarray(m, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        y = np.empty(self.n)
        _check(self._lib.ocsl_ILU_forsub(self._h, m.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p),
                                         y.ctypes.data_as(ctypes.c_void_p), _abi.HOST))
        return y

    def backsub(self, m, y) -> np.ndarray:
        """CMatrix_solver::ILU_backsub with the factor m (slot order): x = U^-1 y."""
        m = np.ascontiguousarray(m, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        x = np.empty(self.n)
        _check(self._lib.ocsl_ILU_backsub(self._h, m.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p),
                                          x.ctypes.data_as(ctypes.c_void_p), _abi.HOST))
        return x

    def time_cg_spmv(self, precision="fp64", reorder=1, reps=50, rows_per_lane=0):
        """(average ms, algorithmic bytes) of one CG SpMV launch on the handle's stream."""
        ms, by = ctypes.c_double(), ctypes.c_double()
        _check(self._lib.ocsl_time_cg_spmv(self._h, PRECISIONS[precision], reorder, rows_per_lane, reps,
                                           ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value

    def time_cg_iteration(self, precision="mixed", pairs=50):
        """One-kernel CG iterations of the tuned launch: ([even ms, odd ms],
        [even bytes, odd bytes]) per launch, from `pairs` timed pairs."""
        ms = (ctypes.c_double * 2)()
        by = (ctypes.c_double * 2)()
        _check(self._lib.ocsl_time_cg_iteration(self._h, PRECISIONS[precision], pairs, ms, by))
        return [ms[0], ms[1]], [by[0], by[1]]

    def tune_cg(self, precision="mixed", iterations=64):
        """Times every CG launch candidate on this system; keeps the fastest in
        the handle.  Returns (best options dict, best us/iteration, table)."""
        best = _abi.Options()
        us = ctypes.c_double()
        cap = 256
        table = (_abi.TuneEntry * cap)()
        ln = ctypes.c_int()
        _check(self._lib.ocsl_tune_cg(self._h, PRECISIONS[precision], iterations, ctypes.byref(best),
                                      ctypes.byref(us), table, cap, ctypes.byref(ln)))
        keys = ("cg_kernels", "rows_per_lane", "block_size", "grid_blocks", "vec_blocks")
        rows = [dict({k: getattr(table[i], k) for k in keys}, us_per_iteration=table[i].us_per_iteration)
                for i in range(min(ln.value, cap))]
        return {k: getattr(best, k) for k in keys}, us.value, rows

    def set_tuned_cg(self, precision="mixed", rows_per_lane=0, block_size=0, grid_blocks=0, vec_blocks=0,
                     cg_kernels=0):
        o = default_options(rows_per_lane=rows_per_lane, block_size=block_size, grid_blocks=grid_blocks,
                            vec_blocks=vec_blocks, cg_kernels=cg_kernels)
        _check(self._lib.ocsl_set_tuned_cg(self._h, PRECISIONS[precision], ctypes.byref(o)))

    # -- differentiation ---------------------------------------------------
    def adjoint(self, g, method="cg", precision="fp64", dLdb=None, dLdA=None, **opts):
        """Adjoint of the last solve for a loss with dL/dx = g.

        Solves A^T lambda = g; returns (dL/db = lambda, dL/dA on the stored
        pattern, report).  Torch CUDA tensors (float64) stay on the device;
        pass ``dLdb``/``dLdA`` tensors to receive the results in place."""
        o = default_options(**opts)
        rep = _abi.Report()
        pg, where, _kg = _ptr(g)
        if where == _abi.DEVICE:
            import torch

            if dLdb is None:
                dLdb = torch.empty(self.n, dtype=torch.float64, device=g.device)
            if dLdA is None:
                dLdA = torch.empty(self.nnz, dtype=torch.float64, device=g.device)
            pb, pa = ctypes.c_void_p(dLdb.data_ptr()), ctypes.c_void_p(dLdA.data_ptr())
        else:
            dLdb = np.empty(self.n) if dLdb is None else dLdb
            dLdA = np.empty(self.nnz) if dLdA is None else dLdA
            pb, pa = dLdb.ctypes.data_as(ctypes.c_void_p), dLdA.ctypes.data_as(ctypes.c_void_p)
        st = self._lib.ocsl_adjoint(self._h, METHODS[method], PRECISIONS[precision], pg, pb, pa, where,
                                    ctypes.byref(o), ctypes.byref(rep))
        _check(st, allow_warn=True)
        d = rep.as_dict()
        d["status"] = st
        return dLdb, dLdA, d

    def sampled_outer(self, alpha, u, v) -> np.ndarray:
        """out[p] = alpha * u[i] * v[j] on the stored pattern (caller slot order)."""
        pu, _w, _ku = _ptr(u)
        pv, _w2, _kv = _ptr(v)
        out = np.empty(self.nnz)
        _check(self._lib.ocsl_sampled_outer(self._h, alpha, pu, pv, out.ctypes.data_as(ctypes.c_void_p),
                                            _abi.HOST))
        return out

    def time_sampled_outer(self, reps=50):
        """(average ms, algorithmic bytes) of one sampled outer-product launch."""
        ms, by = ctypes.c_double(), ctypes.c_double()
        _check(self._lib.ocsl_time_sampled_outer(self._h, reps, ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value


def build_structure(n_cells, face_left, face_right, bc_left, slots=False):
    """The CSR structure of CFVM::define_matrix_structure (cfvm.cpp:74-98),
    built on the device from the mesh's face lists: per row, ascending and
    each once, the cell, the cells across its faces and, when it owns a
    boundary face, the sentinel column 2^64 - 1.  Returns (row_ptr, col) as
    uint64 arrays, and with slots=True also (face_slot[F, 2], diag_slot[n]):
    the slots of each face's two off-diagonal entries (row face_left, then
    row face_right) and of each diagonal."""
    lib = hip_lib()
    fl = np.ascontiguousarray(face_left, dtype=np.int32)
    fr = np.ascontiguousarray(face_right, dtype=np.int32)
    bl = np.ascontiguousarray(bc_left, dtype=np.int32)
    if fl.shape != fr.shape:
        raise ValueError("face_left and face_right differ in length")
    nnz = ctypes.c_int64()
    vp = ctypes.c_void_p
    args = (int(n_cells), fl.shape[0], fl.ctypes.data_as(vp), fr.ctypes.data_as(vp), bl.shape[0],
            bl.ctypes.data_as(vp), _abi.HOST)
    row_ptr = np.empty(int(n_cells) + 1, np.uint64)
    _check(lib.ocsl_build_structure(*args, row_ptr.ctypes.data_as(vp), None, 0, ctypes.byref(nnz), None, None))
    col = np.empty(nnz.value, np.uint64)
    fs = np.empty((fl.shape[0], 2), np.int64) if slots else None
    ds = np.empty(int(n_cells), np.int64) if slots else None
    _check(lib.ocsl_build_structure(*args, row_ptr.ctypes.data_as(vp), col.ctypes.data_as(vp), col.shape[0],
                                    ctypes.byref(nnz), fs.ctypes.data_as(vp) if slots else None,
                                    ds.ctypes.data_as(vp) if slots else None))
    return (row_ptr, col, fs, ds) if slots else (row_ptr, col)


class Batch:
    """Many independent CSR systems solved by CG in the same launches."""

    def __init__(self, systems):
        """systems: list of (row_ptr, column_index) in the reference layout."""
        self._lib = hip_lib()
        self.m = len(systems)
        self._rp = [np.ascontiguousarray(rp, dtype=np.uint64) for rp, _ in systems]
        self._col = [np.ascontiguousarray(c, dtype=np.uint64) for _, c in systems]
        self.n = [int(r.shape[0] - 1) for r in self._rp]
        n_arr = np.array(self.n, np.int64)
        rps = (_u64p * self.m)(*[r.ctypes.data_as(_u64p) for r in self._rp])
        cols = (_u64p * self.m)(*[c.ctypes.data_as(_u64p) for c in self._col])
        h = _abi.c_handle()
        _check(self._lib.ocsl_batch_create(ctypes.byref(h), self.m,
                                           n_arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                           rps, cols))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ocsl_batch_destroy(self._h)
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

    @staticmethod
    def _ptrs(arrs):
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]
        ptrs = (ctypes.c_void_p * len(keep))(*[a.ctypes.data_as(ctypes.c_void_p) for a in keep])
        return ptrs, keep

    def set_values(self, values):
        ptrs, keep = self._ptrs(values)
        _check(self._lib.ocsl_batch_set_values(self._h, ptrs, _abi.HOST))

    def set_rhs(self, rhs):
        ptrs, keep = self._ptrs(rhs)
        _check(self._lib.ocsl_batch_set_rhs(self._h, ptrs, _abi.HOST))

    def solve(self, precision="fp64", x0=None, **opts):
        """Returns (solutions, reports, latency_ms).  x0: None (zeros), host
        arrays (copied), or torch CUDA float64 tensors, which hold the
        initial guesses on entry and are overwritten by the solutions (the
        returned solutions are those tensors)."""
        if x0 is not None and all(_is_torch_cuda(x) for x in x0):
            xs = list(x0)
            for x, n in zip(xs, self.n):
                _ptr(x)  # checks dtype and layout, orders torch's stream before the call
                if x.numel() != n:
                    raise ValueError("x0 sizes must match the systems")
            where = _abi.DEVICE
            ptrs = (ctypes.c_void_p * self.m)(*[x.data_ptr() for x in xs])
        else:
            xs = [np.zeros(n) if x0 is None else np.array(x0[k], dtype=np.float64) for k, n in enumerate(self.n)]
            where = _abi.HOST
            ptrs = (ctypes.c_void_p * self.m)(*[x.ctypes.data_as(ctypes.c_void_p) for x in xs])
        o = default_options(**opts)
        reps = (_abi.Report * self.m)()
        lat = np.zeros(self.m)
        st = self._lib.ocsl_batch_solve(self._h, PRECISIONS[precision], ctypes.byref(o), ptrs, where,
                                        reps, lat.ctypes.data_as(_f64p))
        _check(st, allow_warn=True)
        return xs, [r.as_dict() for r in reps], lat

    # -- launch shape ------------------------------------------------------
    def set_launch(self, rows_per_lane=1, block_size=1024):
        """Rows per lane (1, 2, 4) of the concatenated layout and threads per
        SpMV workgroup (256, 512, 1024); values and rhs follow."""
        _check(self._lib.ocsl_batch_set_launch(self._h, rows_per_lane, block_size))

    def launch(self):
        r, b = ctypes.c_int(), ctypes.c_int()
        _check(self._lib.ocsl_batch_get_launch(self._h, ctypes.byref(r), ctypes.byref(b)))
        return {"rows_per_lane": r.value, "block_size": b.value}

    def time_iterations(self, precision="mixed", iterations=32):
        """Device us per CG iteration of the whole batch in the current shape."""
        us = ctypes.c_double()
        _check(self._lib.ocsl_batch_time_iterations(self._h, PRECISIONS[precision], iterations, ctypes.byref(us)))
        return us.value

    def tune(self, precision="mixed", iterations=32):
        """Times every rows-per-lane x workgroup-size shape, keeps the fastest.
        Returns (best shape dict, best us/iteration, table)."""
        r, b, us, ln = ctypes.c_int(), ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        cap = 32
        table = (_abi.TuneEntry * cap)()
        _check(self._lib.ocsl_batch_tune(self._h, PRECISIONS[precision], iterations, ctypes.byref(r),
                                         ctypes.byref(b), ctypes.byref(us), table, cap, ctypes.byref(ln)))
        keys = ("rows_per_lane", "block_size", "grid_blocks", "vec_blocks")
        rows = [dict({k: getattr(table[i], k) for k in keys}, us_per_iteration=table[i].us_per_iteration)
                for i in range(ln.value)]
        return {"rows_per_lane": r.value, "block_size": b.value}, us.value, rows
