"""Seeded synthetic systems (host only, no GPU).

Thin numpy wrappers over ``libocsl_gen.so`` (include/ocsl_gen.h), which fixes
the RNG contract: std::mt19937_64 engine, U = (e >> 11) * 2^-53, Box-Muller
normals, Fisher-Yates permutation from the top.  Matrices come back in the
reference's CSR layout (uint64 row_ptr / column_index, float64 values), the
layout of ``CMatrix_solver::CCSR`` (cmatrix_solver.h:33-51).
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from ._native import gen_lib

SENTINEL = np.uint64(0xFFFFFFFFFFFFFFFF)

_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


@dataclasses.dataclass
class CSR:
    """Square CSR matrix in the reference layout."""

    row_ptr: np.ndarray  # uint64[n+1]
    col: np.ndarray  # uint64[nnz], sentinel = 2**64-1
    val: np.ndarray  # float64[nnz]

    @property
    def n(self) -> int:
        return int(self.row_ptr.shape[0] - 1)

    @property
    def nnz(self) -> int:
        return int(self.col.shape[0])

    def to_dense(self) -> np.ndarray:
        n = self.n
        a = np.zeros((n, n))
        for i in range(n):
            for p in range(int(self.row_ptr[i]), int(self.row_ptr[i + 1])):
                c = int(self.col[p])
                if c < n:
                    a[i, c] += self.val[p]
        return a

    def matvec(self, x: np.ndarray) -> np.ndarray:
        n = self.n
        rows = np.repeat(np.arange(n), np.diff(self.row_ptr.astype(np.int64)))
        mask = self.col < np.uint64(n)
        out = np.zeros(n)
        np.add.at(out, rows[mask], self.val[mask] * x[self.col[mask].astype(np.int64)])
        return out

    def with_values(self, val: np.ndarray) -> "CSR":
        return CSR(self.row_ptr, self.col, np.ascontiguousarray(val, dtype=np.float64))


def hex_sizes(n: int, sentinels: bool = False) -> tuple[int, int]:
    rows, nnz = ctypes.c_int64(), ctypes.c_int64()
    rc = gen_lib().ocsl_gen_hex_sizes(n, int(sentinels), ctypes.byref(rows), ctypes.byref(nnz))
    if rc != 0:
        raise ValueError(f"ocsl_gen_hex_sizes({n}) failed: {rc}")
    return rows.value, nnz.value


def poisson3d(n: int, seed_mesh: int = 1, seed_rhs: int | None = 2, sigma: float = 0.5,
              leak: float = 1e-6, permute: bool = True, sentinels: bool = False):
    """SPD pressure-correction system on an n^3 hex mesh; returns (A, b, perm).

    ``seed_rhs=None`` draws b from the continuation of the mesh stream.
    """
    rows, nnz = hex_sizes(n, sentinels)
    rp = np.empty(rows + 1, np.uint64)
    col = np.empty(nnz, np.uint64)
    val = np.empty(nnz, np.float64)
    b = np.empty(rows, np.float64)
    perm = np.empty(rows, np.int64)
    sr = 0xFFFFFFFFFFFFFFFF if seed_rhs is None else seed_rhs
    rc = gen_lib().ocsl_gen_poisson3d(n, ctypes.c_uint64(seed_mesh), ctypes.c_uint64(sr),
                                      ctypes.c_double(sigma), ctypes.c_double(leak), int(permute),
                                      int(sentinels), _p(rp, _u64p), _p(col, _u64p),
                                      _p(val, _f64p), _p(b, _f64p), _p(perm, _i64p))
    if rc != 0:
        raise ValueError(f"ocsl_gen_poisson3d failed: {rc}")
    return CSR(rp, col, val), b, perm


def chain1d(n: int, seed: int) -> CSR:
    """Nonsymmetric, diagonally dominant tridiagonal matrix (a 1D
    convection-diffusion row, the shape of the shock-tube systems): every
    level of its ILU(0) triangular solves holds one row.  Off-diagonals
    -1 - U[0,1) (lower) and -0.5 - U[0,1) (upper), diagonal 4 + U[0,1), row by
    row from seed."""
    rng = np.random.default_rng(seed)
    rp, col, val = [0], [], []
    for i in range(n):
        if i > 0:
            col.append(i - 1)
            val.append(-1.0 - rng.random())
        col.append(i)
        val.append(4.0 + rng.random())
        if i + 1 < n:
            col.append(i + 1)
            val.append(-0.5 - rng.random())
        rp.append(len(col))
    return CSR(np.array(rp, np.uint64), np.array(col, np.uint64), np.array(val))


def convdiff3d(n: int, seed_perm: int = 3, rho: float = 1.2, dt: float = 1e-3,
               peclet: float = 10.0, permute: bool = True, sentinels: bool = False):
    """Taylor-Green upwind convection-diffusion system; returns (A, perm)."""
    rows, nnz = hex_sizes(n, sentinels)
    rp = np.empty(rows + 1, np.uint64)
    col = np.empty(nnz, np.uint64)
    val = np.empty(nnz, np.float64)
    perm = np.empty(rows, np.int64)
    rc = gen_lib().ocsl_gen_convdiff3d(n, ctypes.c_uint64(seed_perm), ctypes.c_double(rho),
                                       ctypes.c_double(dt), ctypes.c_double(peclet), int(permute),
                                       int(sentinels), _p(rp, _u64p), _p(col, _u64p),
                                       _p(val, _f64p), _p(perm, _i64p))
    if rc != 0:
        raise ValueError(f"ocsl_gen_convdiff3d failed: {rc}")
    return CSR(rp, col, val), perm


def uniform(seed: int, n: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    out = np.empty(n, np.float64)
    rc = gen_lib().ocsl_gen_uniform(ctypes.c_uint64(seed), ctypes.c_int64(n), ctypes.c_double(lo),
                                    ctypes.c_double(hi), _p(out, _f64p))
    if rc != 0:
        raise ValueError("ocsl_gen_uniform failed")
    return out


def normal(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    rc = gen_lib().ocsl_gen_normal(ctypes.c_uint64(seed), ctypes.c_int64(n), _p(out, _f64p))
    if rc != 0:
        raise ValueError("ocsl_gen_normal failed")
    return out


def batch_plan(seed: int = 7, m: int = 64, lo: int = 22, hi: int = 80):
    """Sizes n_k and per-system (mesh, rhs) seeds of m independent systems."""
    n = np.empty(m, np.int32)
    sm = np.empty(m, np.uint64)
    sr = np.empty(m, np.uint64)
    rc = gen_lib().ocsl_gen_batch_plan(ctypes.c_uint64(seed), m, lo, hi,
                                       n.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       _p(sm, _u64p), _p(sr, _u64p))
    if rc != 0:
        raise ValueError("ocsl_gen_batch_plan failed")
    return n, sm, sr


def sample_slots(seed: int, a: CSR, count: int) -> np.ndarray:
    out = np.empty(count, np.int64)
    rc = gen_lib().ocsl_gen_sample_slots(ctypes.c_uint64(seed), ctypes.c_int64(a.n),
                                         _p(a.row_ptr, _u64p), _p(a.col, _u64p), count,
                                         _p(out, _i64p))
    if rc != 0:
        raise ValueError(f"ocsl_gen_sample_slots failed: {rc}")
    return out


def system_hash(a: CSR, b: np.ndarray | None = None) -> int:
    return int(gen_lib().ocsl_gen_hash(ctypes.c_int64(a.n), _p(a.row_ptr, _u64p),
                                       _p(a.col, _u64p), _p(a.val, _f64p), _p(b, _f64p)))


def tridiag_unit_test_system():
    """The 10x10 system of the reference's CG unit test
    (cmatrix_solver_supply.h:29-51): diagonal 5, off-diagonals 1, b = 2, x0 = 1."""
    n = 10
    rp = [0]
    cols, vals = [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                cols.append(j)
                vals.append(5.0 if j == i else 1.0)
        rp.append(len(cols))
    a = CSR(np.array(rp, np.uint64), np.array(cols, np.uint64), np.array(vals, np.float64))
    return a, np.full(n, 2.0), np.full(n, 1.0)



def hex_mesh(n: int, h: float | None = None, perm: np.ndarray | None = None):
    """Uniform n^3 hexahedral finite-volume mesh of the unit cube (cell size
    h = 1/n unless given) in the CFVM form (cfvm.cpp:100-170): interior faces
    x-, then y-, then z-oriented in natural order, boundary faces of the six
    sides (markers 0..5: x-, x+, y-, y+, z-, z+).  perm[c] renumbers natural
    cell c = i + n (j + n k).  Returns (mesh dict, row_ptr, col) with the
    CSR structure of CFVM::define_matrix_structure (cfvm.cpp:74-98): per row
    the sorted cell and neighbour columns, and one sentinel column
    (2^64 - 1) when the cell has boundary faces."""
    h = 1.0 / n if h is None else float(h)
    N = n ** 3
    nat = np.arange(N, dtype=np.int64).reshape(n, n, n)  # [k, j, i]
    ids = nat if perm is None else np.asarray(perm, dtype=np.int64)[nat]
    A = h * h
    faces_L, faces_R, normals = [], [], []
    for axis in range(3):  # 0: x (i), 1: y (j), 2: z (k)
        ax = 2 - axis
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, n - 1)
        hi[ax] = slice(1, n)
        faces_L.append(ids[tuple(lo)].reshape(-1))
        faces_R.append(ids[tuple(hi)].reshape(-1))
        e = np.zeros(3)
        e[axis] = 1.0
        normals.append(np.tile(e, (faces_L[-1].shape[0], 1)))
    fL = np.concatenate(faces_L).astype(np.int32)
    fR = np.concatenate(faces_R).astype(np.int32)
    fe = np.concatenate(normals)
    F = fL.shape[0]
    bL, bm, bn = [], [], []
    for axis in range(3):
        ax = 2 - axis
        for side, idx in ((0, 0), (1, n - 1)):
            sl = [slice(None)] * 3
            sl[ax] = idx
            cells = ids[tuple(sl)].reshape(-1)
            e = np.zeros(3)
            e[axis] = -1.0 if side == 0 else 1.0
            bL.append(cells)
            bm.append(np.full(cells.shape[0], 2 * axis + side, np.int32))
            bn.append(np.tile(e, (cells.shape[0], 1)))
    bL = np.concatenate(bL).astype(np.int32)
    bm = np.concatenate(bm)
    be = np.concatenate(bn)
    B = bL.shape[0]
    mesh = {
        "cell_volume": np.full(N, h ** 3),
        "face_left": fL, "face_right": fR,
        "face_area_vector": A * fe, "face_unit": fe.copy(), "face_alpha": fe / h,
        "face_area": np.full(F, A), "face_dist": np.full(F, h), "face_ksi": np.full(F, h),
        "face_geo_aver": np.full(F, 0.5),
        "bc_left": bL, "bc_marker": bm,
        "bc_area_vector": A * be, "bc_unit": be.copy(), "bc_alpha": 2.0 * be / h,
        "bc_area": np.full(B, A), "bc_dist": np.full(B, 0.5 * h), "bc_ksi": np.full(B, 2.0 * h),
        "bc_perp": np.full(B, 0.5 * h),
    }
    # matrix structure: cell, neighbours, one sentinel for boundary cells
    sentinel = np.uint64(0xFFFFFFFFFFFFFFFF)
    rows = np.concatenate([np.arange(N), fL.astype(np.int64), fR.astype(np.int64)])
    cols = np.concatenate([np.arange(N), fR.astype(np.int64), fL.astype(np.int64)]).astype(np.uint64)
    bcells = np.unique(bL.astype(np.int64))
    rows = np.concatenate([rows, bcells])
    cols = np.concatenate([cols, np.full(bcells.shape[0], sentinel, np.uint64)])
    order = np.lexsort((cols, rows))
    rows, cols = rows[order], cols[order]
    keep = np.ones(rows.shape[0], bool)
    keep[1:] = (rows[1:] != rows[:-1]) | (cols[1:] != cols[:-1])
    rows, cols = rows[keep], cols[keep]
    row_ptr = np.zeros(N + 1, np.uint64)
    np.add.at(row_ptr, rows + 1, 1)
    row_ptr = np.cumsum(row_ptr).astype(np.uint64)
    return mesh, row_ptr, cols
