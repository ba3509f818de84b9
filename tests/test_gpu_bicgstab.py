"""ILU(0)-BiCGSTAB on the GPU through the C ABI, against the reference's golden
vectors and the oracle."""
import glob
import os

import numpy as np
import pytest

import oracle
from opencslsolver_amd import gen
from opencslsolver_amd.gen import CSR
from opencslsolver_amd.solver import Solver

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
SOLVER_FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "solver_*.npz")))


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _solver(a, b, x0=None):
    s = Solver(a.row_ptr, a.col)
    s.set_values(a.val)
    s.set_rhs(b)
    s.set_x(np.zeros(a.n) if x0 is None else x0)
    return s


@pytest.fixture(autouse=True)
def _one_thread():
    oracle.set_threads(1)


@pytest.mark.parametrize("path", SOLVER_FIXTURES, ids=lambda p: os.path.basename(p))
def test_reference_mode_matches_reference(path):
    d = np.load(path)
    a = CSR(d["row_ptr"], d["col"], d["val"])
    with _solver(a, d["b"], d["x0"]) as s:
        m = s.ilu_decompose(3)
        res = s.ILU_BICGSTAB_solve()
        x = s.get_x()
        hist = s.history()
    assert _rel(m, d["ilu_m"]) <= 1e-12
    assert res == pytest.approx(float(d["bicg_res"]), rel=1e-8, abs=1e-28)
    assert _rel(x, d["bicg_x"]) <= 1e-9
    _, _, h_or = oracle.ilu_bicgstab_reference(a, d["b"], d["x0"])
    assert len(hist) == len(h_or)
    assert np.allclose(hist, h_or, rtol=1e-8, atol=1e-28)


@pytest.mark.parametrize("tag", ["bicgstab", "split"])
def test_reference_mode_on_sod_systems(tag):
    """Shock-tube momentum / enthalpy systems: a tridiagonal pattern, so the
    triangular solves have 1000 dependent levels."""
    d = np.load(os.path.join(GOLDEN, f"sod_{tag}_2steps.npz"))
    rp, col = d["row_ptr"], d["col"]
    checked = 0
    for k in range(d["dump_val"].shape[0]):
        if d["dump_method"][k] != 1:
            continue
        a = CSR(rp, col, d["dump_val"][k])
        with _solver(a, d["dump_b"][k], d["dump_x0"][k]) as s:
            res = s.ILU_BICGSTAB_solve()
            x = s.get_x()
        assert res == pytest.approx(d["res_hist"][k], rel=1e-8, abs=1e-28), k
        assert _rel(x, d["dump_x"][k]) <= 1e-9, k
        checked += 1
    assert checked >= 4


@pytest.mark.parametrize("sweeps", [1, 2, 5])
def test_ilu_sweeps_match_oracle(sweeps):
    a, _ = gen.convdiff3d(8, sentinels=True)
    with Solver(a.row_ptr, a.col) as s:
        s.set_values(a.val)
        m = s.ilu_decompose(sweeps)
    assert _rel(m, oracle.ilu_decompose(a, sweeps)) <= 1e-13


@pytest.mark.parametrize("n,tri", [(16, "exact"), (16, "jacobi"), (48, "exact")])
def test_tolerance_fp64_matches_oracle(n, tri):
    a, _ = gen.convdiff3d(n, seed_perm=3)
    b = gen.uniform(4, a.n)
    tri_sweeps = 3 if tri == "jacobi" else 0
    it_o, x_o, h_o, tr_o = oracle.bicgstab_tol(a, b, tol=1e-10, ilu_sweeps=3, tri_sweeps=tri_sweeps,
                                               tri_w=0.9)
    with _solver(a, b) as s:
        rep = s.solve("ilu_bicgstab", "tolerance", "fp64", tol=1e-10,
                      tri_mode=1 if tri == "jacobi" else 0, tri_sweeps=3, tri_weight=0.9)
        x = s.get_x()
        h = s.history()
    assert rep["status"] == 0, rep
    assert abs(rep["iterations"] - it_o) <= 1
    assert rep["true_relres"] <= 1.5e-10
    assert np.linalg.norm(b - a.matvec(x)) / np.linalg.norm(b) <= 1.5e-10
    k = min(len(h), len(h_o))
    assert np.allclose(h[:k], h_o[:k], rtol=1e-6)
    assert _rel(x, x_o) <= 1e-8


def test_mixed_reaches_fp64_tolerance():
    a, _ = gen.convdiff3d(24, seed_perm=3)
    b = gen.uniform(4, a.n)
    with _solver(a, b) as s:
        s.solve("ilu_bicgstab", "tolerance", "fp64", tol=1e-10)
        x64 = s.get_x()
        s.set_x(np.zeros(a.n))
        rep = s.solve("ilu_bicgstab", "tolerance", "mixed", tol=1e-10)
        xm = s.get_x()
    assert rep["status"] == 0, rep
    assert np.linalg.norm(b - a.matvec(xm)) / np.linalg.norm(b) <= 1e-10
    assert np.linalg.norm(xm - x64) / np.linalg.norm(x64) <= 1e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_three_rhs_equal_single_solves(precision):
    a, _ = gen.convdiff3d(20, seed_perm=3)
    B = np.stack([gen.uniform(s, a.n) for s in (4, 5, 6)])
    with _solver(a, B[0]) as s:
        X, reps = s.solve_multi(B, precision=precision, tol=1e-10)
        singles = []
        for k in range(3):
            s.set_rhs(B[k])
            s.set_x(np.zeros(a.n))
            s.solve("ilu_bicgstab", "tolerance", precision, tol=1e-10)
            singles.append(s.get_x())
    for k in range(3):
        assert reps[k]["status"] == 0, reps[k]
        rel = np.linalg.norm(B[k] - a.matvec(X[k])) / np.linalg.norm(B[k])
        assert rel <= 1e-10
        assert _rel(X[k], singles[k]) <= (1e-12 if precision == "fp64" else 1e-9)
