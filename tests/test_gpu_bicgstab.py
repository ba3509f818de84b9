"""ILU(0)-BiCGSTAB on the GPU through the C ABI, against the reference's golden
vectors and the oracle."""
import glob
import os

import numpy as np
import pytest

import oracle
import parity
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
    """REFERENCE ILU(0)-BiCGSTAB against the reference's own results: factors,
    returned residual, solution, and the per-iteration residuals against the
    oracle's restatement."""
    name = os.path.basename(path)[:-4]
    d = np.load(path)
    a = CSR(d["row_ptr"], d["col"], d["val"])
    with _solver(a, d["b"], d["x0"]) as s:
        m = s.ilu_decompose(3)
        res = s.ILU_BICGSTAB_solve()
        x = s.get_x()
        hist = s.history()
    parity.check(f"{name} ilu_m", parity.rel(m, d["ilu_m"]), 1e-12)
    parity.check(f"{name} bicg residual", parity.res_rel(res, d["bicg_res"]), parity.bar("solver", name, "bicg_res"))
    parity.check(f"{name} bicg x", parity.rel(x, d["bicg_x"]), parity.bar("solver", name, "bicg_x"))
    _, _, h_or = oracle.ilu_bicgstab_reference(a, d["b"], d["x0"])
    parity.check(f"{name} bicg history vs oracle", parity.hist_rel(hist, h_or, floor=parity.RES_FLOOR),
                 parity.FLOOR_BAR)


@pytest.mark.parametrize("tag", ["bicgstab", "split"])
def test_reference_mode_on_sod_systems(tag):
    """Shock-tube momentum / enthalpy systems: a tridiagonal pattern, so the
    triangular solves have 1000 dependent levels."""
    d = np.load(os.path.join(GOLDEN, f"sod_{tag}_2steps.npz"))
    rp, col = d["row_ptr"], d["col"]
    checked = 0
    worst_res = worst_x = 0.0
    for k in range(d["dump_val"].shape[0]):
        if d["dump_method"][k] != 1:
            continue
        a = CSR(rp, col, d["dump_val"][k])
        with _solver(a, d["dump_b"][k], d["dump_x0"][k]) as s:
            res = s.ILU_BICGSTAB_solve()
            x = s.get_x()
        worst_res = max(worst_res, parity.res_rel(res, d["res_hist"][k]))
        worst_x = max(worst_x, parity.rel(x, d["dump_x"][k]))
        checked += 1
    assert checked >= 4
    parity.check(f"sod {tag} systems residual", worst_res, parity.bar("sod_systems", tag, "res"))
    parity.check(f"sod {tag} systems x", worst_x, parity.bar("sod_systems", tag, "x"))


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


def test_convdiff_96cubed_three_rhs_against_oracle():
    """The 96^3 Taylor-Green convection-diffusion system (cells permuted by
    seed 3) with three right-hand sides (seeds 4-6) sharing A: ILU(0)-
    BiCGSTAB to 1e-10 in one multi-RHS solve, FP64 and MIXED (FP32 inner
    iterations, FP64 refinement).  Every solution's relative residual is
    recomputed by a host SpMV; each is compared with the oracle's FP64
    solution of the same right-hand side, and each MIXED solution with the
    FP64 one (errors printed per right-hand side).  The same in the colour
    order (OCSL_BICGSTAB_ORDER_COLOUR): residuals within 1e-10, solutions
    within 1e-9 of the oracle's."""
    oracle.set_threads(8)
    a, _ = gen.convdiff3d(96, seed_perm=3)
    assert (a.n, a.nnz) == (884736, 6137856)
    B = np.stack([gen.uniform(sd, a.n) for sd in (4, 5, 6)])
    ref = [oracle.bicgstab_tol(a, B[k], tol=1e-10) for k in range(3)]
    X = {}
    with Solver(a.row_ptr, a.col) as s:
        s.set_values(a.val)
        for order, tag in ((0, ""), (1, " colour")):
            for prec in ("fp64", "mixed"):
                X[prec + tag], reps = s.solve_multi(B, precision=prec, tol=1e-10, bicgstab_order=order)
                for k in range(3):
                    assert reps[k]["status"] == 0, (prec, tag, k, reps[k])
                    relres = np.linalg.norm(B[k] - a.matvec(X[prec + tag][k])) / np.linalg.norm(B[k])
                    parity.check(f"convdiff96{tag} {prec} rhs{k} host relres", relres, 1e-10)
    for k in range(3):
        # the colour order factors the renumbered matrix: another
        # preconditioner, the same residual target
        for prec in ("fp64", "mixed"):
            parity.check(f"convdiff96 colour {prec} rhs{k} vs oracle fp64",
                         parity.rel(X[prec + " colour"][k], ref[k][1]), 1e-9)
        it_o, x_o, _, _ = ref[k]
        parity.check(f"convdiff96 fp64 rhs{k} vs oracle fp64", parity.rel(X["fp64"][k], x_o), 1e-10)
        err = np.linalg.norm(X["mixed"][k] - X["fp64"][k]) / np.linalg.norm(X["fp64"][k])
        # both solutions have relative residual <= 1e-10; the system is
        # diagonally dominant (rho V / dt), so their distance is of that order
        parity.check(f"convdiff96 mixed rhs{k} error vs fp64", float(err), 1e-9)
        parity.check(f"convdiff96 mixed rhs{k} vs oracle fp64", parity.rel(X["mixed"][k], x_o), 1e-9)


def _renumbered(a, order):
    """P A P^T for new -> old `order`: rows in the new order, each row's
    entries by ascending new column (host numpy)."""
    n = a.n
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)
    rp = a.row_ptr.astype(np.int64)
    rows = np.repeat(np.arange(n), np.diff(rp))
    nr, nc = inv[rows], inv[a.col.astype(np.int64)]
    key = np.lexsort((nc, nr))
    counts = np.bincount(nr, minlength=n)
    out_rp = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    return CSR(out_rp, nc[key].astype(np.uint64), a.val[key].copy())


def _lower_depth(a):
    """Levels of the lower triangle's dependency graph (host)."""
    n = a.n
    rows = np.repeat(np.arange(n), np.diff(a.row_ptr.astype(np.int64)))
    cols = a.col.astype(np.int64)
    m = cols < rows
    r, c = rows[m], cols[m]
    lev = np.zeros(n, np.int64)
    while True:
        new = lev.copy()
        np.maximum.at(new, r, lev[c] + 1)
        if np.array_equal(new, lev):
            return int(lev.max()) + 1
        lev = new


@pytest.mark.parametrize("n", [10, 24])
def test_colour_order_solves_the_renumbered_system(n):
    """OCSL_BICGSTAB_ORDER_COLOUR: the order the handle reports is a
    permutation whose renumbered matrix has a lower triangle of exactly
    n_colours levels (no two rows of a colour coupled; 2 colours on the
    7-point pattern, which is bipartite).  The FP64 solves of three
    right-hand sides equal the oracle's ILU(0)-BiCGSTAB run on that
    renumbered matrix (1e-10, same iteration count); MIXED lands within 1e-9
    of FP64; every residual, recomputed on the host in the caller's
    numbering, is within the 1e-10 target; solutions of the input-order solve
    agree to 1e-9."""
    a, _ = gen.convdiff3d(n, seed_perm=3)
    B = np.stack([gen.uniform(sd, a.n) for sd in (4, 5, 6)])
    with Solver(a.row_ptr, a.col) as s:
        s.set_values(a.val)
        order, nc = s.colour_order()
        assert np.array_equal(np.sort(order), np.arange(a.n))
        ap = _renumbered(a, order)
        assert nc == 2 and _lower_depth(ap) == nc
        X = {}
        for prec in ("fp64", "mixed"):
            X[prec], reps = s.solve_multi(B, precision=prec, tol=1e-10, bicgstab_order=1)
            for k in range(3):
                assert reps[k]["status"] == 0, (prec, k, reps[k])
                relres = np.linalg.norm(B[k] - a.matvec(X[prec][k])) / np.linalg.norm(B[k])
                parity.check(f"colour{n} {prec} rhs{k} host relres", relres, 1e-10)
            if prec == "fp64":
                its = [r["iterations"] for r in reps]
        X_in, _ = s.solve_multi(B, precision="fp64", tol=1e-10)
    for k in range(3):
        it_o, x_o, _, _ = oracle.bicgstab_tol(ap, B[k][order], tol=1e-10)
        x_back = np.empty(a.n)
        x_back[order] = x_o
        assert its[k] == it_o, (k, its[k], it_o)
        parity.check(f"colour{n} fp64 rhs{k} vs oracle on P A P^T", parity.rel(X["fp64"][k], x_back), 1e-10)
        parity.check(f"colour{n} mixed rhs{k} vs fp64", parity.rel(X["mixed"][k], X["fp64"][k]), 1e-9)
        parity.check(f"colour{n} fp64 rhs{k} vs input order", parity.rel(X["fp64"][k], X_in[k]), 1e-9)


def test_colour_order_single_rhs_and_history():
    """ocsl_solve in the colour order: the same solution as the one-column
    multi-RHS solve, a history of iterations + 1 entries starting at the
    initial relative residual 1, and a later input-order solve on the same
    handle unaffected (equal to a fresh handle's)."""
    a, _ = gen.convdiff3d(12, seed_perm=5, dt=1.0)
    b = gen.uniform(7, a.n)
    with _solver(a, b) as s:
        rep = s.solve("ilu_bicgstab", "tolerance", "fp64", tol=1e-10, bicgstab_order=1)
        x1 = s.get_x()
        h = s.history()
        X, _ = s.solve_multi(b[None, :], precision="fp64", tol=1e-10, bicgstab_order=1)
        s.set_x(np.zeros(a.n))
        s.solve("ilu_bicgstab", "tolerance", "fp64", tol=1e-10)
        x_in = s.get_x()
    assert rep["status"] == 0 and rep["iterations"] >= 1
    assert np.array_equal(x1, X[0])
    assert h.size == rep["iterations"] + 1
    assert abs(h[0] - 1.0) <= 1e-12 and np.all(np.isfinite(h))  # x0 = 0: ||r0|| / ||b|| = 1
    with _solver(a, b) as s:
        s.solve("ilu_bicgstab", "tolerance", "fp64", tol=1e-10)
        assert np.array_equal(s.get_x(), x_in)


@pytest.mark.parametrize("n,seed", [(12, 3), (48, 3), (64, 31)])
def test_exact_triangular_solves_match_oracle(n, seed):
    """ocsl_ILU_forsub / ocsl_ILU_backsub with a 3-sweep ILU(0) factor against
    the oracle's sequential substitutions: 12^3 runs every level in one
    workgroup; 48^3 and 64^3 (randomly numbered, 19-24 levels) launch one
    kernel per level.  Same terms in the same order per row: 1e-13."""
    a, _ = gen.convdiff3d(n, seed_perm=seed)
    b = gen.uniform(4, a.n)
    m = oracle.ilu_decompose(a, 3)
    with Solver(a.row_ptr, a.col) as s:
        s.set_values(a.val)
        y = s.forsub(m, b)
        x = s.backsub(m, y)
    y_o = oracle.forsub(a, m, b)
    parity.check(f"forsub {n}^3", parity.rel(y, y_o), 1e-13)
    parity.check(f"backsub {n}^3", parity.rel(x, oracle.backsub(a, m, y_o)), 1e-13)


@pytest.mark.parametrize("kind", ["chain", "box"])
def test_triangular_solve_forms_agree_bitwise(kind):
    """ocsl_ILU_forsub / ocsl_ILU_backsub in every form: chosen from the
    schedule (0), one workgroup (1), one launch per level (2), serial over the
    rows staged in LDS (3).  Each adds the same terms in the same order, so
    the results are bitwise equal.  The chain (1000 rows, one per level) takes
    the serial form by default; the randomly numbered 8^3 box (about 8
    levels of 64 rows) one workgroup."""
    a = gen.chain1d(1000, 5) if kind == "chain" else gen.convdiff3d(8, seed_perm=3)[0]
    b = gen.uniform(6, a.n)
    m = oracle.ilu_decompose(a, 3)
    out = {}
    with Solver(a.row_ptr, a.col) as s:
        s.set_values(a.val)
        for form in (0, 1, 2, 3):
            s.set_tri_form(form)
            y = s.forsub(m, b)
            out[form] = (y, s.backsub(m, y))
    for form in (1, 2, 3):
        assert np.array_equal(out[form][0], out[0][0]), form
        assert np.array_equal(out[form][1], out[0][1]), form
    y_o = oracle.forsub(a, m, b)
    parity.check(f"forsub {kind}", parity.rel(out[0][0], y_o), 1e-13)
    parity.check(f"backsub {kind}", parity.rel(out[0][1], oracle.backsub(a, m, y_o)), 1e-13)


def test_serial_triangular_form_inside_the_solvers():
    """The serial form inside the REFERENCE ILU(0)-BiCGSTAB (FP64, one vector)
    and the TOLERANCE MIXED solve of three right-hand sides (FP32 factor,
    three interleaved vectors) on a 1000-row chain: bitwise the results of
    one launch per level."""
    a = gen.chain1d(1000, 7)
    B = np.stack([gen.uniform(sd, a.n) for sd in (4, 5, 6)])
    res = {}
    for form in (0, 2):
        with Solver(a.row_ptr, a.col) as s:
            s.set_values(a.val)
            s.set_tri_form(form)
            s.set_rhs(B[0])
            s.set_x(np.zeros(a.n))
            r = s.ILU_BICGSTAB_solve()
            x = s.get_x()
            X, reps = s.solve_multi(B, precision="mixed", tol=1e-10)
            res[form] = (r, x, X, [q["iterations"] for q in reps])
    assert res[0][0] == res[2][0]
    assert np.array_equal(res[0][1], res[2][1])
    assert np.array_equal(res[0][2], res[2][2]) and res[0][3] == res[2][3]
    for k in range(3):
        assert np.linalg.norm(B[k] - a.matvec(res[0][2][k])) <= 1e-10 * np.linalg.norm(B[k])


@pytest.mark.parametrize("kind,k", [("chain", 3), ("chain", 1), ("box", 3), ("perm32", 4), ("perm32", 2)])
def test_reference_multi_rhs_equals_single_solves(kind, k):
    """ocsl_ILU_BICGSTAB_solve_multi: each right-hand side's solution,
    residual and history bitwise those of ocsl_ILU_BICGSTAB_solve with that
    right-hand side and guess, on the 1000-row chain (serial triangular
    form), the randomly numbered 8^3 box (one workgroup) and 32^3 system (a
    launch per level).  The chain cases include a zero right-hand side (the
    reference's 0/0 residual) and a guess that is already the solution (the
    reference's early exit); a second call replays the captured graph."""
    if kind == "chain":
        a = gen.chain1d(1000, 5)
    else:
        a = gen.convdiff3d(8 if kind == "box" else 32, seed_perm=3)[0]
    rng = np.random.default_rng(17)
    B = rng.uniform(-1.0, 1.0, (k, a.n))
    X0 = rng.standard_normal((k, a.n)) * 0.01
    if kind == "chain" and k == 3:
        B[1] = 0.0
        X0[1] = 0.0
        dense = np.zeros((a.n, a.n))
        for i in range(a.n):
            for q in range(int(a.row_ptr[i]), int(a.row_ptr[i + 1])):
                if int(a.col[q]) < a.n:
                    dense[i, int(a.col[q])] += a.val[q]
        X0[2] = np.linalg.solve(dense, B[2])
    singles = []
    with Solver(a.row_ptr, a.col) as s:
        s.set_values(a.val)
        for c in range(k):
            s.set_rhs(B[c])
            s.set_x(X0[c])
            r = s.ILU_BICGSTAB_solve()
            singles.append((s.get_x(), r, s.history()))
        X, res = s.ILU_BICGSTAB_solve_multi(B, X0)
        hist = s.history()
        X2, res2 = s.ILU_BICGSTAB_solve_multi(B, X0)
    assert np.array_equal(X, X2, equal_nan=True) and np.array_equal(res, res2, equal_nan=True)
    at = 0
    for c in range(k):
        x1, r1, h1 = singles[c]
        assert np.array_equal(X[c], x1, equal_nan=True), c
        assert np.array_equal(res[c], r1, equal_nan=True), (c, res[c], r1)
        assert np.array_equal(hist[at:at + len(h1)], h1, equal_nan=True), c
        at += len(h1)
    assert at == len(hist)
    if kind == "chain" and k == 3:
        assert len(singles[2][2]) == 1  # the early exit after one residual
        assert np.isnan(res[1])


def test_graph_cache_follows_the_jacobi_weight():
    """TOLERANCE BiCGSTAB with damped-Jacobi triangular sweeps replays its
    iteration chunks as graphs whose sweeps take the weight as an argument: a
    handle that solved with one weight and then solves with another must give
    what a fresh handle gives with the second weight (and the oracle's
    iteration count)."""
    a, _ = gen.convdiff3d(16, seed_perm=3, dt=1.0)
    b = gen.uniform(4, a.n)
    opts = dict(tol=1e-10, tri_mode=1, tri_sweeps=3, check_every=2)
    with _solver(a, b) as s:
        s.solve("ilu_bicgstab", "tolerance", "fp64", tri_weight=0.9, **opts)
        s.set_x(np.zeros(a.n))
        rep = s.solve("ilu_bicgstab", "tolerance", "fp64", tri_weight=0.7, **opts)
        x_reused = s.get_x()
    assert rep["iterations"] > 4, rep  # chunks of check_every iterations replayed as graphs
    with _solver(a, b) as s:
        rep_f = s.solve("ilu_bicgstab", "tolerance", "fp64", tri_weight=0.7, **opts)
        x_fresh = s.get_x()
    assert rep["iterations"] == rep_f["iterations"]
    assert np.array_equal(x_reused, x_fresh)
    it_o, x_o, _, _ = oracle.bicgstab_tol(a, b, tol=1e-10, ilu_sweeps=3, tri_sweeps=3, tri_w=0.7)
    assert abs(rep["iterations"] - it_o) <= 1
    assert parity.rel(x_reused, x_o) <= 1e-7


@pytest.mark.parametrize("method", ["cg", "ilu_bicgstab"])
@pytest.mark.parametrize("n", [12, 48])
def test_reference_graph_replays_match_eager(method, n):
    """REFERENCE solves replay one hipGraph per right-hand-side address after
    their first run.  Against a handle that issues every launch eagerly: the
    same returned residual, solution and history, bit for bit, on every one
    of several solves with new values, right-hand sides and initial guesses,
    with eager work on the replaying handle between the replays.  12^3 runs
    the triangular solves in one workgroup, 48^3 (110,592 rows) one launch
    per level."""
    a, _ = gen.convdiff3d(n, seed_perm=3)
    with Solver(a.row_ptr, a.col) as g, Solver(a.row_ptr, a.col) as e:
        e.set_reference_graphs(False)
        for k in range(5):
            vals = a.val * (1.0 + 0.01 * k)
            b = gen.uniform(40 + k, a.n)
            x0 = gen.uniform(60 + k, a.n)
            out = []
            for s in (g, e):
                s.set_values(vals)
                s.set_rhs(b)
                s.set_x(x0)
                res = s.CG_solve() if method == "cg" else s.ILU_BICGSTAB_solve()
                out.append((res, s.get_x(), s.history()))
            assert out[0][0] == out[1][0], k
            assert np.array_equal(out[0][1], out[1][1]), k
            assert np.array_equal(out[0][2], out[1][2]), k
            # eager work of other kinds on the replaying handle between replays
            g.set_x(np.zeros(a.n))
            g.solve(method, "tolerance", "fp64", tol=1e-6)
            g.gemv(1.0, x0)
