"""CG on the GPU through the C ABI, against the reference's golden vectors and
the oracle."""
import glob
import os

import numpy as np
import pytest

import oracle
import parity
from opencslsolver_amd import gen
from opencslsolver_amd.gen import CSR
from opencslsolver_amd.solver import OcslError, Solver

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


@pytest.mark.parametrize("path", SOLVER_FIXTURES, ids=lambda p: os.path.basename(p))
def test_cg_reference_mode_matches_reference(path):
    name = os.path.basename(path)[:-4]
    d = np.load(path)
    a = CSR(d["row_ptr"], d["col"], d["val"])
    with _solver(a, d["b"], d["x0"]) as s:
        res = s.CG_solve()
        x = s.get_x()
        hist = s.history()
    parity.check(f"{name} cg residual", parity.res_rel(res, d["cg_res"], floor=0.0), parity.bar("solver", name, "cg_res"))
    parity.check(f"{name} cg x", parity.rel(x, d["cg_x"]), parity.bar("solver", name, "cg_x"))
    _, _, h_or = oracle.cg_reference(a, d["b"], d["x0"])
    parity.check(f"{name} cg history vs oracle", parity.hist_rel(hist, h_or, floor=0.0), parity.FLOOR_BAR)


def test_cg_reference_mode_on_sod_systems():
    d = np.load(os.path.join(GOLDEN, "sod_split_2steps.npz"))
    rp, col = d["row_ptr"], d["col"]
    checked = 0
    worst_res = worst_x = 0.0
    for k in range(d["dump_val"].shape[0]):
        if d["dump_method"][k] != 0:
            continue
        a = CSR(rp, col, d["dump_val"][k])
        with _solver(a, d["dump_b"][k], d["dump_x0"][k]) as s:
            res = s.CG_solve()
            x = s.get_x()
        worst_res = max(worst_res, parity.res_rel(res, d["res_hist"][k]))
        worst_x = max(worst_x, parity.rel(x, d["dump_x"][k]))
        checked += 1
    assert checked >= 2
    parity.check("sod split cg systems residual", worst_res, parity.FLOOR_BAR)
    parity.check("sod split cg systems x", worst_x, parity.FLOOR_BAR)


def test_blas_helpers_match_reference():
    d = np.load(os.path.join(GOLDEN, "solver_poisson6s.npz"))
    a = CSR(d["row_ptr"], d["col"], d["val"])
    with _solver(a, d["b"], d["x0"]) as s:
        assert _rel(s.gemv(-1.0, d["x0"], 1.0, d["b"]), d["gemv_r"]) <= 1e-14
        assert s.norm(d["b"], d["b"]) == pytest.approx(float(d["norm_bb"]), rel=1e-13)
        assert s.abs_sum(d["b"]) == pytest.approx(float(d["abssum_b"]), rel=1e-13)
        assert s.sum(d["b"]) == pytest.approx(float(np.sum(d["b"])), rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("kernels", [1, 2])
@pytest.mark.parametrize("reorder,rpl", [(1, 1), (1, 2), (1, 4), (0, 1), (0, 2), (0, 4)])
def test_cg_tolerance_fp64_matches_oracle(reorder, rpl, kernels):
    a, b, _ = gen.poisson3d(16, seed_mesh=1, seed_rhs=2)
    oracle.set_threads(1)
    it_o, x_o, h_o, tr_o = oracle.cg_tol(a, b, tol=1e-8, maxit=5000)
    with _solver(a, b) as s:
        rep = s.solve("cg", "tolerance", "fp64", tol=1e-8, max_iter=5000, reorder=reorder, rows_per_lane=rpl,
                      cg_kernels=kernels)
        x = s.get_x()
        h = s.history()
    assert rep["status"] == 0, rep
    assert abs(rep["iterations"] - it_o) <= 2
    assert rep["true_relres"] <= 1.2e-8
    # residual histories agree closely until rounding differences accumulate
    k = min(50, len(h), len(h_o))
    assert np.allclose(h[:k], h_o[:k], rtol=1e-7)
    r = b - a.matvec(x)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.2e-8
    assert _rel(x, x_o) <= 1e-5


def test_cg_mixed_reaches_fp64_tolerance():
    a, b, _ = gen.poisson3d(24, seed_mesh=1, seed_rhs=2)
    with _solver(a, b) as s:
        rep64 = s.solve("cg", "tolerance", "fp64", tol=1e-8)
        x64 = s.get_x()
        s.set_x(np.zeros(a.n))
        rep = s.solve("cg", "tolerance", "mixed", tol=1e-8)
        xm = s.get_x()
    assert rep["status"] == 0, rep
    assert rep["replacements"] >= 1
    r = b - a.matvec(xm)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-8
    assert rep["iterations"] <= 1.5 * rep64["iterations"]
    # error against the FP64 solution, bounded by the conditioning of the
    # near-null boundary-leak mode
    assert np.linalg.norm(xm - x64) / np.linalg.norm(x64) <= 1e-5


def test_cg_deterministic_and_graph_invariant():
    a, b, _ = gen.poisson3d(20, seed_mesh=3, seed_rhs=4)
    xs = []
    with _solver(a, b) as s:
        for graph in (1, 1, 0):
            s.set_x(np.zeros(a.n))
            s.solve("cg", "tolerance", "fp64", tol=1e-9, use_graph=graph, check_every=16)
            xs.append(s.get_x())
    assert np.array_equal(xs[0], xs[1])
    assert np.array_equal(xs[0], xs[2])


def test_cg_sentinel_input_equals_clean_input():
    a, b, _ = gen.poisson3d(10, sentinels=True)
    a2, b2, _ = gen.poisson3d(10, sentinels=False)
    with _solver(a, b) as s1, _solver(a2, b2) as s2:
        s1.solve("cg", "tolerance", "fp64", tol=1e-9)
        s2.solve("cg", "tolerance", "fp64", tol=1e-9)
        assert np.array_equal(s1.get_x(), s2.get_x())


def test_api_errors_are_reported():
    a, b, _ = gen.poisson3d(4)
    s = Solver(a.row_ptr, a.col)
    with pytest.raises(OcslError) as e:
        s.solve("cg", "tolerance", "fp64")
    assert e.value.status == -5  # values / rhs not set
    s.set_values(a.val)
    s.set_rhs(b)
    with pytest.raises(OcslError):
        s.solve("cg", "reference", "mixed")
    bad = a.col.copy()
    bad[0], bad[1] = bad[1], bad[0]
    with pytest.raises(OcslError) as e:
        Solver(a.row_ptr, bad)
    assert e.value.status == -2
    s.close()


def test_cg_zero_rhs_returns_zero():
    a, _, _ = gen.poisson3d(6)
    with _solver(a, np.zeros(a.n)) as s:
        rep = s.solve("cg", "tolerance", "fp64")
        assert rep["iterations"] == 0
        assert np.all(s.get_x() == 0)


def test_cg_full_size_headline_system():
    """128^3 permuted pressure system: the FP64 residual of both precisions is
    checked by an independent host SpMV."""
    a, b, _ = gen.poisson3d(128, seed_mesh=1, seed_rhs=2)
    with _solver(a, b) as s:
        rep64 = s.solve("cg", "tolerance", "fp64", tol=1e-8)
        x64 = s.get_x()
        s.set_x(np.zeros(a.n))
        repm = s.solve("cg", "tolerance", "mixed", tol=1e-8)
        xm = s.get_x()
    for x, rep in ((x64, rep64), (xm, repm)):
        assert rep["status"] == 0, rep
        rel = np.linalg.norm(b - a.matvec(x)) / np.linalg.norm(b)
        assert rel <= 1.05e-8, (rel, rep)
    # the oracle takes 1993 iterations on this system
    assert abs(rep64["iterations"] - 1993) <= 20


@pytest.mark.parametrize("kernels", [1, 2, 3])
@pytest.mark.parametrize("n", [5, 7, 13])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cg_rows_per_lane_layouts_agree(n, precision, kernels):
    """One, two and four rows per lane: the same per-row SpMV sums, so the
    solves agree to the dot-product rounding; odd row counts end inside a row
    group."""
    a, b, _ = gen.poisson3d(n, seed_mesh=3, seed_rhs=4, sentinels=(n == 7))
    tol = 1e-10 if precision == "fp64" else 1e-8
    out = {}
    with _solver(a, b) as s:
        for rpl in (1, 2, 4):
            s.set_x(np.zeros(a.n))
            rep = s.solve("cg", "tolerance", precision, tol=tol, rows_per_lane=rpl, cg_kernels=kernels)
            assert rep["status"] == 0, (rpl, rep)
            out[rpl] = (rep, s.get_x())
    # FP64 stops on the recursive residual: the true one may sit a few times
    # above the tolerance at 1e-10 on these near-singular systems
    for rpl in (1, 2, 4):
        assert np.linalg.norm(b - a.matvec(out[rpl][1])) / np.linalg.norm(b) <= 10 * tol
        if precision == "fp64":
            assert abs(out[1][0]["iterations"] - out[rpl][0]["iterations"]) <= 2
        assert _rel(out[rpl][1], out[1][1]) <= (1e-6 if precision == "fp64" else 1e-4)


@pytest.mark.parametrize("kernels", [1, 2])
@pytest.mark.parametrize("max_iter", [1, 2, 7, 8, 32, 33])
@pytest.mark.parametrize("graph", [1, 0])
def test_cg_iteration_count_cut(max_iter, graph, kernels):
    """FP64 CG stops at max_iter whatever its parity (the odd one-kernel
    iteration then adds the pending step of its pair) and records the
    residual of the last iterate: x and the history after k iterations are
    the oracle's after k textbook CG iterations."""
    a, b, _ = gen.poisson3d(12, seed_mesh=5, seed_rhs=6)
    oracle.set_threads(1)
    it_o, x_o, h_o, _ = oracle.cg_tol(a, b, tol=1e-30, maxit=max_iter)
    assert it_o == max_iter
    with _solver(a, b) as s:
        rep = s.solve("cg", "tolerance", "fp64", tol=1e-30, max_iter=max_iter, cg_kernels=kernels,
                      use_graph=graph, check_every=4)
        x = s.get_x()
        h = s.history()
    assert rep["iterations"] == max_iter and rep["status"] == 1, rep  # OCSL_WARN_MAXITER
    assert len(h) == max_iter + 1
    assert np.allclose(h, h_o[:max_iter + 1], rtol=1e-9)
    assert _rel(x, x_o) <= 1e-10


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("check_every", [2, 3, 8, 32])
def test_cg_one_kernel_and_split_forms_agree(precision, check_every):
    """Same system, every iteration form (one kernel with the LDS window, split,
    one kernel gathering from memory), chunk lengths that put host checks (and
    MIXED residual replacements) after even and odd iteration counts."""
    a, b, _ = gen.poisson3d(20, seed_mesh=7, seed_rhs=8)
    tol = 1e-10 if precision == "fp64" else 1e-9
    res = {}
    with _solver(a, b) as s:
        for kernels in (1, 2, 3):
            s.set_x(np.zeros(a.n))
            rep = s.solve("cg", "tolerance", precision, tol=tol, cg_kernels=kernels, check_every=check_every)
            assert rep["status"] == 0, (kernels, rep)
            res[kernels] = (rep, s.get_x())
    for kernels in (1, 2, 3):
        rel = np.linalg.norm(b - a.matvec(res[kernels][1])) / np.linalg.norm(b)
        assert rel <= (10 * tol if precision == "fp64" else tol), (kernels, rel)
        assert abs(res[kernels][0]["iterations"] - res[2][0]["iterations"]) <= max(3, res[2][0]["iterations"] // 20)
        assert _rel(res[kernels][1], res[2][1]) <= (1e-7 if precision == "fp64" else 1e-5)


def test_cg_lds_window_overflow_takes_gather_path():
    """Input row order on a permuted mesh: every block's row window spans
    nearly all 64000 rows, more than the LDS holds, so the LDS form falls back
    to gathering; it must agree with the gathering form."""
    a, b, _ = gen.poisson3d(40, seed_mesh=1, seed_rhs=2)
    res = {}
    with _solver(a, b) as s:
        for kernels in (1, 3):
            s.set_x(np.zeros(a.n))
            rep = s.solve("cg", "tolerance", "mixed", tol=1e-8, cg_kernels=kernels, reorder=0)
            assert rep["status"] == 0, (kernels, rep)
            res[kernels] = (rep, s.get_x())
    for kernels in (1, 3):
        assert np.linalg.norm(b - a.matvec(res[kernels][1])) / np.linalg.norm(b) <= 1e-8
    assert abs(res[1][0]["iterations"] - res[3][0]["iterations"]) <= 5
    assert _rel(res[1][1], res[3][1]) <= 1e-5


@pytest.mark.parametrize("max_iter,tol", [(7, 1e-30), (13, 1e-30), (33, 1e-30), (20000, 1e-9)])
@pytest.mark.parametrize("check_every", [8, 32])
def test_cg_mixed_x_period_is_bitwise_neutral(max_iter, tol, check_every):
    """MIXED split CG adds alpha_k p_k to the FP64 x once per period of 2, 4
    or 8 iterations, in iteration order: the same additions as every
    iteration, so x, the iteration count and the history are bitwise equal
    for every period, also when max_iter, convergence or a residual
    replacement cuts a period.  (Host checks, and so replacements, fall
    after whole chunks: check_every is a multiple of every period here.)"""
    a, b, _ = gen.poisson3d(14, seed_mesh=2, seed_rhs=3)
    out = {}
    with _solver(a, b) as s:
        for per in (2, 4, 8):
            s.set_x(np.zeros(a.n))
            rep = s.solve("cg", "tolerance", "mixed", tol=tol, max_iter=max_iter, cg_kernels=2, x_period=per,
                          check_every=check_every)
            out[per] = (rep, s.get_x(), s.history())
    for per in (4, 8):
        assert out[per][0]["iterations"] == out[2][0]["iterations"]
        assert out[per][0]["replacements"] == out[2][0]["replacements"]
        assert np.array_equal(out[per][1], out[2][1])
        assert np.array_equal(out[per][2], out[2][2])
    if tol > 1e-20:
        assert out[2][0]["status"] == 0
        assert np.linalg.norm(b - a.matvec(out[2][1])) / np.linalg.norm(b) <= tol
    else:
        assert out[2][0]["iterations"] == max_iter


def test_cg_one_kernel_timing_hook():
    a, b, _ = gen.poisson3d(32, seed_mesh=1, seed_rhs=2)
    with _solver(a, b) as s:
        x0 = s.get_x()
        for precision in ("fp64", "mixed"):
            ms, by = s.time_cg_iteration(precision, pairs=5)
            assert all(m > 0 for m in ms) and by[1] > by[0] > 0
        assert np.array_equal(s.get_x(), x0)


def test_cg_autotuner_keeps_x_and_tuned_solves_converge():
    a, b, _ = gen.poisson3d(24, seed_mesh=1, seed_rhs=2)
    x0 = gen.uniform(9, a.n)
    with _solver(a, b, x0) as s:
        best, us, table = s.tune_cg("mixed", iterations=8)
        assert np.array_equal(s.get_x(), x0)
        assert best["rows_per_lane"] in (1, 2, 4) and best["block_size"] in (256, 512, 1024)
        assert best["cg_kernels"] in (1, 2, 3)
        assert {t["cg_kernels"] for t in table} == {1, 2, 3}
        assert len(table) >= 36 and all(t["us_per_iteration"] > 0 for t in table)
        # the pick is the fastest of the sweep's leaders re-timed over more
        # iterations: one of the table's configurations
        assert us > 0
        keys = ("cg_kernels", "rows_per_lane", "block_size", "grid_blocks", "vec_blocks")
        assert any(all(t[k] == best[k] for k in keys) for t in table)
        s.set_x(np.zeros(a.n))
        rep = s.solve("cg", "tolerance", "mixed", tol=1e-8)
        assert rep["status"] == 0, rep
        assert np.linalg.norm(b - a.matvec(s.get_x())) / np.linalg.norm(b) <= 1e-8
        s.set_tuned_cg("mixed")  # cleared: defaults again
        s.set_x(np.zeros(a.n))
        assert s.solve("cg", "tolerance", "mixed", tol=1e-8)["status"] == 0


@pytest.mark.parametrize("kernels", [1, 2])
@pytest.mark.parametrize("check_every", [1, 8, 32])
def test_cg_mixed_replacement_right_after_trigger(kernels, check_every):
    # Pressure systems with a near-null constant mode (smallest eigenvalue
    # ~1e-6, condition ~1e7) and right-hand sides that excite it: the FP32
    # recurrence drifts fast, and iterations run on past the replacement
    # trigger (to the end of a chunk) let it diverge for good on the 27-row
    # system.  The chunk stops at the trigger, so every chunk length and both
    # kernel forms converge with the same number of iterations.
    its = []
    for n in (3, 5, 7):
        a, _, _ = gen.poisson3d(n, seed_mesh=100, seed_rhs=101)
        b = gen.uniform(70, a.n)
        with Solver(a.row_ptr, a.col) as s:
            s.set_values(a.val)
            s.set_rhs(b)
            s.set_x(np.zeros(a.n))
            rep = s.solve("cg", "tolerance", "mixed", tol=1e-8, check_every=check_every, max_iter=3000,
                          cg_kernels=kernels)
            x = s.get_x()
        assert rep["status"] == 0, (n, rep)
        assert np.linalg.norm(b - a.matvec(x)) / np.linalg.norm(b) <= 1.05e-8
        assert rep["replacements"] >= 1
        its.append(rep["iterations"])
    # the same runs with the shortest chunk (same kernel form: the forms add
    # their dot products in different orders, which moves the count a little
    # on systems this ill-conditioned)
    ref = []
    for n in (3, 5, 7):
        a, _, _ = gen.poisson3d(n, seed_mesh=100, seed_rhs=101)
        b = gen.uniform(70, a.n)
        with Solver(a.row_ptr, a.col) as s:
            s.set_values(a.val)
            s.set_rhs(b)
            s.set_x(np.zeros(a.n))
            ref.append(s.solve("cg", "tolerance", "mixed", tol=1e-8, check_every=1, max_iter=3000,
                               cg_kernels=kernels)["iterations"])
    assert all(abs(u - v) <= 2 for u, v in zip(its, ref)), (its, ref)


def test_stream_triad_timing_hook():
    import ctypes

    from opencslsolver_amd._native import hip_lib

    lib = hip_lib()
    ms, by = ctypes.c_double(), ctypes.c_double()
    assert lib.ocsl_time_stream_triad(1 << 22, 10, ctypes.byref(ms), ctypes.byref(by)) == 0
    assert by.value == 24.0 * (1 << 22) and ms.value > 0.0
    gbps = by.value / (ms.value * 1e-3) / 1e9
    assert 100.0 < gbps < 20000.0, gbps
    assert lib.ocsl_time_stream_triad(3, 10, ctypes.byref(ms), ctypes.byref(by)) != 0  # odd length
