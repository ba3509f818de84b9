"""Batched independent systems on the GPU (one set of launches for all)."""
import numpy as np
import pytest

import oracle
from opencslsolver_amd import gen
from opencslsolver_amd.solver import Batch, Solver

pytestmark = pytest.mark.gpu


def _systems(sizes, seed0=100):
    out = []
    for i, n in enumerate(sizes):
        a, b, _ = gen.poisson3d(int(n), seed_mesh=seed0 + 2 * i, seed_rhs=seed0 + 2 * i + 1)
        out.append((a, b))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_single_solves(precision):
    sysl = _systems([6, 11, 17, 9, 23, 14])
    with Batch([(a.row_ptr, a.col) for a, _ in sysl]) as bt:
        bt.set_values([a.val for a, _ in sysl])
        bt.set_rhs([b for _, b in sysl])
        xs, reps, lat = bt.solve(precision, tol=1e-8, check_every=8)
    for k, (a, b) in enumerate(sysl):
        rel = np.linalg.norm(b - a.matvec(xs[k])) / np.linalg.norm(b)
        assert reps[k]["status"] == 0, (k, reps[k])
        assert rel <= 1.05e-8, (k, rel)
        assert reps[k]["true_relres"] == pytest.approx(rel, rel=1e-3, abs=1e-12)
        _, x_o, _, _ = oracle.cg_tol(a, b, tol=1e-10)
        assert np.linalg.norm(xs[k] - x_o) / np.linalg.norm(x_o) <= 1e-5
        assert lat[k] > 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_device_resident_guesses(precision):
    """Guesses and solutions in device memory (torch tensors, solved in
    place) give bitwise the solutions of host arrays."""
    import torch

    sysl = _systems([7, 13, 10, 19])
    rng = np.random.default_rng(5)
    x0 = [rng.standard_normal(a.n) * 0.1 for a, _ in sysl]
    with Batch([(a.row_ptr, a.col) for a, _ in sysl]) as bt:
        bt.set_values([a.val for a, _ in sysl])
        bt.set_rhs([b for _, b in sysl])
        xs, reps, _ = bt.solve(precision, x0=x0, tol=1e-8, check_every=8)
        xd = [torch.tensor(x, dtype=torch.float64, device="cuda") for x in x0]
        out, reps_d, _ = bt.solve(precision, x0=xd, tol=1e-8, check_every=8)
        assert all(o is t for o, t in zip(out, xd))
        for k in range(len(sysl)):
            assert np.array_equal(xs[k], xd[k].cpu().numpy()), k
            assert reps_d[k]["iterations"] == reps[k]["iterations"]
        with pytest.raises(ValueError):
            bt.solve(precision, x0=[x[:-1] for x in xd], tol=1e-8)


def test_batch_zero_rhs_and_reuse():
    sysl = _systems([5, 8, 12])
    bs = [b for _, b in sysl]
    bs[1] = np.zeros_like(bs[1])
    with Batch([(a.row_ptr, a.col) for a, _ in sysl]) as bt:
        bt.set_values([a.val for a, _ in sysl])
        bt.set_rhs(bs)
        xs, reps, _ = bt.solve("mixed", tol=1e-8)
        assert np.all(xs[1] == 0.0) and reps[1]["iterations"] == 0
        # second solve on the same batch with a new right-hand side
        bs2 = [gen.uniform(50 + k, a.n) for k, (a, _) in enumerate(sysl)]
        bt.set_rhs(bs2)
        xs2, reps2, _ = bt.solve("fp64", tol=1e-9)
    for k, (a, _) in enumerate(sysl):
        assert np.linalg.norm(bs2[k] - a.matvec(xs2[k])) / np.linalg.norm(bs2[k]) <= 1.5e-9


def test_batch_with_sentinel_layout_and_unequal_structures():
    a1, b1, _ = gen.poisson3d(7, sentinels=True)
    a2, _ = gen.convdiff3d(6, sentinels=True)  # not SPD: CG may not converge, must not fault
    a3, b3, _ = gen.poisson3d(10)
    b2 = gen.uniform(9, a2.n)
    with Batch([(a.row_ptr, a.col) for a in (a1, a2, a3)]) as bt:
        bt.set_values([a1.val, a2.val, a3.val])
        bt.set_rhs([b1, b2, b3])
        xs, reps, _ = bt.solve("fp64", tol=1e-8, max_iter=400)
    for a, b, x, r in ((a1, b1, xs[0], reps[0]), (a3, b3, xs[2], reps[2])):
        assert np.linalg.norm(b - a.matvec(x)) / np.linalg.norm(b) <= 1.5e-8


@pytest.mark.parametrize("rpl,block", [(1, 256), (2, 512), (4, 1024), (4, 256), (2, 1024)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_launch_shapes_match_oracle(rpl, block, precision):
    # sizes whose row counts are not multiples of 64 R, a 1-slice system and
    # one spanning several workgroups at every shape
    sysl = _systems([3, 7, 13, 26, 5])
    with Batch([(a.row_ptr, a.col) for a, _ in sysl]) as bt:
        bt.set_values([a.val for a, _ in sysl])
        bt.set_rhs([b for _, b in sysl])
        bt.set_launch(rpl, block)
        assert bt.launch() == {"rows_per_lane": rpl, "block_size": block}
        xs, reps, _ = bt.solve(precision, tol=1e-8, check_every=4)
        # values and rhs set after the shape change land in the new layout too
        bs2 = [gen.uniform(70 + k, a.n) for k, (a, _) in enumerate(sysl)]
        bt.set_rhs(bs2)
        xs2, reps2, _ = bt.solve(precision, tol=1e-8)
    for k, (a, b) in enumerate(sysl):
        assert reps[k]["status"] == 0, (k, reps[k])
        assert np.linalg.norm(b - a.matvec(xs[k])) / np.linalg.norm(b) <= 1.05e-8
        _, x_o, _, _ = oracle.cg_tol(a, b, tol=1e-10)
        assert np.linalg.norm(xs[k] - x_o) / np.linalg.norm(x_o) <= 1e-5
        assert np.linalg.norm(bs2[k] - a.matvec(xs2[k])) / np.linalg.norm(bs2[k]) <= 1.05e-8


def test_batch_shape_change_keeps_fp64_iterations():
    # FP64 CG in exact arithmetic is shape-independent; in floating point the
    # reductions change order, so iteration counts agree to within one or two
    sysl = _systems([9, 15, 21])
    its = {}
    for rpl in (1, 2, 4):
        with Batch([(a.row_ptr, a.col) for a, _ in sysl]) as bt:
            bt.set_values([a.val for a, _ in sysl])
            bt.set_rhs([b for _, b in sysl])
            bt.set_launch(rpl, 512)
            _, reps, _ = bt.solve("fp64", tol=1e-9, check_every=1)
            its[rpl] = [r["iterations"] for r in reps]
    for rpl in (2, 4):
        assert all(abs(u - v) <= 2 for u, v in zip(its[1], its[rpl])), its


def test_batch_autotuner_applies_fastest_shape():
    sysl = _systems([8, 14, 19, 11])
    with Batch([(a.row_ptr, a.col) for a, _ in sysl]) as bt:
        bt.set_values([a.val for a, _ in sysl])
        bt.set_rhs([b for _, b in sysl])
        best, us, table = bt.tune("mixed", iterations=8)
        assert len(table) == 9
        assert {(t["rows_per_lane"], t["block_size"]) for t in table} == \
            {(r, b) for r in (1, 2, 4) for b in (256, 512, 1024)}
        assert all(t["us_per_iteration"] > 0 for t in table)
        assert us == min(t["us_per_iteration"] for t in table)
        assert bt.launch() == best
        assert bt.time_iterations("fp64", 8) > 0
        xs, reps, _ = bt.solve("mixed", tol=1e-8)
    for k, (a, b) in enumerate(sysl):
        assert reps[k]["status"] == 0
        assert np.linalg.norm(b - a.matvec(xs[k])) / np.linalg.norm(b) <= 1.05e-8


@pytest.mark.parametrize("check_every", [1, 8, 32, 64])
def test_batch_mixed_replacement_independent_of_check_every(check_every):
    # Pressure systems with a near-null constant mode (smallest eigenvalue
    # ~1e-6, condition ~1e7) and right-hand sides that excite it: FP32
    # recurrences drift fast, so a residual replacement that waited for the
    # end of a chunk let the recurrence diverge.  The replacement now takes
    # the iteration after its trigger, so the iteration counts do not depend
    # on the chunk length.
    sysl = [gen.poisson3d(n, seed_mesh=100, seed_rhs=101)[0] for n in (3, 5, 7, 13)]
    bs = [gen.uniform(70 + k, a.n) for k, a in enumerate(sysl)]
    with Batch([(a.row_ptr, a.col) for a in sysl]) as bt:
        bt.set_values([a.val for a in sysl])
        bt.set_rhs(bs)
        xs, reps, _ = bt.solve("mixed", tol=1e-8, check_every=check_every, max_iter=2000)
        _, ref, _ = bt.solve("mixed", tol=1e-8, check_every=1, max_iter=2000)
    for k, (a, b) in enumerate(zip(sysl, bs)):
        assert reps[k]["status"] == 0, (k, reps[k])
        assert np.linalg.norm(b - a.matvec(xs[k])) / np.linalg.norm(b) <= 1.05e-8
        assert reps[k]["replacements"] >= 1
        assert abs(reps[k]["iterations"] - ref[k]["iterations"]) <= 3, (k, reps[k], ref[k])


@pytest.mark.parametrize("check_every", [1, 5, 32])
def test_batch_x_period_is_bitwise_neutral(check_every):
    # x takes the period's steps at its end, before each FP64 residual
    # replacement and when a system stops, one FP64 addition per step in
    # iteration order: every solution, iteration count and replacement count
    # equals that of an update every iteration (x_period 1), for chunk
    # lengths that cut periods anywhere.  The systems include near-null-mode
    # pressure systems (several replacements) and a zero right-hand side.
    sysl = [gen.poisson3d(n, seed_mesh=100, seed_rhs=101)[0] for n in (3, 5, 7, 13)]
    sysl += [a for a, _ in _systems([9, 17, 24], seed0=300)]
    bs = [gen.uniform(70 + k, a.n) for k, a in enumerate(sysl)]
    bs[5] = np.zeros_like(bs[5])
    runs = {}
    with Batch([(a.row_ptr, a.col) for a in sysl]) as bt:
        bt.set_values([a.val for a in sysl])
        bt.set_rhs(bs)
        for m in (1, 2, 8):
            runs[m] = bt.solve("mixed", tol=1e-8, check_every=check_every, max_iter=3000, x_period=m)
    xs1, reps1, _ = runs[1]
    assert sum(r["replacements"] for r in reps1) >= 4
    for m in (2, 8):
        xs, reps, _ = runs[m]
        for k in range(len(sysl)):
            assert reps[k]["iterations"] == reps1[k]["iterations"], (m, k, reps[k], reps1[k])
            assert reps[k]["replacements"] == reps1[k]["replacements"], (m, k)
            assert np.array_equal(xs[k], xs1[k]), (m, k, np.max(np.abs(xs[k] - xs1[k])))
    for k, (a, b) in enumerate(zip(sysl, bs)):
        assert reps1[k]["status"] == 0, (k, reps1[k])
        nb = np.linalg.norm(b)
        assert np.linalg.norm(b - a.matvec(xs1[k])) <= 1.05e-8 * nb, k


def test_batch_x_period_at_the_iteration_cap():
    # systems stopped by max_iter in the middle of a period: x holds every
    # step taken, as with an update every iteration
    sysl = [a for a, _ in _systems([10, 19, 27], seed0=400)]
    bs = [gen.uniform(80 + k, a.n) for k, a in enumerate(sysl)]
    out = {}
    with Batch([(a.row_ptr, a.col) for a in sysl]) as bt:
        bt.set_values([a.val for a in sysl])
        bt.set_rhs(bs)
        for m in (1, 8):
            out[m] = bt.solve("mixed", tol=1e-12, check_every=4, max_iter=21, x_period=m)
    for k in range(len(sysl)):
        assert out[8][1][k]["iterations"] == out[1][1][k]["iterations"] == 21
        assert np.array_equal(out[8][0][k], out[1][0][k]), k
        assert np.linalg.norm(out[1][0][k]) > 0


def test_batch_full_plan_64_systems():
    """The full uneven batch: 64 pressure systems of n^3 cells, n log-uniform
    in [22, 80] from seed 7 (about 7.5 M cells), MIXED CG to 1e-8 in one
    batched solve.  Every system's FP64 relative residual is recomputed by a
    host SpMV; the smallest, the median and the largest systems are compared
    with the oracle's FP64 solutions."""
    import parity

    oracle.set_threads(8)
    ns, sm, sr = gen.batch_plan(7, 64, 22, 80)
    assert len(ns) == 64 and min(ns) >= 22 and max(ns) <= 80
    sysl = [gen.poisson3d(int(ns[k]), seed_mesh=int(sm[k]), seed_rhs=int(sr[k]))[:2] for k in range(64)]
    with Batch([(a.row_ptr, a.col) for a, _ in sysl]) as bt:
        bt.set_values([a.val for a, _ in sysl])
        bt.set_rhs([b for _, b in sysl])
        xs, reps, lat = bt.solve("mixed", tol=1e-8)
    worst = 0.0
    for k, (a, b) in enumerate(sysl):
        assert reps[k]["status"] == 0, (k, reps[k])
        worst = max(worst, float(np.linalg.norm(b - a.matvec(xs[k])) / np.linalg.norm(b)))
    parity.check("batch64 worst host relres", worst, 1.05e-8)
    order = np.argsort([a.n for a, _ in sysl])
    for k in (order[0], order[32], order[-1]):
        a, b = sysl[k]
        _, x_o, _, _ = oracle.cg_tol(a, b, tol=1e-10)
        err = float(np.linalg.norm(xs[k] - x_o) / np.linalg.norm(x_o))
        parity.check(f"batch64 system {k} (n={a.n}) error vs oracle fp64", err, 1e-5)
