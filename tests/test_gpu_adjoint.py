"""Adjoint (transpose solve + sampled outer product) on the GPU, against the
oracle's FP64 CPU gradient, dense solves and central finite differences."""
import numpy as np
import pytest
import torch

import oracle
from opencslsolver_amd import gen
from opencslsolver_amd.autograd import sparse_solve
from opencslsolver_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _row_of(a, p):
    return int(np.searchsorted(a.row_ptr, p, side="right") - 1)


@pytest.mark.parametrize("n,sentinels", [(7, True), (33, True), (40, False)])
def test_sampled_outer_is_exact(n, sentinels):
    """Small and multi-workgroup systems, sentinel slots included: bit-exact."""
    a, _ = gen.convdiff3d(n, sentinels=sentinels)
    u, v = gen.normal(1, a.n), gen.normal(2, a.n)
    with Solver(a.row_ptr, a.col) as s:
        out = s.sampled_outer(-1.0, u, v)
    assert np.array_equal(out, oracle.sampled_outer(a, u, v))


def test_adjoint_pressure_system_vs_oracle_and_fd():
    """64^3 permuted pressure system (mesh seed 8, b from the continued
    stream), L = 0.5 ||x - x_t||^2 with x_t ~ N(0,1) from seed 9."""
    oracle.set_threads(16)
    a, b, _ = gen.poisson3d(64, seed_mesh=8, seed_rhs=None)
    xt = gen.normal(9, a.n)
    with Solver(a.row_ptr, a.col) as s:
        s.set_values(a.val)
        s.set_rhs(b)
        s.set_x(np.zeros(a.n))
        rep = s.solve("cg", "tolerance", "fp64", tol=1e-12)
        assert rep["status"] == 0, rep
        x = s.get_x()
        g = x - xt
        lam, dA, arep = s.adjoint(g, "cg", "fp64", tol=1e-12)
        assert arep["status"] == 0, arep
        # FP64 CPU reference gradient (oracle transpose + CG + outer product)
        _, x_o, _, _ = oracle.cg_tol(a, b, tol=1e-13)
        _, lam_o, _, _ = oracle.cg_tol(oracle.transpose(a), x_o - xt, tol=1e-13)
        dA_o = oracle.sampled_outer(a, lam_o, x_o)
        assert _rel(lam, lam_o) <= 1e-6
        assert _rel(dA, dA_o) <= 1e-6
        # central differences, h = 1e-6, on 32 seeded entries.  The perturbed
        # solutions are exact rank-one updates of x (Sherman-Morrison):
        # (A + t e_i e_j^T)^{-1} b = x - t x_j w / (1 + t w_i), A w = e_i, and
        # x(b + t e_i) = x + t w; L is differenced without cancellation.
        h = 1e-6
        slots = gen.sample_slots(9, a, 32)
        gmax = np.abs(dA_o[slots]).max()
        for p in slots:
            i, j = _row_of(a, p), int(a.col[p])
            e = np.zeros(a.n)
            e[i] = 1.0
            s.set_rhs(e)
            s.set_x(np.zeros(a.n))
            assert s.solve("cg", "tolerance", "fp64", tol=1e-13)["status"] == 0
            w = s.get_x()
            dp = -h * x[j] * w / (1.0 + h * w[i])
            dm = h * x[j] * w / (1.0 - h * w[i])
            fd = (g @ (dp - dm) + 0.5 * (dp @ dp - dm @ dm)) / (2 * h)
            assert abs(fd - dA[p]) <= 1e-6 * max(abs(fd), 1e-3 * gmax), (p, fd, dA[p])
            fdb = (g @ (2 * h * w)) / (2 * h)
            assert abs(fdb - lam[i]) <= 1e-6 * max(abs(fdb), 1e-3 * np.abs(lam).max()), (i, fdb, lam[i])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_adjoint_nonsymmetric_vs_dense(precision):
    a, _ = gen.convdiff3d(10, seed_perm=3, dt=10.0)
    b = gen.uniform(4, a.n)
    xt = gen.normal(9, a.n)
    dense = a.to_dense()
    x_d = np.linalg.solve(dense, b)
    with Solver(a.row_ptr, a.col) as s:
        s.set_values(a.val)
        s.set_rhs(b)
        s.set_x(np.zeros(a.n))
        assert s.solve("ilu_bicgstab", "tolerance", "fp64", tol=1e-13)["status"] == 0
        x = s.get_x()
        lam, dA, rep = s.adjoint(x - xt, "ilu_bicgstab", precision, tol=1e-11)
    assert rep["status"] == 0, rep
    lam_d = np.linalg.solve(dense.T, x_d - xt)
    assert _rel(lam, lam_d) <= 1e-8
    assert _rel(dA, oracle.sampled_outer(a, lam_d, x_d)) <= 1e-8


def test_adjoint_sentinel_layout():
    a1, b, _ = gen.poisson3d(8, sentinels=True)
    a0, b0, _ = gen.poisson3d(8, sentinels=False)
    assert np.array_equal(b, b0)
    xt = gen.normal(9, a0.n)
    out = []
    for a in (a0, a1):
        with Solver(a.row_ptr, a.col) as s:
            s.set_values(a.val)
            s.set_rhs(b)
            s.set_x(np.zeros(a.n))
            s.solve("cg", "tolerance", "fp64", tol=1e-12)
            x = s.get_x()
            lam, dA, _ = s.adjoint(x - xt, "cg", "fp64", tol=1e-12)
            out.append((lam, dA))
    real = a1.col < np.uint64(a1.n)
    assert np.all(out[1][1][~real] == 0.0)
    assert np.array_equal(out[1][0], out[0][0])
    assert np.array_equal(out[1][1][real], out[0][1])


def test_autograd_function_gradcheck():
    a, _ = gen.convdiff3d(3, seed_perm=3, dt=10.0)
    dev = torch.device("cuda:0")
    vals = torch.tensor(a.val, dtype=torch.float64, device=dev, requires_grad=True)
    b = torch.tensor(gen.uniform(4, a.n), dtype=torch.float64, device=dev, requires_grad=True)
    with Solver(a.row_ptr, a.col) as s:
        f = lambda v, r: sparse_solve(s, v, r, method="ilu_bicgstab", tol=1e-13, max_iter=200)
        assert torch.autograd.gradcheck(f, (vals, b), eps=1e-6, atol=1e-6, rtol=1e-4)


def test_autograd_loss_matches_adjoint():
    a, b_np, _ = gen.poisson3d(12, seed_mesh=8, seed_rhs=None)
    xt_np = gen.normal(9, a.n)
    dev = torch.device("cuda:0")
    vals = torch.tensor(a.val, dtype=torch.float64, device=dev, requires_grad=True)
    b = torch.tensor(b_np, dtype=torch.float64, device=dev, requires_grad=True)
    xt = torch.tensor(xt_np, dtype=torch.float64, device=dev)
    with Solver(a.row_ptr, a.col) as s:
        x = sparse_solve(s, vals, b, method="cg", tol=1e-12)
        loss = 0.5 * ((x - xt) ** 2).sum()
        loss.backward()
    _, x_o, _, _ = oracle.cg_tol(a, b_np, tol=1e-13)
    _, lam_o, _, _ = oracle.cg_tol(oracle.transpose(a), x_o - xt_np, tol=1e-13)
    assert _rel(b.grad.cpu().numpy(), lam_o) <= 1e-8
    assert _rel(vals.grad.cpu().numpy(), oracle.sampled_outer(a, lam_o, x_o)) <= 1e-8
