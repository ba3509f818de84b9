"""torch.autograd binding of the sparse solve x = A^{-1} b.

The forward pass runs the HIP solver on the tensor's device; the backward
pass runs the hand-written adjoint (transpose solve A^T lambda = dL/dx, then
dL/db = lambda and dL/dA_ij = -lambda_i x_j on the CSR pattern).  Values,
right-hand side and solution are float64 CUDA tensors; the CSR structure
(the reference's u64 row_ptr / column_index layout) is fixed per Solver.
"""
from __future__ import annotations

import torch

from .solver import Solver


class SparseSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, solver: Solver, method: str, precision: str, opts: dict):
        if not (values.is_cuda and b.is_cuda):
            raise TypeError("SparseSolve needs CUDA tensors")
        # private copies: the handle keeps device pointers to them
        values = values.detach().to(torch.float64).clone()
        b = b.detach().to(torch.float64).clone()
        x = torch.zeros_like(b)
        solver.set_values(values)
        solver.set_rhs(b)
        solver.set_x(x)
        rep = solver.solve(method, "tolerance", precision, **opts)
        if rep["status"] != 0:
            raise RuntimeError(f"forward solve did not converge: {rep}")
        solver.get_x(out=x)
        ctx.save_for_backward(values, x)
        ctx.solver, ctx.method, ctx.precision, ctx.opts = solver, method, precision, opts
        return x

    @staticmethod
    def backward(ctx, gx):
        values, x = ctx.saved_tensors
        solver = ctx.solver
        g = gx.detach().contiguous().to(torch.float64)
        # the handle may have solved something else since the forward pass
        solver.set_values(values)
        solver.set_x(x)
        dLdb, dLdA, rep = solver.adjoint(g, ctx.method, ctx.precision, **ctx.opts)
        if rep["status"] != 0:
            raise RuntimeError(f"adjoint solve did not converge: {rep}")
        return (dLdA if ctx.needs_input_grad[0] else None, dLdb if ctx.needs_input_grad[1] else None,
                None, None, None, None)


def sparse_solve(solver: Solver, values, b, method="cg", precision="fp64", **opts):
    """x = A(values)^{-1} b, differentiable in ``values`` and ``b``."""
    return SparseSolve.apply(values, b, solver, method, precision, opts)
