"""Differentiable batched block-tridiagonal solve.

``block_tridiagonal_solve(plan, a, b, c, d)`` returns x with T x = d for the batch the
plan describes (arrays in the plan's layout, flat float64 device tensors). Its backward
pass is the hand-written adjoint kernel behind the C ABI (btd_solve_adjoint): it reuses
the pivoted block factors saved by the forward solve to compute lambda = T^{-T} dL/dx,
and returns dL/dd = lambda and dL/dA_q = -lambda_q x_{q-1}^T, dL/dB_q = -lambda_q x_q^T,
dL/dC_q = -lambda_q x_{q+1}^T (zero for the ignored A of the first row and C of the last).
"""
from __future__ import annotations

import torch

from . import BtdmaGpuPlan, layout_elems, solve_adjoint, solve_noncyclic_save


class BlockTridiagonalSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan: BtdmaGpuPlan, a, b, c, d):
        if plan.npart != 1:
            raise ValueError("the differentiable solve needs a plan with npart == 1")
        a, b = a.contiguous(), b.contiguous()
        cw = c.detach().clone()       # the solve overwrites C with C'
        x = d.detach().clone()        # and d with x
        factors = torch.empty_like(cw)
        pivots = torch.empty(layout_elems(plan.layout, plan.nsys, plan.nrow, 1),
                             dtype=torch.int64, device=d.device)
        solve_noncyclic_save(plan, a.detach(), b.detach(), cw, x, factors, pivots)
        ctx.plan = plan
        ctx.save_for_backward(a.detach(), cw, factors, pivots, x)
        return x

    @staticmethod
    def backward(ctx, w):
        a, cw, factors, pivots, x = ctx.saved_tensors
        w = w.contiguous()
        lam = torch.empty_like(x)
        da, db, dc = torch.empty_like(a), torch.empty_like(a), torch.empty_like(a)
        solve_adjoint(ctx.plan, a, cw, factors, pivots, x, w, lam, da, db, dc)
        return None, da, db, dc, lam


def block_tridiagonal_solve(plan: BtdmaGpuPlan, a, b, c, d):
    """x = T^{-1} d, differentiable in a, b, c and d."""
    return BlockTridiagonalSolve.apply(plan, a, b, c, d)
