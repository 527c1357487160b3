"""Differentiable batched block-tridiagonal solve.

``block_tridiagonal_solve(plan, a, b, c, d)`` returns x with T x = d for the batch the
plan describes (arrays in the plan's layout, flat float64 device tensors). Its backward
pass is the hand-written adjoint kernel behind the C ABI (btd_solve_adjoint): it reuses
the pivoted block factors saved by the forward solve to compute lambda = T^{-T} dL/dx,
and returns dL/dd = lambda and dL/dA_q = -lambda_q x_{q-1}^T, dL/dB_q = -lambda_q x_q^T,
dL/dC_q = -lambda_q x_{q+1}^T (zero for the ignored A of the first row and C of the last).

``block_tridiagonal_solve_f32r(plan, a, b, c, d, sweeps)`` is the same for FP32-stored blocks
(float32 a, b, c; float64 d): the forward is btd_solve_noncyclic_f32r, the backward solves
the transposed system from the same FP32 blocks with FP64-residual refinement
(btd_solve_transposed_f32r) and writes FP32 block gradients (btd_outer_grads_f32).
"""
from __future__ import annotations

import torch

from . import (BtdmaGpuPlan, f32r_work_elems, layout_elems, outer_grads_f32, solve_adjoint,
               solve_noncyclic_f32r, solve_noncyclic_save_out, solve_transposed_f32r)


class BlockTridiagonalSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan: BtdmaGpuPlan, a, b, c, d):
        if plan.npart != 1:
            raise ValueError("the differentiable solve needs a plan with npart == 1")
        a, b = a.contiguous(), b.contiguous()
        c, d = c.detach().contiguous(), d.detach().contiguous()
        cw = torch.empty_like(c)      # C' (c and d are only read)
        x = torch.empty_like(d)
        factors = torch.empty_like(cw)
        pivots = torch.empty(layout_elems(plan.layout, plan.nsys, plan.nrow, 1),
                             dtype=torch.int64, device=d.device)
        solve_noncyclic_save_out(plan, a.detach(), b.detach(), c, d, cw, x, factors, pivots)
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


class BlockTridiagonalSolveF32R(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan: BtdmaGpuPlan, a, b, c, d, sweeps: int = 2):
        if plan.npart != 1:
            raise ValueError("the differentiable solve needs a plan with npart == 1")
        a, b, c = (t.detach().contiguous() for t in (a, b, c))
        x = torch.empty_like(d)
        work = torch.empty(f32r_work_elems(plan), dtype=torch.float32, device=d.device)
        solve_noncyclic_f32r(plan, a, b, c, d.detach().contiguous(), x, work, sweeps)
        ctx.plan, ctx.sweeps = plan, sweeps
        ctx.save_for_backward(a, b, c, x)
        return x

    @staticmethod
    def backward(ctx, w):
        a, b, c, x = ctx.saved_tensors
        lam = torch.empty_like(x)
        work = torch.empty(f32r_work_elems(ctx.plan), dtype=torch.float32, device=x.device)
        solve_transposed_f32r(ctx.plan, a, b, c, w.contiguous(), lam, work, ctx.sweeps)
        da, db, dc = torch.empty_like(a), torch.empty_like(a), torch.empty_like(a)
        outer_grads_f32(ctx.plan, lam, x, da, db, dc)
        return None, da, db, dc, lam, None


def block_tridiagonal_solve_f32r(plan: BtdmaGpuPlan, a, b, c, d, sweeps: int = 2):
    """x = T^{-1} d for FP32-stored blocks (FP64 d and x), differentiable in a, b, c, d."""
    return BlockTridiagonalSolveF32R.apply(plan, a, b, c, d, sweeps)
