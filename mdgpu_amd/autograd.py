"""torch.autograd wrapper of the reverse-mode C ABI (test and user
convenience; the kernels are in csrc/selinv.hip and csrc/kernels.hip).

``logdet_solve(vals, b, sym)`` takes the values of a symmetric positive
definite matrix's lower triangle (caller entry order of the analysis
``sym``, float64 CUDA tensor) and a right-hand side, factors on the GPU in
FP64 and returns ``(log det A, x = A^{-1} b)``.  Backward uses the selected
inverse (d log det A / d a = Z on the pattern, doubled off the diagonal) and
the solve adjoint (bbar = A^{-1} xbar, d/da = -sym(bbar x^T) on the
pattern)."""
from __future__ import annotations

import torch

from .numeric import Factorization


class LogdetSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, b, sym):
        F = Factorization(sym, "fp64").factorize(vals.detach())
        x = F.solve(b.detach())
        ld = torch.tensor(F.logdet(), dtype=torch.float64, device=vals.device)
        ctx.F = F
        ctx.save_for_backward(x)
        return ld, x

    @staticmethod
    def backward(ctx, g_ld, g_x):
        F = ctx.F
        (x,) = ctx.saved_tensors
        gvals = torch.zeros(F.sym.nnz_a, dtype=torch.float64, device=x.device)
        gb = None
        if g_ld is not None:
            gvals = gvals + g_ld * F.logdet_grad()
        if g_x is not None:
            bbar, ga = F.solve_adjoint(x, g_x)
            gvals = gvals + ga
            gb = bbar
        return gvals, gb, None


def logdet_solve(vals, b, sym):
    return LogdetSolve.apply(vals, b, sym)
