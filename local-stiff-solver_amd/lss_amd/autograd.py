"""Differentiable hybrid solve: x = A(vals)^{-1} b as a torch.autograd.Function.

The reference solver has no adjoint.  Here the backward pass reuses the
factorized preconditioner of the forward solve: for an upstream gradient
g = dL/dx it solves A^T lam = g (the operator is symmetric, so with the same
handle) and returns

    dL/db    = lam
    dL/dvals = -lam_j x_i   for the stored entry (j, i)  (CSC column j, row i;
                            column j is operator row j, GPUHybrid.cu:401-433)

computed by the HIP kernel behind lss_adjoint_pattern_grad.  Per-element
gradients follow with `element_grad`, the chain rule through the sparse
Jacobian d vals / d coef (lss_adjoint_chain).

All tensors live on the GPU: colptr/rowidx int32, vals/b float64, contiguous.
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import LssError, load
from .solver import Solver, require_device

_P = ctypes.c_void_p


def _dp(t: torch.Tensor):
    return _P(t.data_ptr())


def _check_free(rc: int):
    if rc != 0:
        raise LssError(rc, load().lss_last_error(None).decode())


def _cur_stream():
    return _P(int(torch.cuda.current_stream().cuda_stream))


def pattern_grad(colptr: torch.Tensor, rowidx: torch.Tensor, lam: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """-lam_j x_i on every stored entry (CSC column j, row rowidx[e])."""
    require_device(colptr, torch.int32, "colptr")
    n = int(colptr.numel()) - 1
    require_device(rowidx, torch.int32, "rowidx")
    require_device(lam, torch.float64, "lam", numel=n)
    require_device(x, torch.float64, "x", numel=n)
    grad = torch.empty(rowidx.numel(), dtype=torch.float64, device=x.device)
    _check_free(load().lss_adjoint_pattern_grad(n, int(rowidx.numel()), _dp(colptr), _dp(rowidx), _dp(lam), _dp(x),
                                                _dp(grad), _cur_stream()))
    return grad


def element_grad(seg_ptr: torch.Tensor, slot: torch.Tensor, coef: torch.Tensor, grad_vals: torch.Tensor) -> torch.Tensor:
    """out[e] = sum_k coef[k] grad_vals[slot[k]] over element e's Jacobian entries."""
    require_device(seg_ptr, torch.int64, "seg_ptr")
    nseg = int(seg_ptr.numel()) - 1
    require_device(slot, torch.int32, "slot")
    require_device(coef, torch.float64, "coef", numel=slot.numel())
    grad_vals = grad_vals.contiguous()
    require_device(grad_vals, torch.float64, "grad_vals")
    out = torch.empty(nseg, dtype=torch.float64, device=grad_vals.device)
    _check_free(load().lss_adjoint_chain(nseg, _dp(seg_ptr), _dp(slot), _dp(coef), _dp(grad_vals),
                                         _dp(out), _cur_stream()))
    return out


class HybridSolve(torch.autograd.Function):
    """x = solve(A(vals), b) with the GPU hybrid solver; differentiable in vals and b.

    forward(vals, b, colptr, rowidx, solver): factorizes A with `solver` (a
    lss_amd.solver.Solver) and solves from x0 = 0.  The handle must not be
    refactorized before backward runs (backward reuses its factorization).
    """

    @staticmethod
    def forward(ctx, vals, b, colptr, rowidx, solver: Solver):
        n = int(b.numel())
        stream = torch.cuda.current_stream()
        solver.factorize_device(n, colptr, rowidx, vals.detach().contiguous(), stream=stream)
        b = b.detach().contiguous()
        x = torch.zeros_like(b)
        solver.solve_device(b, x, stream=stream)
        ctx.solver = solver
        ctx.forward_info = solver.get_info()
        ctx.save_for_backward(x, colptr, rowidx)
        return x

    @staticmethod
    def backward(ctx, gx):
        x, colptr, rowidx = ctx.saved_tensors
        lam = torch.zeros_like(x)
        ctx.solver.solve_device(gx.contiguous(), lam, stream=torch.cuda.current_stream())
        ctx.adjoint_info = ctx.solver.get_info()
        grad_vals = pattern_grad(colptr, rowidx, lam, x) if ctx.needs_input_grad[0] else None
        grad_b = lam if ctx.needs_input_grad[1] else None
        return grad_vals, grad_b, None, None, None


def hybrid_solve(vals: torch.Tensor, b: torch.Tensor, colptr: torch.Tensor, rowidx: torch.Tensor,
                 solver: Solver | None = None, config: dict | None = None) -> torch.Tensor:
    """Differentiable x = A^{-1} b; creates a solver from `config` when none is given."""
    if solver is None:
        solver = Solver.create(config or {"solver": "GPUHybrid"})
    return HybridSolve.apply(vals, b, colptr, rowidx, solver)
