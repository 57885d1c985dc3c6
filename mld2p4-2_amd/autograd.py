"""Differentiable AMG solve for the 7-point finite-volume diffusion operator.

    x = AmgSolve.apply(theta, b, idim)       # A(theta) x = b

A(theta) is the operator of ``problems.fv7_matrix`` with cell coefficients
k = exp(theta).  The forward pass assembles A on the host, builds the
multilevel hierarchy (the reference's setup path, host side) and solves on
the device with the preconditioned Krylov method.  The backward pass reuses
that hierarchy for the adjoint solve A^T lam = dL/dx (the hierarchy applied
with trans = 'T') and forms

    dL/db = lam,   dL/dtheta_c = -lam^T (dA/dtheta_c) x

with the device edge-product kernel (``fv7_grad_theta``).  There is no
reference counterpart: the reference computes no gradients.
"""
from __future__ import annotations

import numpy as np
import torch

from . import MldPrec, psb_krylov, psb_krylov_adjoint, fv7_grad_theta
from . import problems

DEFAULT_PREC = (("SMOOTHER_TYPE", "L1-JACOBI"), ("SMOOTHER_SWEEPS", 2), ("AGGR_THRESH", 0.0),
                ("AGGR_PROL", "SMOOTHED"), ("MIN_COARSE_SIZE", 200), ("COARSE_SOLVE", "BJAC"),
                ("COARSE_SUBSOLVE", "ILU"), ("COARSE_SWEEPS", 4))


class SolveOptions:
    """Krylov method, tolerance and preconditioner settings of AmgSolve."""

    def __init__(self, method="CG", eps=1e-10, itmax=1000, prec=DEFAULT_PREC, precision="double"):
        self.method, self.eps, self.itmax, self.prec, self.precision = method, eps, itmax, prec, precision
        self.last = {}  # iterations / errors of the last forward and adjoint solves


def build_prec(theta: torch.Tensor, idim: int, opts: SolveOptions):
    """Host assembly of A(theta) and the hierarchy; device upload."""
    kc = np.exp(theta.detach().double().cpu().numpy())
    irp, ja, val = problems.fv7_matrix(kc)
    p = MldPrec(device=theta.device.index or 0).init("ML")
    for k, v in opts.prec:
        p.set(k, v)
    p.set_precision(opts.precision)
    p.build((irp, ja, val))
    return p, (irp, ja, val)


class AmgSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, b, idim, opts=None):
        opts = opts or SolveOptions()
        if theta.dtype != torch.float64 or b.dtype != torch.float64 or not theta.is_cuda or not b.is_cuda:
            raise TypeError("AmgSolve: float64 CUDA tensors expected")
        prec, a = build_prec(theta, idim, opts)
        x, it, err = psb_krylov(opts.method, a, prec, b.contiguous(), eps=opts.eps, itmax=opts.itmax)
        if err > opts.eps:
            raise RuntimeError(f"AmgSolve: no convergence ({it} iterations, err {err:.3e})")
        opts.last = {"forward_iters": it, "forward_err": err}
        ctx.prec, ctx.idim, ctx.opts = prec, idim, opts
        ctx.save_for_backward(theta, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        theta, x = ctx.saved_tensors
        opts = ctx.opts
        lam, it, err = psb_krylov_adjoint(opts.method, ctx.prec, gx.contiguous(), eps=opts.eps, itmax=opts.itmax)
        if err > opts.eps:
            raise RuntimeError(f"AmgSolve adjoint: no convergence ({it} iterations, err {err:.3e})")
        opts.last.update({"adjoint_iters": it, "adjoint_err": err})
        gtheta = fv7_grad_theta(ctx.idim, theta.contiguous(), x, lam) if ctx.needs_input_grad[0] else None
        gb = lam if ctx.needs_input_grad[1] else None
        return gtheta, gb, None, None


def solve(theta, b, idim, opts=None):
    return AmgSolve.apply(theta, b, idim, opts)
