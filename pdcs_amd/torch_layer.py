"""torch.autograd bindings of the solve and its adjoint.

``ConicSolve.apply(c, h, vals, problem)`` returns the unscaled primal solution
x*(c, h, G) of ``min c'x  s.t.  G x - h in K, bl <= x[:nb] <= bu`` for the
sparsity pattern, cones and bounds of ``problem`` (a ConicProblem); the
backward pass returns dL/dc, dL/dh and dL/dG (values on the CSR pattern) from
the library's adjoint (pdcs_adjoint / pdcs_batch_adjoint).  The solve and the
derivative both run on the GPU through libpdcs_hip.so; tensors are moved to
host memory at the boundary (the C ABI takes host arrays).
"""
from __future__ import annotations

import copy

import numpy as np
import torch

from .solver import Batch, Problem


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy()


class ConicSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, h, vals, problem, solve_kw=None):
        P = copy.copy(problem)
        P.c, P.h, P.vals = _np(c), _np(h), _np(vals)
        prob = Problem.from_instance(P)
        res = prob.solve(**(solve_kw or {}))
        if res["exit_code"] != 0:
            raise RuntimeError(f"solve did not converge: exit code {res['exit_code']}")
        x, _, _ = prob.solution()
        ctx.prob = prob
        ctx.dev = c.device
        return torch.as_tensor(x, dtype=torch.float64, device=c.device)

    @staticmethod
    def backward(ctx, gx):
        dc, dh, dG = ctx.prob.adjoint(_np(gx))
        ctx.prob.close()
        t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=ctx.dev)
        return t(dc), t(dh), t(dG), None, None


class BatchConicSolve(torch.autograd.Function):
    """Batched form: lists of per-instance tensors in, a list of solutions out
    (one persistent launch for the solves, one for the adjoints)."""

    @staticmethod
    def forward(ctx, problems, solve_kw, *flat):
        k = len(problems)
        Ps = []
        for i, P0 in enumerate(problems):
            P = copy.copy(P0)
            P.c, P.h, P.vals = _np(flat[3 * i]), _np(flat[3 * i + 1]), _np(flat[3 * i + 2])
            Ps.append(P)
        B = Batch(Ps)
        res, _ = B.solve(**(solve_kw or {}))
        bad = [i for i, r in enumerate(res) if r["exit_code"] != 0]
        if bad:
            raise RuntimeError(f"instances {bad[:8]} did not converge")
        ctx.batch = B
        ctx.dev = flat[0].device
        return tuple(torch.as_tensor(x, dtype=torch.float64, device=ctx.dev) for x, _ in B.solutions()[:k])

    @staticmethod
    def backward(ctx, *gxs):
        grads = ctx.batch.adjoint([_np(g) for g in gxs])
        ctx.batch.close()
        out = [None, None]
        for g in grads:
            if g is None:
                raise RuntimeError("singular KKT Jacobian at a solution")
            out += [torch.as_tensor(a, dtype=torch.float64, device=ctx.dev) for a in g]
        return tuple(out)
