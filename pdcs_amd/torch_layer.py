"""torch.autograd bindings of the solve and its adjoint.

``ConicSolve.apply(c, h, vals, problem)`` returns the unscaled primal solution
x*(c, h, G) of ``min c'x  s.t.  G x - h in K, bl <= x[:nb] <= bu`` for the
sparsity pattern, cones and bounds of ``problem`` (a ConicProblem); the
backward pass returns dL/dc, dL/dh and dL/dG (values on the CSR pattern) from
the library's adjoint (pdcs_adjoint / pdcs_batch_adjoint).  The solve and the
derivative both run on the GPU through libpdcs_hip.so; tensors are moved to
host memory at the boundary (the C ABI takes host arrays).

``ConicLayer(problem)`` keeps one problem resident on the GPU and re-solves it
with each call's (c, h, vals) on the fixed sparsity pattern (the cvxpy-style
re-solve of the reference's Python bridge, update_G! / update_solver!,
src/pdcs_gpu/cvxpy_wrapper/data_updating.jl:4-75): GPU tensors go in through
pdcs_problem_update_device, and a backward pass whose forward data has since
been replaced re-solves that data first (a re-solve after an update equals
the first solve bit for bit).
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


class _ResidentBatchSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, layer, *flat):
        layer._load(flat)
        xs = layer._solve()
        ctx.layer = layer
        ctx.gen = layer._gen
        ctx.data = tuple(t.detach() for t in flat)
        ctx.dev = flat[0].device
        return tuple(torch.as_tensor(x, dtype=torch.float64, device=ctx.dev) for x in xs)

    @staticmethod
    def backward(ctx, *gxs):
        layer = ctx.layer
        if layer._gen != ctx.gen:            # a later call replaced the data: solve this call's again
            layer._load(ctx.data)
            layer._solve()
        grads = layer.batch.adjoint([_np(g) for g in gxs])
        out = [None]
        for g in grads:
            if g is None:
                raise RuntimeError("singular KKT Jacobian at a solution")
            out += [torch.as_tensor(a, dtype=torch.float64, device=ctx.dev) for a in g]
        return tuple(out)


class BatchConicLayer(torch.nn.Module):
    """Batched ConicLayer: one resident batch (pdcs_batch_*) re-solved per call
    with that call's per-instance (c, h, G values) through
    pdcs_batch_update_host; the forward takes c_0, h_0, vals_0, c_1, ... and
    returns one solution per instance."""

    def __init__(self, problems, ctx=None, **solve_kw):
        super().__init__()
        self.batch = Batch(problems, ctx)
        self.count = len(problems)
        self.solve_kw = solve_kw
        self._gen = 0

    def _load(self, flat):
        k = self.count
        self.batch.update(c=[_np(flat[3 * i]) for i in range(k)], h=[_np(flat[3 * i + 1]) for i in range(k)],
                          vals=[_np(flat[3 * i + 2]) for i in range(k)])
        self._gen += 1

    def _solve(self):
        res, _ = self.batch.solve(**self.solve_kw)
        bad = [i for i, r in enumerate(res) if r["exit_code"] != 0]
        if bad:
            raise RuntimeError(f"instances {bad[:8]} did not converge")
        return [x for x, _ in self.batch.solutions()]

    def forward(self, *flat):
        if len(flat) != 3 * self.count:
            raise ValueError(f"expected {3 * self.count} tensors (c, h, vals per instance), got {len(flat)}")
        return _ResidentBatchSolve.apply(self, *flat)

    def close(self):
        self.batch.close()


class _ResidentSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, h, vals, layer):
        layer._load(c, h, vals)
        x = layer._solve()
        ctx.layer = layer
        ctx.gen = layer._gen
        ctx.data = (c.detach(), h.detach(), vals.detach())
        ctx.dev = c.device
        return torch.as_tensor(x, dtype=torch.float64, device=c.device)

    @staticmethod
    def backward(ctx, gx):
        layer = ctx.layer
        if layer._gen != ctx.gen:            # a later call replaced the data: solve this call's again
            layer._load(*ctx.data)
            layer._solve()
        dc, dh, dG = layer.prob.adjoint(_np(gx))
        t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=ctx.dev)
        return t(dc), t(dh), t(dG), None


class ConicLayer(torch.nn.Module):
    """x*(c, h, G) on the pattern, cones and bounds of `problem`, with the
    problem kept resident between calls.  solve_kw: pdcs_params fields."""

    def __init__(self, problem, ctx=None, **solve_kw):
        super().__init__()
        self.prob = Problem.from_instance(problem, ctx)
        self.solve_kw = solve_kw
        self._gen = 0

    def _load(self, c, h, vals):
        if c.is_cuda and h.is_cuda and vals.is_cuda:
            d = lambda t: t.detach().to(torch.float64).contiguous()
            self.prob.update_tensors(c=d(c), vals=d(vals), h=d(h))
        else:
            self.prob.update(c=_np(c), vals=_np(vals), h=_np(h))
        self._gen += 1

    def _solve(self):
        res = self.prob.solve(**self.solve_kw)
        if res["exit_code"] != 0:
            raise RuntimeError(f"solve did not converge: exit code {res['exit_code']}")
        x, _, _ = self.prob.solution()
        return x

    def forward(self, c, h, vals):
        return _ResidentSolve.apply(c, h, vals, self)

    def close(self):
        self.prob.close()
