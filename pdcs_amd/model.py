"""Conic model front-end with MathOptInterface semantics.

The Python counterpart of the MathOptInterface optimizer in
julia/src/MOI_wrapper.jl, which mirrors the reference's ``PDCS_GPU.Optimizer``
(src/pdcs_gpu/MOI_wrapper/MOI_wrapper.jl:165-873).  A model is

    min or max  c'x + c0   s.t.   A_k x + b_k in K_k  (k = 1..K),   l <= x <= u,

each K_k one of "zeros", "nonnegatives", "soc", "rsoc", "exp", "dualexp"
(MOI.Zeros, Nonnegatives, SecondOrderCone, RotatedSecondOrderCone,
ExponentialCone, DualExponentialCone; exponential cones in MOI's (x, y, z)
order).  ``standard_form`` groups the rows by cone type in that order (the
order of the C ABI's cone descriptor; within a type, the order the
constraints were added) and returns  min c'x  s.t.  G x - h in K  with G = A,
h = -b; a maximization is solved as the minimization of -c'x and its
objective values are negated back.

Results follow MOI's conventions: ``constraint_dual`` is y on the rows of a
constraint (y in K*, a0 = A'y + reduced costs for a minimization, unchanged
for a maximization), ``constraint_primal`` the slack A x + b of its rows.
Status names are MOI's.  Differences from the reference wrapper (exit codes
2/3 map to INFEASIBLE and 4/5 to DUAL_INFEASIBLE; certificates only on the
side that has one; ConstraintPrimal as the row slack) are listed in
julia/src/MOI_wrapper.jl.
"""
from __future__ import annotations

import numpy as np

from . import instances
from .solver import Context, Problem, default_context, make_cones

CONES = ("zeros", "nonnegatives", "soc", "rsoc", "exp", "dualexp")


def termination_status(exit_code: int) -> str:
    """MOI.TerminationStatus of a library exit code (-1: no solve)."""
    if exit_code < 0:
        return "OPTIMIZE_NOT_CALLED"
    if exit_code == 0:
        return "OPTIMAL"
    if exit_code in (1, 7):
        return "ITERATION_LIMIT"
    if exit_code in (2, 3):
        return "INFEASIBLE"
    if exit_code in (4, 5):
        return "DUAL_INFEASIBLE"
    if exit_code == 6:
        return "TIME_LIMIT"
    return "NUMERICAL_ERROR"


def primal_status(exit_code: int) -> str:
    if exit_code < 0:
        return "NO_SOLUTION"
    if exit_code == 0:
        return "FEASIBLE_POINT"
    if exit_code in (4, 5):
        return "INFEASIBILITY_CERTIFICATE"
    if exit_code in (2, 3, 8, 9):
        return "NO_SOLUTION"
    return "UNKNOWN_RESULT_STATUS"


def dual_status(exit_code: int) -> str:
    if exit_code < 0:
        return "NO_SOLUTION"
    if exit_code == 0:
        return "FEASIBLE_POINT"
    if exit_code in (2, 3):
        return "INFEASIBILITY_CERTIFICATE"
    if exit_code in (4, 5, 8, 9):
        return "NO_SOLUTION"
    return "UNKNOWN_RESULT_STATUS"


class Model:
    """A conic model; constraints and variables are referred to by the
    integer indices the add_* methods return."""

    def __init__(self):
        self.lower: list[float] = []
        self.upper: list[float] = []
        self._cons: list[tuple[np.ndarray, np.ndarray, str]] = []   # (A dense d x n_at_add, b, cone)
        self.c: dict[int, float] = {}
        self.c0 = 0.0
        self.sense = "min"
        self.x_start: dict[int, float] = {}
        self.dual_start: dict[int, np.ndarray] = {}
        self.options: dict = {}
        self._sol = None

    # ---- building -------------------------------------------------------
    @property
    def n(self) -> int:
        return len(self.lower)

    def add_variables(self, k: int, lower=-np.inf, upper=np.inf) -> np.ndarray:
        lo = np.broadcast_to(np.asarray(lower, float), (k,))
        up = np.broadcast_to(np.asarray(upper, float), (k,))
        if np.any(lo > up):
            raise ValueError("Not all elements of bl are less than or equal to the corresponding elements in bu")
        first = self.n
        self.lower += lo.tolist()
        self.upper += up.tolist()
        self._sol = None
        return np.arange(first, first + k)

    def set_bounds(self, j: int, lower=None, upper=None):
        lo = self.lower[j] if lower is None else float(lower)
        up = self.upper[j] if upper is None else float(upper)
        if lo > up:
            raise ValueError("Not all elements of bl are less than or equal to the corresponding elements in bu")
        self.lower[j], self.upper[j] = lo, up
        self._sol = None

    def add_constraint(self, A, b, cone: str) -> int:
        """A x + b in cone; A is a (d x k) array over the first k variables (k <= n)."""
        if cone not in CONES:
            raise ValueError(f"unsupported cone {cone!r}")
        A = np.atleast_2d(np.asarray(A, float))
        b = np.asarray(b, float).reshape(-1)
        d = A.shape[0]
        if b.shape != (d,) or A.shape[1] > self.n:
            raise ValueError(f"constraint shapes {A.shape}, {b.shape} do not fit {self.n} variables")
        if cone in ("exp", "dualexp") and d != 3:
            raise ValueError("exponential cones have dimension 3")
        if cone in ("soc", "rsoc") and d < (2 if cone == "soc" else 3):
            raise ValueError("The size of the SOC must be greater than 1")
        self._cons.append((A, b, cone))
        self._sol = None
        return len(self._cons) - 1

    def set_objective(self, coeffs, constant: float = 0.0, sense: str = "min"):
        """coeffs: {variable: coefficient} or a length-n array."""
        if sense not in ("min", "max"):
            raise ValueError("sense is 'min' or 'max'")
        if isinstance(coeffs, dict):
            self.c = {int(k): float(v) for k, v in coeffs.items()}
        else:
            self.c = {j: float(v) for j, v in enumerate(np.asarray(coeffs, float).reshape(-1)) if v != 0.0}
        self.c0 = float(constant)
        self.sense = sense
        self._sol = None

    def set_start(self, x=None, duals=None):
        """Warm start: primal values {variable: value} or an array, and row
        duals {constraint: array} (VariablePrimalStart / ConstraintDualStart)."""
        if x is not None:
            self.x_start = dict(x) if isinstance(x, dict) else dict(enumerate(np.asarray(x, float).reshape(-1)))
        if duals is not None:
            self.dual_start = {int(k): np.asarray(v, float).reshape(-1) for k, v in duals.items()}

    # ---- standard form --------------------------------------------------
    def standard_form(self):
        """(ConicProblem, rows) with rows[k] the row slice of constraint k."""
        n = self.n
        if not self._cons:
            raise ValueError("The constraint matrix G is empty, PDHG solver cannot be applied.")
        order = sorted(range(len(self._cons)), key=lambda k: (CONES.index(self._cons[k][2]), k))
        blocks, hs, rows = [], [], {}
        r = 0
        count = {cn: 0 for cn in CONES}
        soc, rsoc = [], []
        for k in order:
            A, b, cone = self._cons[k]
            d = A.shape[0]
            Af = np.zeros((d, n))
            Af[:, :A.shape[1]] = A
            blocks.append(Af)
            hs.append(-b)
            rows[k] = slice(r, r + d)
            r += d
            count[cone] += d
            if cone == "soc":
                soc.append(d)
            elif cone == "rsoc":
                rsoc.append(d)
        G = np.vstack(blocks)
        h = np.concatenate(hs)
        m = G.shape[0]
        ri, ci = np.nonzero(G)
        rowptr, colidx, vals = instances.csr_from_coo(m, ri, ci, G[ri, ci])
        sgn = -1.0 if self.sense == "max" else 1.0
        c = np.zeros(n)
        for j, v in self.c.items():
            c[j] += sgn * v
        P = instances.ConicProblem(n=n, m=m, nb=n, c=c, rowptr=rowptr, colidx=colidx, vals=vals, h=h,
                                   bl=np.asarray(self.lower, float), bu=np.asarray(self.upper, float),
                                   mzero=count["zeros"], mnonneg=count["nonnegatives"], soc=soc,
                                   nexp=count["exp"] // 3, ndexp=count["dualexp"] // 3, name="model")
        P.rsoc = rsoc
        return P, rows

    # ---- solving --------------------------------------------------------
    def optimize(self, ctx: Context | None = None, **params):
        """Solve on the GPU; params are pdcs_params fields (defaults of the
        reference wrapper: rel_tol = abs_tol = 1e-6, time_limit 1000 s)."""
        P, rows = self.standard_form()
        kw = dict(rel_tol=1e-6, abs_tol=1e-6, time_limit=1000.0)
        kw.update(self.options)
        kw.update(params)
        ctx = ctx or default_context()
        cones = make_cones(P.mzero, P.mnonneg, P.soc, P.rsoc, P.nexp, P.ndexp)
        prob = Problem(ctx, P.n, P.m, P.nb, P.c, P.rowptr, P.colidx, P.vals, P.h, P.bl, P.bu, cones)
        try:
            x0 = y0 = None
            if self.x_start or self.dual_start:
                x0 = np.zeros(P.n)
                for j, v in self.x_start.items():
                    x0[j] = v
                y0 = np.zeros(P.m)
                for k, v in self.dual_start.items():
                    y0[rows[k]] = v
            res = prob.solve(x0=x0, y0=y0, **kw)
            x, y, s = prob.solution()
        finally:
            prob.close()
        sgn = -1.0 if self.sense == "max" else 1.0
        self._sol = dict(res=res, x=x, y=y, reduced=s, slack=P.matvec(x) - P.h, rows=rows,
                         pobj=sgn * res["pobj"], dobj=sgn * res["dobj"])
        return self

    # ---- results --------------------------------------------------------
    def _need(self):
        if self._sol is None:
            raise RuntimeError("no solve of the current model")
        return self._sol

    @property
    def exit_code(self) -> int:
        return -1 if self._sol is None else int(self._sol["res"]["exit_code"])

    def termination_status(self) -> str:
        return termination_status(self.exit_code)

    def primal_status(self) -> str:
        return primal_status(self.exit_code)

    def dual_status(self) -> str:
        return dual_status(self.exit_code)

    def objective_value(self) -> float:
        S = self._need()
        ray = self.primal_status() == "INFEASIBILITY_CERTIFICATE"
        return S["pobj"] + (0.0 if ray else self.c0)

    def dual_objective_value(self) -> float:
        S = self._need()
        ray = self.dual_status() == "INFEASIBILITY_CERTIFICATE"
        return S["dobj"] + (0.0 if ray else self.c0)

    def value(self, j=None):
        x = self._need()["x"]
        return x.copy() if j is None else x[j]

    def constraint_primal(self, k: int) -> np.ndarray:
        S = self._need()
        return S["slack"][S["rows"][k]]

    def constraint_dual(self, k: int) -> np.ndarray:
        S = self._need()
        return S["y"][S["rows"][k]]

    def bound_dual(self, j: int, side: str) -> float:
        """Dual of the bound x_j >= l ("lower", >= 0) or x_j <= u ("upper", <= 0):
        the part of the reduced cost c - G'y on that side."""
        r = float(self._need()["reduced"][j])
        return max(r, 0.0) if side == "lower" else min(r, 0.0)

    def solve_time(self) -> float:
        return float(self._need()["res"]["time_solve"])

    def iterations(self) -> int:
        return int(self._need()["res"]["iter"])

    def raw_status(self) -> str:
        from .solver import EXIT_STATUS
        return EXIT_STATUS.get(self.exit_code, "unknown")
