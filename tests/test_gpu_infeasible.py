"""GPU infeasibility certificates (exit codes 2-5; termination.jl:96-174,
rpdhg_alg_gpu_gen_scaling.jl:185-246).  With the reference's default
thresholds (1e-12 / 1e-16) the iterates must diverge to ~1e12 before a
certificate is declared, so these cases loosen them (eps_*_low_acc = 1e-3,
eps_*_high_acc = 1e-6): the detection logic, the ray quantities and the exit
paths are exercised in a test's time budget."""
import numpy as np
import pytest

from pdcs_amd import instances
from pdcs_amd.solver import Problem

pytestmark = pytest.mark.gpu

LOOSE = dict(eps_primal_infeasible_low_acc=1e-3, eps_dual_infeasible_low_acc=1e-3,
             eps_primal_infeasible_high_acc=1e-6, eps_dual_infeasible_high_acc=1e-6,
             time_limit=60.0, check_terminate_freq=200)


def _lp(G, h, c, mzero, mnonneg):
    G = np.asarray(G, dtype=float)
    m, n = G.shape
    rows, cols = np.nonzero(G)
    rowptr, colidx, vals = instances.csr_from_coo(m, rows, cols, G[rows, cols])
    return instances.ConicProblem(n=n, m=m, nb=0, c=np.asarray(c, float), rowptr=rowptr, colidx=colidx,
                                  vals=vals, h=np.asarray(h, float), bl=np.zeros(0), bu=np.zeros(0),
                                  mzero=mzero, mnonneg=mnonneg, name="lp")


def _random_infeasible(seed):
    """A x >= b with a planted Farkas certificate: rows 0..k-1 random, the last
    row is -(sum of the others) with a right side that contradicts their sum."""
    rng = np.random.default_rng(seed)
    n, k = 8, 12
    A = rng.standard_normal((k, n))
    b = rng.uniform(0.1, 1.0, k)
    G = np.vstack([A, -A.sum(axis=0)])
    h = np.concatenate([b, [-b.sum() + 1.0]])   # sum(A x) >= sum(b) and <= sum(b) - 1
    return _lp(G, h, rng.standard_normal(n), 0, k + 1)


def test_primal_infeasible_tiny(ctx):
    # x >= 1 and -x >= 0
    P = _lp([[1.0], [-1.0]], [1.0, 0.0], [0.0], 0, 2)
    res = Problem.from_instance(P, ctx).solve(**LOOSE)
    assert res["exit_code"] in (2, 3), res
    assert res["dobj"] > 0 and res["sequence"] in (0, 1)


def test_primal_infeasible_random(ctx):
    res = Problem.from_instance(_random_infeasible(5), ctx).solve(**LOOSE)
    assert res["exit_code"] in (2, 3), res
    assert res["dobj"] > 0


def test_dual_infeasible_unbounded(ctx):
    # min -x1 - x2 s.t. x1 >= 0, x2 >= 0, x1 - x2 = 0 (unbounded along (1, 1))
    P = _lp([[1.0, -1.0], [1.0, 0.0], [0.0, 1.0]], [0.0, 0.0, 0.0], [-1.0, -1.0], 1, 2)
    res = Problem.from_instance(P, ctx).solve(**LOOSE)
    assert res["exit_code"] in (4, 5), res
    assert res["pobj"] < 0


def test_feasible_not_flagged_with_loose_thresholds(ctx):
    P = instances.lp_box(60, 10, 30, seed=4)
    res = Problem.from_instance(P, ctx).solve(**LOOSE)
    assert res["exit_code"] == 0, res
