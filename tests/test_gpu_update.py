"""GPU: new data on a fixed sparsity pattern (pdcs_problem_update_host), the
re-solve path of the reference's Python bridge (update_G! / update_solver!,
src/pdcs_gpu/cvxpy_wrapper/data_updating.jl:4-75).

A problem created with data D1 and updated to D2 must solve exactly as one
created with D2 (the scaling, the scaled copies of both precisions and the
norms are all rebuilt); the schedule autotuner is off so both run the same
summation order.
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle
from pdcs_amd import _lib, instances
from pdcs_amd.instances import _soc_interior
from pdcs_amd.solver import Problem

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _same_pattern(P, seed):
    """New feasible data on the pattern of P: G values rescaled entrywise,
    h = G x0 - s0 and c = G' y0 + noise as in soc_program, a new box."""
    rng = np.random.default_rng(seed)
    Q = dataclasses.replace(P, vals=P.vals * rng.uniform(0.5, 1.5, P.vals.shape))
    x0 = rng.uniform(-4.0, 4.0, P.n)
    Q.h = Q.matvec(x0) - _soc_interior(rng, P.soc)
    Q.c = Q.rmatvec(_soc_interior(rng, P.soc)) + rng.normal(0.0, 0.1, P.n)
    Q.bl = np.full(P.n, -8.0)
    Q.bu = np.full(P.n, 8.0)
    return Q


def _fields(r):
    return (r["exit_code"], r["iter"], r["pobj"], r["dobj"], r["rel_primal_res"], r["rel_dual_res"], r["rel_gap"])


@pytest.mark.parametrize("precision", [0, 1])
def test_update_matches_fresh_problem(ctx, precision):
    P1 = instances.soc_program(300, 600, 0.02, 10, seed=31)
    P2 = _same_pattern(P1, 32)
    ref, _, _ = oracle.solve(P2, rel_tol=1e-8, abs_tol=1e-12)
    assert ref["exit_code"] == 0
    fresh = Problem.from_instance(P2, ctx)
    r_fresh = fresh.solve(precision=precision, autotune=0, time_limit=100.0)
    x_fresh, y_fresh, _ = fresh.solution()
    prob = Problem.from_instance(P1, ctx)
    r1 = prob.solve(precision=precision, autotune=0, time_limit=100.0)   # fills the scaled copies of D1
    assert r1["exit_code"] == 0
    prob.update(c=P2.c, vals=P2.vals, h=P2.h, bl=P2.bl, bu=P2.bu)
    r2 = prob.solve(precision=precision, autotune=0, time_limit=100.0)
    x2, y2, _ = prob.solution()
    print(f"prec {precision}: fresh {_fields(r_fresh)} updated {_fields(r2)}", flush=True)
    assert _fields(r2) == _fields(r_fresh)
    assert np.array_equal(x2, x_fresh) and np.array_equal(y2, y_fresh)
    assert r2["exit_code"] == 0
    for k in ("pobj", "dobj"):
        assert abs(r2[k] - ref[k]) <= 2e-6 * (1 + abs(ref[k])), (k, r2[k], ref[k])
    chk = oracle.kkt(P2, x2, y2)
    assert max(chk["rel_primal_res"], chk["rel_dual_res"], chk["rel_gap"]) < 1e-6, chk


def test_partial_update_and_rejections(ctx):
    P1 = instances.soc_program(200, 400, 0.03, 10, seed=41)
    P2 = _same_pattern(P1, 42)
    Pc = dataclasses.replace(P1, c=P2.c)          # only c changes
    fresh = Problem.from_instance(Pc, ctx)
    r_fresh = fresh.solve(autotune=0, time_limit=100.0)
    prob = Problem.from_instance(P1, ctx)
    r1 = prob.solve(autotune=0, time_limit=100.0)
    prob.update(c=P2.c)
    r2 = prob.solve(autotune=0, time_limit=100.0)
    assert _fields(r2) == _fields(r_fresh)
    # bl > bu is rejected before anything changes
    with pytest.raises(_lib.PdcsError):
        prob.update(bl=np.full(P1.n, 1.0), bu=np.full(P1.n, -1.0))
    r3 = prob.solve(autotune=0, time_limit=100.0)
    assert _fields(r3) == _fields(r2)
    # the solution of the old data is dropped by an update
    prob.update(c=P1.c)
    with pytest.raises(_lib.PdcsError):
        prob.solution()
    assert _fields(prob.solve(autotune=0, time_limit=100.0)) == _fields(r1)
    # rotated-SOC problems keep a rotated pattern: updates are refused
    R = instances.soc_program(60, 120, 0.1, 6, seed=43)
    R = dataclasses.replace(R, soc=R.soc[:-2], rsoc=[6, 6])
    pr = Problem.from_instance(R, ctx)
    with pytest.raises(_lib.PdcsError):
        pr.update(c=R.c)


@pytest.mark.parametrize("precision", [0, 1])
def test_warm_start_matches_oracle(ctx, precision):
    """pdcs_solve with params.warm_start: the unscaled (x0, y0) times (Dr, Dl)
    is the first iterate (rpdhg_alg_gpu_gen.jl:1350-1372), as in the oracle."""
    P = _same_pattern(instances.soc_program(300, 600, 0.02, 10, seed=31), 32)
    rng = np.random.default_rng(52)
    x0 = rng.uniform(-1.0, 1.0, P.n)
    y0 = _soc_interior(rng, P.soc) * 0.1
    ref, _, _ = oracle.solve(P, x0=x0, y0=y0)
    opt, _, _ = oracle.solve(P, rel_tol=1e-8, abs_tol=1e-12)
    prob = Problem.from_instance(P, ctx)
    r = prob.solve(precision=precision, x0=x0, y0=y0, time_limit=30.0)
    x, y, _ = prob.solution()
    print(f"prec {precision}: oracle iter {ref['iter']} gpu iter {r['iter']}", flush=True)
    assert r["exit_code"] == 0
    for k in ("pobj", "dobj"):
        assert abs(r[k] - opt[k]) <= 2e-6 * (1 + abs(opt[k])), (k, r[k], opt[k])
    chk = oracle.kkt(P, x, y)
    assert max(chk["rel_primal_res"], chk["rel_dual_res"], chk["rel_gap"]) < 1e-6, chk
    # a zero warm start is the cold solve
    cold = prob.solve(precision=precision, autotune=0, time_limit=30.0)
    zero = prob.solve(precision=precision, autotune=0, x0=np.zeros(P.n), y0=np.zeros(P.m), time_limit=30.0)
    assert _fields(zero) == _fields(cold)
