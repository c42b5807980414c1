"""GPU: the MathOptInterface-style front-end end to end (pdcs_amd/model.py,
the counterpart of julia/src/MOI_wrapper.jl and of the reference's
PDCS_GPU.Optimizer, src/pdcs_gpu/MOI_wrapper/MOI_wrapper.jl:382-873), on
models whose optimum and duals are known in closed form."""
import math

import numpy as np
import pytest

from pdcs_amd.model import Model

pytestmark = pytest.mark.gpu

ACC = dict(rel_tol=1e-9, abs_tol=1e-10, time_limit=60.0)
LOOSE = dict(eps_primal_infeasible_low_acc=1e-3, eps_dual_infeasible_low_acc=1e-3,
             eps_primal_infeasible_high_acc=1e-6, eps_dual_infeasible_high_acc=1e-6,
             time_limit=60.0, check_terminate_freq=200)


def _lp():
    # max x + y + 1  s.t.  x + 2y <= 4,  3x + y <= 6,  x, y >= 0
    m = Model()
    m.add_variables(2, lower=0.0)
    k1 = m.add_constraint([[-1.0, -2.0]], [4.0], "nonnegatives")
    k2 = m.add_constraint([[-3.0, -1.0]], [6.0], "nonnegatives")
    m.set_objective([1.0, 1.0], constant=1.0, sense="max")
    return m, k1, k2


def test_lp_maximize_with_duals(ctx):
    m, k1, k2 = _lp()
    m.optimize(ctx, **ACC)
    print("lp", m.termination_status(), m.objective_value(), m.value(), m.iterations(), flush=True)
    assert (m.termination_status(), m.primal_status(), m.dual_status()) == ("OPTIMAL", "FEASIBLE_POINT",
                                                                             "FEASIBLE_POINT")
    assert abs(m.objective_value() - 3.8) < 1e-6 and abs(m.dual_objective_value() - 3.8) < 1e-6
    assert np.allclose(m.value(), [1.6, 1.2], atol=1e-6)
    # a0 + A'y = 0 with y >= 0 (MOI's dual of a maximization): y = (0.4, 0.2)
    assert np.allclose(m.constraint_dual(k1), [0.4], atol=1e-6)
    assert np.allclose(m.constraint_dual(k2), [0.2], atol=1e-6)
    assert np.allclose(m.constraint_primal(k1), [0.0], atol=1e-6)
    assert abs(m.bound_dual(0, "lower")) < 1e-6 and m.iterations() > 0 and m.solve_time() > 0
    # warm start at that solution: same optimum
    m.set_start(x=m.value(), duals={k1: m.constraint_dual(k1), k2: m.constraint_dual(k2)})
    m.optimize(ctx, **ACC)
    assert m.termination_status() == "OPTIMAL" and abs(m.objective_value() - 3.8) < 1e-6


def test_soc_distance_to_box(ctx):
    # min t  s.t.  (t, x1 - 3, x2 - 4) in SOC,  x in [-1, 1]^2  ->  t = |(3, 4) - (1, 1)| = sqrt(13)
    m = Model()
    m.add_variables(2, lower=-1.0, upper=1.0)
    t = m.add_variables(1)[0]
    k = m.add_constraint([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], [0.0, -3.0, -4.0], "soc")
    m.set_objective({t: 1.0})
    m.optimize(ctx, **ACC)
    assert m.termination_status() == "OPTIMAL"
    assert abs(m.objective_value() - math.sqrt(13.0)) < 1e-6
    assert np.allclose(m.value()[:2], [1.0, 1.0], atol=1e-6)
    s = m.constraint_primal(k)
    assert s[0] >= np.linalg.norm(s[1:]) - 1e-6
    # upper-bound duals carry the reduced cost (<= 0) at the active bounds
    assert m.bound_dual(0, "upper") < -1e-3 and m.bound_dual(1, "upper") < -1e-3


def test_exp_and_rotated_soc(ctx):
    # min z  s.t.  (x, y, z) in K_exp = {y exp(x / y) <= z},  x = 1, y = 1  ->  z = e
    m = Model()
    m.add_variables(3)
    m.add_constraint([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], [-1.0, -1.0], "zeros")
    m.add_constraint(np.eye(3), np.zeros(3), "exp")
    m.set_objective([0.0, 0.0, 1.0])
    m.optimize(ctx, **ACC)
    assert m.termination_status() == "OPTIMAL" and abs(m.objective_value() - math.e) < 1e-6
    # min u + w  s.t.  (u, w, z) in RSOC = {2 u w >= z^2},  z = 2  ->  u = w = sqrt 2
    r = Model()
    r.add_variables(2)
    z = r.add_variables(1, lower=2.0, upper=2.0)[0]
    r.add_constraint(np.eye(3), np.zeros(3), "rsoc")
    r.set_objective({0: 1.0, 1: 1.0})
    r.optimize(ctx, **ACC)
    assert r.termination_status() == "OPTIMAL"
    assert abs(r.objective_value() - 2.0 * math.sqrt(2.0)) < 1e-6 and abs(r.value(z) - 2.0) < 1e-9


def test_infeasible_model_statuses(ctx):
    # x - 1 >= 0 and -x >= 0
    m = Model()
    m.add_variables(1)
    m.add_constraint([[1.0]], [-1.0], "nonnegatives")
    m.add_constraint([[-1.0]], [0.0], "nonnegatives")
    m.set_objective([0.0])
    m.optimize(ctx, **LOOSE)
    assert (m.termination_status(), m.primal_status(), m.dual_status()) == (
        "INFEASIBLE", "NO_SOLUTION", "INFEASIBILITY_CERTIFICATE"), m.raw_status()
