"""CPU: the MathOptInterface-style front-end's standard form and status maps
(pdcs_amd/model.py; the reference's MOI wrapper builds the same form,
src/pdcs_gpu/MOI_wrapper/MOI_wrapper.jl:382-712: rows grouped by cone type,
G = A, h = -b, a maximization negated)."""
import numpy as np
import pytest

from pdcs_amd import model as M


def _dense(P):
    G = np.zeros((P.m, P.n))
    for i in range(P.m):
        for k in range(P.rowptr[i], P.rowptr[i + 1]):
            G[i, P.colidx[k]] += P.vals[k]
    return G


def test_rows_grouped_by_cone_type_in_insertion_order():
    m = M.Model()
    x = m.add_variables(3, lower=[-1.0, 0.0, -np.inf], upper=[1.0, np.inf, 5.0])
    A_soc = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    k_soc = m.add_constraint(A_soc, [0.0, 1.0, 2.0], "soc")
    k_nn1 = m.add_constraint([[1.0, 1.0, 0.0]], [-1.0], "nonnegatives")
    k_z = m.add_constraint([[0.0, 2.0, -1.0], [1.0, 0.0, 0.0]], [3.0, 4.0], "zeros")
    k_exp = m.add_constraint(np.eye(3), [0.0, 1.0, 0.0], "exp")
    k_nn2 = m.add_constraint([[0.0, 0.0, 1.0]], [7.0], "nonnegatives")
    k_r = m.add_constraint(np.eye(3), [0.0, 0.0, 0.0], "rsoc")
    m.set_objective({0: 1.0, 2: -2.0}, constant=3.0, sense="max")
    P, rows = m.standard_form()
    assert (P.n, P.m) == (3, 2 + 2 + 3 + 3 + 3)
    assert (P.mzero, P.mnonneg, list(P.soc), list(P.rsoc), P.nexp, P.ndexp) == (2, 2, [3], [3], 1, 0)
    assert rows[k_z] == slice(0, 2) and rows[k_nn1] == slice(2, 3) and rows[k_nn2] == slice(3, 4)
    assert rows[k_soc] == slice(4, 7) and rows[k_r] == slice(7, 10) and rows[k_exp] == slice(10, 13)
    G = _dense(P)
    assert np.array_equal(G[0:2], [[0.0, 2.0, -1.0], [1.0, 0.0, 0.0]])
    assert np.array_equal(G[2:4], [[1.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    assert np.array_equal(G[4:7], A_soc) and np.array_equal(G[10:13], np.eye(3))
    assert np.array_equal(P.h, -np.array([3.0, 4.0, -1.0, 7.0, 0.0, 1.0, 2.0, 0, 0, 0, 0.0, 1.0, 0.0]))
    assert np.array_equal(P.c, [-1.0, 0.0, 2.0])            # max: negated
    assert P.nb == 3 and np.array_equal(P.bl, [-1.0, 0.0, -np.inf]) and np.array_equal(P.bu, [1.0, np.inf, 5.0])
    assert x.tolist() == [0, 1, 2]


def test_constraint_over_earlier_variables_is_padded():
    m = M.Model()
    m.add_variables(1)
    k0 = m.add_constraint([[2.0]], [1.0], "nonnegatives")
    m.add_variables(2)
    m.add_constraint([[1.0, 1.0, 1.0]], [0.0], "zeros")
    P, rows = m.standard_form()
    assert np.array_equal(_dense(P)[rows[k0]], [[2.0, 0.0, 0.0]])


def test_rejections():
    m = M.Model()
    with pytest.raises(ValueError):
        m.add_variables(2, lower=1.0, upper=0.0)
    m.add_variables(2)
    with pytest.raises(ValueError):
        m.add_constraint(np.eye(2), [0.0, 0.0], "psd")
    with pytest.raises(ValueError):
        m.add_constraint(np.eye(2), [0.0, 0.0], "exp")
    with pytest.raises(ValueError):
        m.add_constraint(np.ones((1, 3)), [0.0], "zeros")       # more columns than variables
    with pytest.raises(ValueError):
        m.standard_form()                                        # no rows
    with pytest.raises(RuntimeError):
        m.value()


def test_status_maps():
    # exit codes of the library (solver.EXIT_STATUS) -> MOI statuses
    expect = {-1: ("OPTIMIZE_NOT_CALLED", "NO_SOLUTION", "NO_SOLUTION"),
              0: ("OPTIMAL", "FEASIBLE_POINT", "FEASIBLE_POINT"),
              1: ("ITERATION_LIMIT", "UNKNOWN_RESULT_STATUS", "UNKNOWN_RESULT_STATUS"),
              2: ("INFEASIBLE", "NO_SOLUTION", "INFEASIBILITY_CERTIFICATE"),
              3: ("INFEASIBLE", "NO_SOLUTION", "INFEASIBILITY_CERTIFICATE"),
              4: ("DUAL_INFEASIBLE", "INFEASIBILITY_CERTIFICATE", "NO_SOLUTION"),
              5: ("DUAL_INFEASIBLE", "INFEASIBILITY_CERTIFICATE", "NO_SOLUTION"),
              6: ("TIME_LIMIT", "UNKNOWN_RESULT_STATUS", "UNKNOWN_RESULT_STATUS"),
              7: ("ITERATION_LIMIT", "UNKNOWN_RESULT_STATUS", "UNKNOWN_RESULT_STATUS"),
              8: ("NUMERICAL_ERROR", "NO_SOLUTION", "NO_SOLUTION"),
              9: ("NUMERICAL_ERROR", "NO_SOLUTION", "NO_SOLUTION")}
    for code, (t, p, d) in expect.items():
        assert (M.termination_status(code), M.primal_status(code), M.dual_status(code)) == (t, p, d), code
