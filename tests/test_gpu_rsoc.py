"""Rotated second-order cones {(u, w, z): 2uw >= ||z||^2, u, w >= 0} (A8).

The library maps each rotated block onto an SOC with the orthogonal rotation
(u, w) -> ((u + w)/sqrt2, (u - w)/sqrt2) of its first two rows.  The oracle has
no rotated cones, so the reference objective comes from the oracle on the
rotated SOC formulation built here in numpy (the rotation identity is the
pin); the returned solution is checked in the caller's coordinates against the
rotated cone itself (membership of y and of Gx - h)."""
import numpy as np
import pytest

from oracle import oracle
from pdcs_amd import instances
from pdcs_amd.solver import Problem

pytestmark = pytest.mark.gpu

RS = np.sqrt(0.5)


def _rsoc_interior(rng, dims):
    out = []
    for q in dims:
        z = rng.standard_normal(q - 2)
        u = rng.uniform(0.5, 1.5)
        w = z @ z / (2 * u) + rng.uniform(0.1, 1.0)
        out += [u, w] + list(z)
    return np.array(out)


def _soc_interior(rng, dims):
    out = []
    for q in dims:
        z = rng.standard_normal(q - 1)
        out += [np.linalg.norm(z) + rng.uniform(0.1, 1.0)] + list(z)
    return np.array(out)


def _program(seed):
    rng = np.random.default_rng(seed)
    n, soc, rsoc = 120, [10] * 10, [6] * 20 + [40]
    m = sum(soc) + sum(rsoc)
    A = rng.standard_normal((m, n)) * (rng.uniform(size=(m, n)) < 0.1)
    x0 = rng.uniform(-1, 1, n)
    s0 = np.concatenate([_soc_interior(rng, soc), _rsoc_interior(rng, rsoc)])
    y0 = np.concatenate([_soc_interior(rng, soc), _rsoc_interior(rng, rsoc)])
    h = A @ x0 - s0
    c = A.T @ y0 + rng.normal(0, 0.1, n)
    return A, h, c, soc, rsoc


def _instance(A, h, c, soc, rsoc, name):
    m, n = A.shape
    rows, cols = np.nonzero(A)
    rowptr, colidx, vals = instances.csr_from_coo(m, rows, cols, A[rows, cols])
    return instances.ConicProblem(n=n, m=m, nb=n, c=c, rowptr=rowptr, colidx=colidx, vals=vals, h=h,
                                  bl=np.full(n, -10.0), bu=np.full(n, 10.0), soc=list(soc), rsoc=list(rsoc),
                                  name=name)


def _rotate(v, starts):
    v = v.copy()
    for r in starts:
        a, b = v[r], v[r + 1]
        v[r], v[r + 1] = RS * (a + b), RS * (a - b)
    return v


def _in_rsoc(v, starts, dims, tol):
    for r, q in zip(starts, dims):
        u, w, z = v[r], v[r + 1], v[r + 2:r + q]
        scale = 1 + abs(u) + abs(w) + np.linalg.norm(z)
        assert u >= -tol * scale and w >= -tol * scale, (r, u, w)
        assert 2 * u * w - z @ z >= -tol * scale ** 2, (r, 2 * u * w, z @ z)


@pytest.mark.parametrize("precision", [0, 1])
def test_rotated_soc_program(ctx, precision):
    A, h, c, soc, rsoc = _program(31)
    starts = list(np.cumsum([sum(soc)] + rsoc[:-1]))
    P = _instance(A, h, c, soc, rsoc, "rsoc")
    # the same program with the rotated rows as SOCs, for the oracle
    Ar = A.copy()
    for r in starts:
        Ar[[r, r + 1]] = RS * np.array([A[r] + A[r + 1], A[r] - A[r + 1]])
    Pr = _instance(Ar, _rotate(h, starts), c, list(soc) + list(rsoc), [], "rsoc_as_soc")
    ref, _, _ = oracle.solve(Pr, rel_tol=1e-8, abs_tol=1e-12)
    assert ref["exit_code"] == 0
    prob = Problem.from_instance(P, ctx)
    res = prob.solve(precision=precision, time_limit=100.0)
    assert res["exit_code"] == 0, res
    for k in ("pobj", "dobj"):
        assert abs(res[k] - ref[k]) <= 2e-6 * (1 + abs(ref[k])), (k, res[k], ref[k])
    x, y, _ = prob.solution()
    # caller's coordinates: y in the (self-dual) rotated cones, Gx - h in them too
    _in_rsoc(y, starts, rsoc, 1e-6)
    _in_rsoc(A @ x - h, starts, rsoc, 1e-5)
    # and the rotated solution is a 1e-6 solution of the SOC formulation
    chk = oracle.kkt(Pr, x, _rotate(y, starts))
    assert max(chk["rel_primal_res"], chk["rel_dual_res"], chk["rel_gap"]) < 2e-6, chk
    assert abs(c @ x - ref["pobj"]) <= 2e-6 * (1 + abs(ref["pobj"]))
