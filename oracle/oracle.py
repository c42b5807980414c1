"""ctypes wrapper of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results against.  The product (pdcs_amd) never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

c_i64 = C.c_int64
c_dp = C.POINTER(C.c_double)


class OracleProblem(C.Structure):
    _fields_ = [("n", c_i64), ("m", c_i64), ("nb", c_i64),
                ("rowptr", C.POINTER(C.c_int32)), ("colidx", C.POINTER(C.c_int32)),
                ("vals", c_dp), ("c", c_dp), ("h", c_dp), ("bl", c_dp), ("bu", c_dp),
                ("mzero", c_i64), ("mnonneg", c_i64), ("nsoc", c_i64),
                ("soc", C.POINTER(c_i64)), ("nexp", c_i64), ("ndexp", c_i64)]


class OracleParams(C.Structure):
    _fields_ = [("max_outer_iter", c_i64), ("max_inner_iter", c_i64),
                ("abs_tol", C.c_double), ("rel_tol", C.c_double), ("time_limit", C.c_double),
                ("check_terminate_freq", c_i64), ("duality_gap_restart_freq", c_i64),
                ("kkt_restart_freq", c_i64), ("use_kkt_restart", C.c_int32),
                ("use_duality_gap_restart", C.c_int32), ("use_preconditioner", C.c_int32),
                ("ruiz_iters", C.c_int32), ("pock_chambolle", C.c_int32),
                ("warm_start", C.c_int32), ("verbose", C.c_int32), ("bisect_intent", C.c_int32),
                ("code_primal_denominator", C.c_int32), ("pessimistic_at", c_i64)]


class OracleResult(C.Structure):
    _fields_ = [("exit_code", C.c_int32), ("sequence", C.c_int32), ("iter", c_i64),
                ("restarts", c_i64), ("pobj", C.c_double), ("dobj", C.c_double),
                ("rel_primal_res", C.c_double), ("rel_dual_res", C.c_double),
                ("rel_gap", C.c_double), ("abs_primal_res", C.c_double),
                ("abs_dual_res", C.c_double), ("abs_gap", C.c_double),
                ("l2_rel_primal_res", C.c_double), ("l2_rel_dual_res", C.c_double),
                ("rel_primal_res_code", C.c_double), ("omega", C.c_double),
                ("time_solve", C.c_double), ("pessimistic_flag", C.c_int32),
                ("pessimistic_norm", C.c_double), ("pessimistic_iter", c_i64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.oracle_solve.restype = C.c_int
        L.oracle_time_attempts.restype = C.c_double
        L.oracle_default_params.restype = None
        _lib = L
    return _lib


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class _Held:
    """Keeps the numpy buffers behind an OracleProblem alive."""

    def __init__(self, P):
        self.rowptr = np.ascontiguousarray(P.rowptr, dtype=np.int32)
        self.colidx = np.ascontiguousarray(P.colidx, dtype=np.int32)
        self.vals = np.ascontiguousarray(P.vals, dtype=np.float64)
        self.c = np.ascontiguousarray(P.c, dtype=np.float64)
        self.h = np.ascontiguousarray(P.h, dtype=np.float64)
        self.bl = np.ascontiguousarray(P.bl, dtype=np.float64)
        self.bu = np.ascontiguousarray(P.bu, dtype=np.float64)
        self.soc = np.ascontiguousarray(np.asarray(P.soc, dtype=np.int64).reshape(-1))
        if self.soc.size == 0:
            self.soc = np.zeros(1, dtype=np.int64)
        self.s = OracleProblem(P.n, P.m, P.nb, _p(self.rowptr, C.c_int32), _p(self.colidx, C.c_int32),
                               _p(self.vals), _p(self.c), _p(self.h), _p(self.bl), _p(self.bu),
                               P.mzero, P.mnonneg, len(P.soc), _p(self.soc, c_i64), P.nexp, P.ndexp)


def default_params(**kw) -> OracleParams:
    p = OracleParams()
    lib().oracle_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def solve(P, x0=None, y0=None, **kw):
    """Oracle solve; returns (result dict, x, y) with x, y unscaled."""
    H = _Held(P)
    prm = default_params(**kw)
    x = np.zeros(P.n)
    y = np.zeros(P.m)
    res = OracleResult()
    xw = _p(np.ascontiguousarray(x0, dtype=np.float64)) if x0 is not None else None
    yw = _p(np.ascontiguousarray(y0, dtype=np.float64)) if y0 is not None else None
    if x0 is not None:
        prm.warm_start = 1
    rc = lib().oracle_solve(C.byref(H.s), C.byref(prm), xw, yw, _p(x), _p(y), C.byref(res))
    if rc != 0:
        raise RuntimeError(f"oracle_solve failed: {rc}")
    return res.as_dict(), x, y


def time_attempts(P, attempts: int):
    H = _Held(P)
    done = c_i64(0)
    t = lib().oracle_time_attempts(C.byref(H.s), c_i64(attempts), C.byref(done))
    return t, done.value


def kkt(P, x, y):
    H = _Held(P)
    res = OracleResult()
    lib().oracle_kkt(C.byref(H.s), _p(np.ascontiguousarray(x, dtype=np.float64)),
                     _p(np.ascontiguousarray(y, dtype=np.float64)), C.byref(res))
    return res.as_dict()


def tangent(P, x, y, dc=None, dh=None, dG=None, iters=200000, tol=1e-15):
    """Dual-number forward mode (oracle_tangent): returns (x, y, dx, dy, iterations)
    with (x, y) the refined solution and (dx, dy) its derivative along (dc, dh, dG)."""
    H = _Held(P)
    x = np.array(x, dtype=np.float64)
    y = np.array(y, dtype=np.float64)
    dx = np.zeros(P.n)
    dy = np.zeros(P.m)
    keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (dc, dh, dG)]
    done = c_i64(0)
    rc = lib().oracle_tangent(C.byref(H.s), _p(x), _p(y), *[None if a is None else _p(a) for a in keep],
                              c_i64(iters), C.c_double(tol), _p(dx), _p(dy), C.byref(done))
    if rc != 0:
        raise RuntimeError(f"oracle_tangent failed: {rc}")
    return x, y, dx, dy, done.value


def spmv(rowptr, colidx, vals, x):
    m = len(rowptr) - 1
    y = np.zeros(m)
    rp = np.ascontiguousarray(rowptr, dtype=np.int32)
    ci = np.ascontiguousarray(colidx, dtype=np.int32)
    va = np.ascontiguousarray(vals, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    lib().oracle_spmv(c_i64(m), _p(rp, C.c_int32), _p(ci, C.c_int32), _p(va), _p(xx), _p(y))
    return y


def proj_soc(v):
    v = np.array(v, dtype=np.float64)
    lib().oracle_proj_soc(_p(v), c_i64(len(v)))
    return v


def proj_soc_scaled(v, d, xi_ws=1.0, abs_tol=1e-16, rel_tol=1e-12):
    v = np.array(v, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    d2 = d * d
    ws = C.c_double(xi_ws)
    tmp = np.zeros(len(v))
    lib().oracle_proj_soc_scaled(_p(v), _p(d), _p(d2), c_i64(len(v)), C.byref(ws),
                                 C.c_double(abs_tol), C.c_double(rel_tol), _p(tmp))
    return v


def proj_exp(v):
    v = np.array(v, dtype=np.float64)
    lib().oracle_proj_exp(_p(v))
    return v


def proj_dual_exp(v):
    v = np.array(v, dtype=np.float64)
    lib().oracle_proj_dual_exp(_p(v))
    return v


def scale(P, ruiz_iters=10, pock=1):
    H = _Held(P)
    Dr = np.zeros(P.n)
    Dl = np.zeros(P.m)
    vals = np.zeros(P.nnz)
    lib().oracle_scale(C.byref(H.s), C.c_int(ruiz_iters), C.c_int(pock), _p(Dr), _p(Dl), _p(vals))
    return Dr, Dl, vals
