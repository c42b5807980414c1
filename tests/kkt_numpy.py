"""Independent numpy evaluator of the README convergence criteria (README.md:75-88).

Written from the README formulas, not from the oracle: the primal residual is
(Gx - h) - proj_K(Gx - h), the dual residual lambda - proj_Lambda(lambda) with
lambda = c - G'y, the gap |c'x - (h'y + l'lambda+ + u'lambda-)|.  SOC
projections use the closed form of the Euclidean projection onto
{(t, z): ||z|| <= t}.
"""
import numpy as np


def proj_soc(v):
    t, z = v[0], v[1:]
    nz = np.linalg.norm(z)
    if nz <= t:
        return v.copy()
    if nz <= -t:
        return np.zeros_like(v)
    a = (t + nz) / 2.0
    return np.concatenate([[a], z * (a / nz)])


def proj_K(P, r):
    out = r.copy()
    out[:P.mzero] = 0.0
    out[P.mzero:P.mzero + P.mnonneg] = np.maximum(out[P.mzero:P.mzero + P.mnonneg], 0.0)
    off = P.mzero + P.mnonneg
    for q in P.soc:
        out[off:off + q] = proj_soc(r[off:off + q])
        off += q
    assert P.nexp == 0 and P.ndexp == 0
    return out


def readme_residuals(P, x, y):
    Gx = P.matvec(x)
    r = Gx - P.h
    pr = proj_K(P, r)
    primal = np.max(np.abs(r - pr)) / (1 + max(np.max(np.abs(P.h)), np.max(np.abs(Gx)), np.max(np.abs(pr))))
    Gty = P.rmatvec(y)
    lam = P.c - Gty
    lo = np.where(P.bu == np.inf, 0.0, -np.inf)   # Lambda_1 = [lo, hi] per box variable
    hi = np.where(P.bl == -np.inf, 0.0, np.inf)
    lamp = np.clip(lam[:P.nb], lo, hi)
    dres = np.concatenate([lam[:P.nb] - lamp, lam[P.nb:]])
    dual = np.max(np.abs(dres)) / (1 + max(np.max(np.abs(P.c)), np.max(np.abs(Gty))))
    l = np.where(np.isinf(P.bl), 0.0, P.bl)
    u = np.where(np.isinf(P.bu), 0.0, P.bu)
    pobj = P.c @ x
    dobj = P.h @ y + l @ np.maximum(lam[:P.nb], 0) + u @ np.minimum(lam[:P.nb], 0)
    gap = abs(pobj - dobj) / (1 + max(abs(pobj), abs(dobj)))
    return {"primal": primal, "dual": dual, "gap": gap, "pobj": pobj, "dobj": dobj}


def dual_cone_violation(P, y):
    """distance of y from K* (K self-dual for zero^*=free, R+, SOC)"""
    v = 0.0
    off = P.mzero
    v = max(v, np.max(np.maximum(-y[off:off + P.mnonneg], 0.0), initial=0.0))
    off += P.mnonneg
    for q in P.soc:
        b = y[off:off + q]
        v = max(v, np.linalg.norm(b - proj_soc(b)))
        off += q
    return v
