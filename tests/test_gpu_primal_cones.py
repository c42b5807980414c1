"""GPU: primal-side cones (pdcs_problem_create_host_x; the reference's soc_x,
rsoc_x, exp_x, dual_exp_x arguments of rpdhg_gpu_solve,
src/pdcs_gpu/rpdhg_alg_gpu_gen.jl:1234-1237, whose projections the reference
leaves undefined, def_rpdhg_gen.jl:153,161,299-313).

Pins: (1) the problem equals the one with each primal block written as the
rows I x_B - 0 in K_B after the caller's rows of that cone type, built here
independently: bit-identical solves, y on the caller's rows, s = c - G'y
carrying the block multipliers; (2) the oracle on that form; (3) closed-form
optima of one-block programs."""
import math

import numpy as np
import pytest

from oracle import oracle
from pdcs_amd import _lib, instances
from pdcs_amd.solver import Problem, make_cones, make_primal_cones, rpdhg_gpu_solve

pytestmark = pytest.mark.gpu

KW = dict(autotune=0, time_limit=60.0)


def _fields(r):
    return (r["exit_code"], r["iter"], r["pobj"], r["dobj"], r["rel_primal_res"], r["rel_dual_res"], r["rel_gap"])


def _soc_in(rng, d):
    t = rng.standard_normal(d - 1)
    return np.concatenate([[np.linalg.norm(t) + rng.uniform(0.1, 1.0)], t])


def _exp_in(rng):          # y exp(x / y) < z, y > 0
    x, y = rng.uniform(-1, 1), rng.uniform(0.5, 1.5)
    return np.array([x, y, y * math.exp(x / y) + rng.uniform(0.1, 1.0)])


def _dexp_in(rng):         # u < 0, -u exp(v / u) < e w
    u, v = -rng.uniform(0.5, 1.5), rng.uniform(-1, 1)
    return np.array([u, v, -u * math.exp(v / u) / math.e + rng.uniform(0.1, 1.0)])


XSOC, XRSOC, XEXP, XDEXP, NB, NFREE = [4, 3], [3], 1, 1, 10, 2


def _rsoc_in(rng, d):      # 2 u w > |z|^2, u, w > 0
    z = rng.standard_normal(d - 2)
    t = np.linalg.norm(z) + rng.uniform(0.5, 1.0)
    return np.concatenate([[t, t], z])
MZ, MN, SOC, URSOC, EXPG = 2, 5, [5, 4], [3], 1


def _program(seed):
    """Feasible and bounded: h = G x0 - s0, c = G'y0 + lam0 with x0, s0, y0,
    lam0 interior to their cones."""
    rng = np.random.default_rng(seed)
    n = NB + sum(XSOC) + sum(XRSOC) + 3 * (XEXP + XDEXP) + NFREE
    m = MZ + MN + sum(SOC) + sum(URSOC) + 3 * EXPG
    G = rng.standard_normal((m, n)) * (rng.uniform(size=(m, n)) < 0.4)
    x0 = np.concatenate([rng.uniform(-2, 2, NB)] + [_soc_in(rng, d) for d in XSOC] +
                        [np.abs(_soc_in(rng, d)) + 0.5 for d in XRSOC] +          # 2uw > |z|^2
                        [_exp_in(rng) for _ in range(XEXP)] + [_dexp_in(rng) for _ in range(XDEXP)] +
                        [rng.standard_normal(NFREE)])
    for j0, d in zip(np.cumsum([NB + sum(XSOC)] + XRSOC)[:-1], XRSOC):     # make 2 u w > |z|^2
        x0[j0] = x0[j0 + 1] = np.linalg.norm(x0[j0 + 2:j0 + d]) + 1.0
    s0 = np.concatenate([np.zeros(MZ), rng.uniform(0.1, 1, MN)] + [_soc_in(rng, d) for d in SOC] +
                        [_rsoc_in(rng, d) for d in URSOC] + [_exp_in(rng) for _ in range(EXPG)])
    y0 = np.concatenate([rng.standard_normal(MZ), rng.uniform(0.1, 1, MN)] + [_soc_in(rng, d) for d in SOC] +
                        [_rsoc_in(rng, d) for d in URSOC] + [_dexp_in(rng) for _ in range(EXPG)])
    lam = np.concatenate([rng.standard_normal(NB)] + [_soc_in(rng, d) for d in XSOC] +
                         [np.concatenate([[1.0, 1.0], 0.1 * rng.standard_normal(d - 2)]) for d in XRSOC] +
                         [_dexp_in(rng) for _ in range(XEXP)] + [_exp_in(rng) for _ in range(XDEXP)] +
                         [np.zeros(NFREE)])
    h = G @ x0 - s0
    c = G.T @ y0 + lam
    return n, m, G, h, c


def _csr(G):
    r, cidx = np.nonzero(G)
    return instances.csr_from_coo(G.shape[0], r, cidx, G[r, cidx])


def _lowered(n, m, G, h):
    """The extra rows written out: each group of the caller's rows, then the
    identity rows of the primal blocks of that cone type."""
    g1 = MZ + MN + sum(SOC)
    g2 = g1 + sum(URSOC)
    xs = NB
    xr = xs + sum(XSOC)
    xe = xr + sum(XRSOC)
    xd = xe + 3 * XEXP
    eye = lambda j0, k: np.eye(n)[j0:j0 + k]
    parts = [G[:g1], eye(xs, xr - xs), G[g1:g2], eye(xr, xe - xr), G[g2:], eye(xe, xd - xe), eye(xd, 3 * XDEXP)]
    hs = [h[:g1], np.zeros(xr - xs), h[g1:g2], np.zeros(xe - xr), h[g2:], np.zeros(xd - xe), np.zeros(3 * XDEXP)]
    user = np.concatenate([np.arange(g1), g1 + (xr - xs) + np.arange(g2 - g1),
                           g2 + (xr - xs) + (xe - xr) + np.arange(m - g2)])
    return np.vstack(parts), np.concatenate(hs), user


def test_primal_cones_equal_their_row_form(ctx):
    n, m, G, h, c = _program(3)
    bl, bu = np.full(NB, -5.0), np.full(NB, 5.0)
    rp, ci, va = _csr(G)
    xk = make_primal_cones(XSOC, XRSOC, XEXP, XDEXP)
    px = Problem(ctx, n, m, NB, c, rp, ci, va, h, bl, bu, make_cones(MZ, MN, SOC, URSOC, EXPG, 0), xk)
    GA, hA, user = _lowered(n, m, G, h)
    rpa, cia, vaa = _csr(GA)
    pa = Problem(ctx, n, GA.shape[0], NB, c, rpa, cia, vaa, hA, bl, bu,
                 make_cones(MZ, MN, SOC + XSOC, URSOC + XRSOC, EXPG + XEXP, XDEXP))
    for prec in (0, 1):
        rx = px.solve(precision=prec, **KW)
        ra = pa.solve(precision=prec, **KW)
        print(f"prec {prec}: {_fields(rx)}", flush=True)
        assert rx["exit_code"] == 0 and _fields(rx) == _fields(ra)
        x1, y1, s1 = px.solution()
        x2, y2, s2 = pa.solution()
        assert np.array_equal(x1, x2) and np.array_equal(y1, y2[user])
        extra = np.setdiff1d(np.arange(GA.shape[0]), user)
        cols = [int(np.nonzero(GA[r])[0][0]) for r in extra]
        assert np.allclose(s1[cols], s2[cols] + y2[extra], rtol=0, atol=1e-9 * (1 + np.abs(s1).max()))
        # pdcs_evaluate on the caller's (x, y) lifts y with the reduced costs
        ev = px.evaluate(x1, y1)
        assert max(ev["rel_primal_res"], ev["rel_dual_res"], ev["rel_gap"]) < 1e-6, ev
    # the oracle on the row form, its rotated-SOC blocks turned into SOCs by
    # the orthogonal map (u, w) -> ((u + w)/sqrt2, (u - w)/sqrt2) of their first two rows
    GR, hR = GA.copy(), hA.copy()
    rot = np.array([[1.0, 1.0], [1.0, -1.0]]) / math.sqrt(2.0)
    r0 = MZ + MN + sum(SOC) + sum(XSOC)
    for d in URSOC + XRSOC:
        GR[[r0, r0 + 1]] = rot @ GA[[r0, r0 + 1]]
        hR[[r0, r0 + 1]] = rot @ hA[[r0, r0 + 1]]
        r0 += d
    rpr, cir, var = _csr(GR)
    PA = instances.ConicProblem(n=n, m=GR.shape[0], nb=NB, c=c, rowptr=rpr, colidx=cir, vals=var, h=hR,
                                bl=bl, bu=bu, mzero=MZ, mnonneg=MN, soc=SOC + XSOC + URSOC + XRSOC, nexp=EXPG + XEXP,
                                ndexp=XDEXP, name="lowered")
    ref, _, _ = oracle.solve(PA, rel_tol=1e-9, abs_tol=1e-12)
    assert ref["exit_code"] == 0
    for k in ("pobj", "dobj"):
        assert abs(rx[k] - ref[k]) <= 2e-6 * (1 + abs(ref[k])), (k, rx[k], ref[k])
    with pytest.raises(_lib.PdcsError):
        px.update(c=c)
    with pytest.raises(_lib.PdcsError):
        px.adjoint(np.ones(n))
    # the same problem from GPU tensors (pdcs_problem_create_device_x)
    torch = pytest.importorskip("torch")
    dev = lambda a, dt=torch.float64: torch.as_tensor(np.asarray(a), dtype=dt).to("cuda").contiguous()
    pt = Problem.from_tensors(n, m, NB, dev(c), dev(rp, torch.int32), dev(ci, torch.int32), dev(va), dev(h), dev(bl),
                              dev(bu), make_cones(MZ, MN, SOC, URSOC, EXPG, 0), ctx, xcones=xk)
    rt = pt.solve(precision=1, **KW)
    assert _fields(rt) == _fields(rx)
    assert all(np.array_equal(a, b) for a, b in zip(pt.solution(), px.solution()))
    with pytest.raises(_lib.PdcsError):
        pt.update_tensors(h=dev(h))


def _one_block(G, h, c, **xc):
    n = len(c)
    return rpdhg_gpu_solve(n=n, m=len(h), nb=0, c=np.asarray(c, float), G=_csr(np.asarray(G, float)),
                           h=np.asarray(h, float), mGzero=len(h), mGnonnegative=0, socG=[], bl=np.zeros(0),
                           bu=np.zeros(0), rel_tol=1e-9, abs_tol=1e-10, time_limit=60.0, verbose=0, **xc)


def test_closed_form_primal_blocks(ctx):
    # x = (t, x1, x2) in SOC, t = 1:  min 3 x1 - 4 x2 = -5
    r = _one_block([[1.0, 0, 0]], [1.0], [0.0, 3.0, -4.0], soc_x=[3])
    assert r.info.exit_code == 0 and abs(r.info.pObj + 5.0) < 1e-6
    # x in K_exp, x1 = x2 = 1:  min x3 = e
    r = _one_block([[1.0, 0, 0], [0, 1.0, 0]], [1.0, 1.0], [0.0, 0.0, 1.0], exp_x=1)
    assert r.info.exit_code == 0 and abs(r.info.pObj - math.e) < 1e-6
    # x in K_exp*, u = -1, v = 0:  min w = 1/e
    r = _one_block([[1.0, 0, 0], [0, 1.0, 0]], [-1.0, 0.0], [0.0, 0.0, 1.0], dual_exp_x=1)
    assert r.info.exit_code == 0 and abs(r.info.pObj - 1.0 / math.e) < 1e-6
    # x = (u, w, z) in RSOC, z = 2:  min u + w = 2 sqrt 2
    r = _one_block([[0, 0, 1.0]], [2.0], [1.0, 1.0, 0.0], rsoc_x=[3])
    assert r.info.exit_code == 0 and abs(r.info.pObj - 2.0 * math.sqrt(2.0)) < 1e-6
    # more block variables than n
    with pytest.raises(_lib.PdcsError):
        _one_block([[1.0, 0, 0]], [1.0], [0.0, 3.0, -4.0], soc_x=[4])
