"""GPU: the torch bridge (replaces the reference's juliacall/cupy bridge,
src/pdcs_gpu/cvxpy_wrapper/py2jl.jl:5-29, and its fixed-pattern re-solve,
data_updating.jl:4-75): problems created and updated from GPU tensors through
pdcs_problem_create_device / pdcs_problem_update_device, and the resident
ConicLayer, whose gradients equal those of the per-call ConicSolve."""
import numpy as np
import pytest

from pdcs_amd import instances
from pdcs_amd.solver import Problem, make_cones

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

KW = dict(rel_tol=1e-8, abs_tol=1e-14, autotune=0, time_limit=30.0)


def _fields(r):
    return (r["exit_code"], r["iter"], r["pobj"], r["dobj"], r["rel_primal_res"], r["rel_dual_res"], r["rel_gap"])


def _cuda(a, dt=torch.float64):
    return torch.as_tensor(np.asarray(a), dtype=dt).to("cuda").contiguous()


def test_create_and_update_from_tensors(ctx):
    P = instances.soc_program(300, 600, 0.02, 10, seed=31)
    ref = Problem.from_instance(P, ctx)
    r_ref = ref.solve(autotune=0, time_limit=30.0)
    x_ref, y_ref, _ = ref.solution()
    cones = make_cones(P.mzero, P.mnonneg, P.soc, (), P.nexp, P.ndexp)
    dev = Problem.from_tensors(P.n, P.m, P.nb, _cuda(P.c), _cuda(P.rowptr, torch.int32), _cuda(P.colidx, torch.int32),
                               _cuda(P.vals), _cuda(P.h), _cuda(P.bl[:P.nb]), _cuda(P.bu[:P.nb]), cones, ctx)
    r_dev = dev.solve(autotune=0, time_limit=30.0)
    x_dev, y_dev, _ = dev.solution()
    assert _fields(r_dev) == _fields(r_ref)
    assert np.array_equal(x_dev, x_ref) and np.array_equal(y_dev, y_ref)
    # the same update from host arrays and from GPU tensors
    rng = np.random.default_rng(5)
    c2 = P.c + rng.normal(0.0, 0.05, P.n)
    v2 = P.vals * rng.uniform(0.9, 1.1, P.nnz)
    ref.update(c=c2, vals=v2)
    dev.update_tensors(c=_cuda(c2), vals=_cuda(v2))
    assert _fields(dev.solve(autotune=0, time_limit=30.0)) == _fields(ref.solve(autotune=0, time_limit=30.0))
    with pytest.raises(ValueError):
        dev.update_tensors(c=_cuda(c2[:-1]))
    with pytest.raises(ValueError):
        dev.update_tensors(c=_cuda(c2, torch.float32))


def test_conic_layer_matches_conic_solve(ctx):
    from pdcs_amd.torch_layer import ConicLayer, ConicSolve
    P = instances.workload4_layer(9)
    rng = np.random.default_rng(13)
    xt = _cuda(rng.standard_normal(P.n))
    data = [(P.c, P.h, P.vals),
            (P.c + rng.normal(0.0, 0.05, P.n), P.h, P.vals * rng.uniform(0.95, 1.05, P.nnz))]

    def leaf(a):
        return _cuda(a).requires_grad_(True)

    # per-call reference: a fresh problem for each call
    ref = []
    for c, h, v in data:
        tc, th, tv = leaf(c), leaf(h), leaf(v)
        x = ConicSolve.apply(tc, th, tv, P, KW)
        ((x - xt) ** 2).sum().backward()
        ref.append((x.detach().cpu().numpy(), tc.grad.cpu().numpy(), th.grad.cpu().numpy(), tv.grad.cpu().numpy()))
    layer = ConicLayer(P, ctx, **KW)
    # one call at a time
    for (c, h, v), (xr, gc, gh, gv) in zip(data, ref):
        tc, th, tv = leaf(c), leaf(h), leaf(v)
        x = layer(tc, th, tv)
        assert x.device.type == "cuda"
        ((x - xt) ** 2).sum().backward()
        assert np.array_equal(x.detach().cpu().numpy(), xr)
        for g, r in ((tc.grad, gc), (th.grad, gh), (tv.grad, gv)):
            assert np.array_equal(g.cpu().numpy(), r)
    # two calls, then one backward: the first call's data is solved again
    leaves = [[leaf(a) for a in d] for d in data]
    xs = [layer(*lv) for lv in leaves]
    sum(((x - xt) ** 2).sum() for x in xs).backward()
    for lv, (_, gc, gh, gv) in zip(leaves, ref):
        for g, r in zip((t.grad for t in lv), (gc, gh, gv)):
            assert np.array_equal(g.cpu().numpy(), r)
    layer.close()


def test_batch_conic_layer_matches_batch_conic_solve(ctx):
    """The resident batched layer (Batch.update per call) against the per-call
    BatchConicSolve: the same solutions and gradients bit for bit, also when
    two calls precede one backward."""
    from pdcs_amd.torch_layer import BatchConicLayer, BatchConicSolve
    probs = [instances.workload4_layer(k) for k in range(5)]
    rng = np.random.default_rng(21)
    xts = [_cuda(rng.standard_normal(P.n)) for P in probs]
    kw = dict(rel_tol=1e-8, abs_tol=1e-14, time_limit=30.0)

    def data(shift):
        return [(P.c + shift * rng.normal(0.0, 0.05, P.n), P.h, P.vals * (1 + shift * rng.uniform(-0.05, 0.05, P.nnz)))
                for P in probs]

    datas = [data(0.0), data(1.0)]

    def leaves(d):
        return [_cuda(a).requires_grad_(True) for trip in d for a in trip]

    def loss(xs):
        return sum(((x - t) ** 2).sum() for x, t in zip(xs, xts))

    ref = []
    for d in datas:
        lv = leaves(d)
        xs = BatchConicSolve.apply(probs, kw, *lv)
        loss(xs).backward()
        ref.append(([x.detach().cpu().numpy() for x in xs], [t.grad.cpu().numpy() for t in lv]))
    layer = BatchConicLayer(probs, ctx, **kw)
    for d, (xr, gr) in zip(datas, ref):
        lv = leaves(d)
        xs = layer(*lv)
        loss(xs).backward()
        assert all(np.array_equal(x.detach().cpu().numpy(), r) for x, r in zip(xs, xr))
        assert all(np.array_equal(t.grad.cpu().numpy(), r) for t, r in zip(lv, gr))
    lvs = [leaves(d) for d in datas]
    outs = [layer(*lv) for lv in lvs]
    (loss(outs[0]) + loss(outs[1])).backward()
    for lv, (_, gr) in zip(lvs, ref):
        assert all(np.array_equal(t.grad.cpu().numpy(), r) for t, r in zip(lv, gr))
    layer.close()
