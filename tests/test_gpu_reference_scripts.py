"""GPU: the reference's GPU example scripts as tests (test/test_soc_gpu.jl:1-91,
test/test_exp_gpu.jl:1-126; the reference's scripts assert nothing).  Each
builds min 10 1'x s.t. (A x - b) in {0}^B x R+^B x K, x >= 0, with A dense
5B x 2B uniform, b = A x_fea minus the projection of A x_fea onto K (so x_fea
is feasible), and solves it three ways, as the scripts do:
  1. the modeling front-end (the scripts' JuMP model over PDCS_GPU.Optimizer);
  2. rpdhg_gpu_solve with host data;
  3. the same problem from GPU tensors (rpdhg_gpu_solve_input_gpu_data).
The scripts use B = 100; B = 40 here keeps the oracle's reference solve to a
few seconds.  Paths 1 and 3 build the same problem and, with the schedule
autotuner off, return the same solve bit for bit; the oracle pins the optimum
of all three."""
import numpy as np
import pytest

from oracle import oracle
from pdcs_amd import instances
from pdcs_amd.model import Model
from pdcs_amd.solver import Problem, make_cones, rpdhg_gpu_solve, rpdhg_gpu_solve_input_gpu_data

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

BASE = 40
N, M, MZ, MN = 2 * BASE, 5 * BASE, BASE, BASE
KW = dict(rel_tol=1e-6, abs_tol=1e-6, time_limit=120.0)


def _instance(cone, seed):
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.0, 1.0, (M, N))              # sprand(m, n, 1.0)
    x_fea = rng.uniform(0.0, 1.0, N)
    b = A @ x_fea
    proj = b.copy()
    proj[:MZ] = 0.0
    proj[MZ:MZ + MN] = np.maximum(b[MZ:MZ + MN], 0.0)
    if cone == "soc":
        proj[MZ + MN:] = oracle.proj_soc(b[MZ + MN:])
    else:
        for i in range(MZ + MN, M, 3):
            proj[i:i + 3] = oracle.proj_exp(b[i:i + 3])
    return A, b - proj, np.full(N, 10.0)


def _front_end(A, b, c, cone):
    """(model, constraint indices in row order)"""
    mdl = Model()
    mdl.add_variables(N, lower=0.0)
    mdl.set_objective(c)
    ks = [mdl.add_constraint(A[:MZ], -b[:MZ], "zeros"),
          mdl.add_constraint(A[MZ:MZ + MN], -b[MZ:MZ + MN], "nonnegatives")]
    if cone == "soc":
        ks.append(mdl.add_constraint(A[MZ + MN:], -b[MZ + MN:], "soc"))
    else:
        ks += [mdl.add_constraint(A[i:i + 3], -b[i:i + 3], "exp") for i in range(MZ + MN, M, 3)]
    return mdl, ks


@pytest.mark.parametrize("cone", ["soc", "exp"])
def test_reference_example_script(ctx, cone):
    A, b, c = _instance(cone, 2 if cone == "soc" else 1)
    socG = [M - MZ - MN] if cone == "soc" else []
    expG = 0 if cone == "soc" else (M - MZ - MN) // 3
    r, ci = np.nonzero(A)
    rowptr, colidx, vals = instances.csr_from_coo(M, r, ci, A[r, ci])
    # 2. host data
    sol = rpdhg_gpu_solve(n=N, m=M, nb=N, c=c, G=(rowptr, colidx, vals), h=b, mGzero=MZ, mGnonnegative=MN,
                          socG=socG, expG=expG, bl=np.zeros(N), bu=np.full(N, np.inf), verbose=0, ctx=ctx, **KW)
    assert sol.info.exit_code == 0, sol.info.exit_status
    # 1. the front-end (rows and signs as the scripts' constraints)
    mdl, ks = _front_end(A, b, c, cone)
    mdl.optimize(ctx, autotune=0, **KW)
    assert mdl.termination_status() == "OPTIMAL"
    # 3. GPU tensors: the same problem as the front-end's, the same solve
    cones = make_cones(MZ, MN, socG, (), expG, 0)
    dev = lambda a, dt=torch.float64: torch.as_tensor(np.asarray(a), dtype=dt).to("cuda").contiguous()
    pt = Problem.from_tensors(N, M, N, dev(c), dev(rowptr, torch.int32), dev(colidx, torch.int32), dev(vals),
                              dev(b), dev(np.zeros(N)), dev(np.full(N, np.inf)), cones, ctx)
    rt = pt.solve(autotune=0, **KW)
    xt, _, _ = pt.solution()
    assert rt["exit_code"] == 0 and rt["pobj"] == mdl.objective_value() and rt["iter"] == mdl.iterations()
    assert np.array_equal(xt, mdl.value())
    # the reference's second entry point on the same tensors
    sg = rpdhg_gpu_solve_input_gpu_data(n=N, m=M, nb=N, c=dev(c), G=(dev(rowptr, torch.int32),
                                        dev(colidx, torch.int32), dev(vals)), h=dev(b), mGzero=MZ, mGnonnegative=MN,
                                        socG=socG, expG=expG, bl=dev(np.zeros(N)), bu=dev(np.full(N, np.inf)),
                                        verbose=0, ctx=ctx, **KW)
    assert sg.info.exit_code == 0
    assert abs(sg.info.pObj - sol.info.pObj) <= 2e-6 * (1 + abs(sol.info.pObj))
    # the oracle's optimum
    P = instances.ConicProblem(n=N, m=M, nb=N, c=c, rowptr=rowptr, colidx=colidx, vals=vals, h=b,
                               bl=np.zeros(N), bu=np.full(N, np.inf), mzero=MZ, mnonneg=MN, soc=socG, nexp=expG,
                               name=f"reference_{cone}_gpu_script")
    ref, _, _ = oracle.solve(P, rel_tol=1e-8, abs_tol=1e-10)
    print(f"{cone}: gpu pobj {sol.info.pObj} ({sol.info.iter} iters), front-end {mdl.objective_value()}; "
          f"oracle {ref['pobj']} ({ref['exit_code']}, {ref['iter']} iters)", flush=True)
    assert ref["exit_code"] == 0
    for v in (sol.info.pObj, mdl.objective_value()):
        assert abs(v - ref["pobj"]) <= 2e-6 * (1 + abs(ref["pobj"])), (v, ref["pobj"])
    y = np.concatenate([mdl.constraint_dual(k) for k in ks])
    chk = oracle.kkt(P, mdl.value(), y)
    assert max(chk["rel_primal_res"], chk["rel_dual_res"], chk["rel_gap"]) < 1e-6, chk
