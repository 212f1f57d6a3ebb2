"""GPU parity of LINE_SEARCH state-vector scaling (device backtracking in the Newton-step kernel, through the C ABI)
against the oracle's (pinned in tests/test_oracle_line_search.py):

* perturbed starts on the 118-bus grid, many of which make full steps overshoot, solved as one batch (FP64 and FP32 +
  refinement): per start the same status as the oracle and, for converged ones, |V| and angle within 1e-6 (FP64 also
  the same iteration count);
* N-1 at 2,000 buses with LINE_SEARCH: statuses and states as the oracle's N-1 under the same scaling."""
import numpy as np
import pytest

import pyoracle as po
from olfgpu import _native as N
from olfgpu import gridgen
from olfgpu.solver import GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters, NewtonRaphsonParameters

pytestmark = pytest.mark.gpu


def _starts(net, count, seed):
    x0, tg, md, _ = net.system()
    rng = np.random.default_rng(seed)
    xs = np.tile(x0, (count, 1))
    xs[:, 0::2] += rng.uniform(-0.1, 0.1, (count, net.n_bus))
    xs[:, 1::2] *= rng.uniform(0.9, 1.1, (count, net.n_bus))
    xs[:, 0] = 0.0
    fixed = md == 0
    xs[:, fixed] = tg[fixed]
    return xs


@pytest.mark.parametrize("lu", ["BATCHED_LU", "BATCHED_LU_FP32_REFINED"])
def test_perturbed_starts_match_the_oracle(grid, lu):
    net = grid("grid118")
    print(f"[ls] grid118 perturbed starts {lu}", flush=True)
    x0, tg, md, _ = net.system()
    xs = _starts(net, 24, 2)
    o = po.OracleSystem(net)
    prm = po.params(reactive_limits=False, scaling_mode=1, batch_of_one=False)
    ref = [o.solve(x, md, tg, prm=prm) for x in xs]
    st = N.Structure.from_network(net)
    nrp = NewtonRaphsonParameters(stateVectorScalingMode="LINE_SEARCH", reactiveLimits=False)
    sa = GpuBatchedSecurityAnalysisSolver(st, None, GpuLoadFlowParameters(linearSolver=lu), nrp)
    ctg = np.full((len(xs), 1), -1, np.int32)
    r = sa.solve_batch_n1(x0, tg, md, None, ctg, x0_per_scenario=xs)
    want = np.array([q[4].status for q in ref])
    assert np.array_equal(r.status, want), (r.status, want)
    plain = [o.solve(x, md, tg, prm=po.params(reactive_limits=False, batch_of_one=False)) for x in xs]
    print(f"[ls] {int((want == 0).sum())} of {len(xs)} converge with LINE_SEARCH, "
          f"{int(sum(q[4].status == 0 for q in plain))} without", flush=True)
    for j, q in enumerate(ref):
        if q[4].status != 0:
            continue
        dv = np.abs(r.x[j][1::2] - q[0][1::2]).max()
        da = np.abs(r.x[j][0::2] - q[0][0::2]).max()
        assert dv <= 1e-6 and da <= 1e-6, (j, dv, da)
        if lu == "BATCHED_LU":
            assert r.iterations[j] == q[4].iterations, (j, r.iterations[j], q[4].iterations)


def test_n1_with_line_search(grid):
    net = grid("grid2000")
    print("[ls] grid2000 N-1", flush=True)
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    xb, mb, tb, ltb, rb = o.solve(x0, md, tg)
    assert rb.status == 0
    ctg = gridgen.sample_contingencies(net, 40, 0.25, 9)
    ref = o.n1(xb, mb, tb, ltb, ctg, prm=po.params(scaling_mode=1, batch_of_one=False))
    st = N.Structure.from_network(net)
    nrp = NewtonRaphsonParameters(stateVectorScalingMode="LINE_SEARCH")
    sa = GpuBatchedSecurityAnalysisSolver(st, net.reactive_limits(),
                                          GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"), nrp)
    r = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
    assert np.array_equal(r.status, ref["status"])
    ok = ref["status"] == 0
    assert np.abs(r.x[ok][:, 1::2] - ref["x"][ok][:, 1::2]).max() <= 1e-6
    assert np.abs(r.x[ok][:, 0::2] - ref["x"][ok][:, 0::2]).max() <= 1e-6
