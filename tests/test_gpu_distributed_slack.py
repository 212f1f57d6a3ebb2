"""GPU parity of the distributed-slack outer loop (device check in the Newton-step kernel, through the C ABI)
against the oracle's (pinned in tests/test_oracle_distributed_slack.py):

* base case at 2,000 buses, FP64 and FP32 + refinement, batch of one: status identical, |V| and angle within 1e-6,
  every participant's final target within 1e-6 pu (the targets follow the state through the slack's injection);
* N-1 at 2,000 buses (with a participant cut off by its outaged branches) and 9,241 buses from the distributed base:
  the same bar per contingency;
* a distribution that cannot place the imbalance: CONVERGED under LEAVE_ON_SLACK_BUS, REJECTED under FAIL, as in
  the oracle."""
import numpy as np
import pytest

import pyoracle as po
from olfgpu import _native as N
from olfgpu import gridgen
from olfgpu.solver import (GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters, GpuNewtonRaphson,
                           NewtonRaphsonParameters)

pytestmark = pytest.mark.gpu

TOL = 1e-6
SB = gridgen.SB


def _compare(xg, xo, tg, to, ds, what):
    dv = np.abs(xg[..., 1::2] - xo[..., 1::2]).max()
    da = np.abs(xg[..., 0::2] - xo[..., 0::2]).max()
    dt = np.abs(tg[..., 2 * ds["bus"]] - to[..., 2 * ds["bus"]]).max()
    assert dv <= TOL and da <= TOL and dt <= TOL, f"{what}: |dV|={dv:.3e} |dth|={da:.3e} |dP target|={dt:.3e}"


@pytest.mark.parametrize("lu", ["BATCHED_LU", "BATCHED_LU_FP32_REFINED"])
def test_base_case_with_distributed_slack(grid, lu):
    net = grid("grid2000")
    print(f"[ds] grid2000 base {lu}", flush=True)
    ds = gridgen.distributed_slack(net)
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    xo, mo, to, lto, ro = o.solve(x0, md, tg, ds=ds, ds_tol=1.0 / SB)
    assert po.STATUS[ro.status] == "CONVERGED" and ro.outer_switches > 0
    st = N.Structure.from_network(net)
    g = GpuNewtonRaphson(st, net.reactive_limits(), GpuLoadFlowParameters(linearSolver=lu),
                         distributed_slack=ds).run(x0, tg, md)
    assert g.status == "CONVERGED"
    _compare(g.x, xo, g.target, to, ds, f"base {lu}")
    np.testing.assert_array_equal(g.row_mode, mo)
    if lu == "BATCHED_LU":
        assert g.total_iterations == ro.total_iterations


@pytest.mark.parametrize("name, count, bridges", [("grid2000", 40, 0.25), ("grid9241", 40, 0.1)])
def test_n1_with_distributed_slack(grid, name, count, bridges):
    net = grid(name)
    print(f"[ds] {name} N-1", flush=True)
    ds = gridgen.distributed_slack(net)
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    xb, mb, tb, ltb, rb = o.solve(x0, md, tg, ds=ds, ds_tol=1.0 / SB)
    assert po.STATUS[rb.status] == "CONVERGED"
    ctg = np.full((count, 4), -1, np.int32)
    ctg[:, 0] = gridgen.sample_contingencies(net, count, bridges, 5)
    # the last one cuts a participant off: every branch at a PV bus of degree <= 4
    deg = np.bincount(np.concatenate([net.bus1, net.bus2]), minlength=net.n_bus)
    bus = int(next(b for b in ds["bus"] if deg[b] <= 4))
    outs = np.nonzero((net.bus1 == bus) | (net.bus2 == bus))[0]
    ctg[-1] = -1
    ctg[-1, :len(outs)] = outs
    ref = o.n1(xb, mb, tb, ltb, ctg, ds=ds, ds_tol=1.0 / SB, want_targets=True)
    st = N.Structure.from_network(net)
    rl = net.reactive_limits()
    sa = GpuBatchedSecurityAnalysisSolver(st, rl, GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"),
                                          distributed_slack=ds)
    r = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
    assert np.array_equal(r.status, ref["status"]), (r.status, ref["status"])
    _, t_dev, _ = sa._ctx.read_outer_loop_state(len(ctg), len(rl["voltage_row"]))
    ok = ref["status"] == 0
    _compare(r.x[ok], ref["x"][ok], t_dev[ok], ref["target"][ok], ds, f"{name} N-1")
    moved = np.abs(t_dev[ok][:, 2 * ds["bus"]] - tb[2 * ds["bus"]]).max(axis=1)
    print(f"[ds] {name}: {int((moved > 1e-6).sum())} of {int(ok.sum())} contingencies redistributed", flush=True)
    assert (moved > 1e-6).sum() > 0


def test_failed_distribution_leave_or_reject(grid):
    net = grid("grid118")
    net.pload[7] += 400.0
    print("[ds] grid118 failed distribution", flush=True)
    ds = gridgen.distributed_slack(net, headroom=1.01)
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    st = N.Structure.from_network(net)
    for behaviour, reject in (("LEAVE_ON_SLACK_BUS", False), ("FAIL", True)):
        xo, mo, to, lto, ro = o.solve(x0, md, tg, ds=ds, ds_tol=1.0 / SB, ds_reject=reject)
        g = GpuNewtonRaphson(st, net.reactive_limits(), GpuLoadFlowParameters(),
                             NewtonRaphsonParameters(slackDistributionFailureBehavior=behaviour),
                             distributed_slack=ds).run(x0, tg, md)
        assert g.status == po.STATUS[ro.status], behaviour
        if g.status == "CONVERGED":
            _compare(g.x, xo, g.target, to, ds, behaviour)
    assert g.status == "REJECTED"
