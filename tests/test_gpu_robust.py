"""GPU parity of the robust mode (N.LU_ROBUST: FP64 factors with static pivoting, every Newton step's linear solve
finished by GMRES on the FP64 Jacobian) and of the robust retry of the scenarios that the unpivoted batch rejects,
against the FP64 CPU oracle, whose sparse LU pivots by rows (threshold partial pivoting).

Bar: identical statuses; bus |V| within 1e-6 pu and angle within 1e-6 rad of the oracle. The robust mode's linear
solves are accurate to 1e-10 relative, so its Newton iteration counts equal the oracle's as well.

The test network for pivoting is grid118 with two series-resonant paths (gridgen.Network.with_series_resonance): a
load-free bus between reactances x and -x, whose own 2x2 Jacobian block is singular at the flat start and at the
solution while the Jacobian is not. The unpivoted block LU eliminates that bus early (it has two neighbours) and
meets a zero or nearly zero pivot block, so the batch rejects most such scenarios; the oracle pivots past it.
"""
import numpy as np
import pytest

import pyoracle as po
from olfgpu import _native as N
from olfgpu import gridgen
from olfgpu.solver import GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters, GpuNewtonRaphson

pytestmark = pytest.mark.gpu

V_TOL = 1e-6
ANGLE_TOL = 1e-6
PAIRS_BESIDE = [0, 40]   # grid118 branches whose ends have three or more neighbours, away from the slack


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if N.device_count() == 0:
        pytest.fail("no GPU visible: the gpu-marked tests need an MI355X")


def compare_states(xg, xo, what=""):
    dv = np.abs(xg[..., 1::2] - xo[..., 1::2]).max()
    da = np.abs(xg[..., 0::2] - xo[..., 0::2]).max()
    assert dv <= V_TOL and da <= ANGLE_TOL, f"{what}: |dV|={dv:.3e} |dtheta|={da:.3e}"


@pytest.fixture(scope="module")
def resonant(grid):
    base = grid("grid118")
    net = base.with_series_resonance(PAIRS_BESIDE, x=0.05)
    osys = po.OracleSystem(net)
    x0, tgt, md, _ = net.system()
    xb, mb, tb, ltb, res = osys.solve(x0, md, tgt)
    assert po.STATUS[res.status] == "CONVERGED"
    return base, net, osys, (x0, tgt, md), (xb, mb, tb, ltb, res)


def test_robust_mode_n1_matches_oracle(grid):
    """Every scenario in the robust mode (forced) on a network that needs no pivoting: same answers as the oracle,
    iteration for iteration."""
    net = grid("grid2000")
    print("[robust] grid2000 forced robust mode, 24 contingencies", flush=True)
    osys = po.OracleSystem(net)
    x0, tgt, md, _ = net.system()
    xb, mb, tb, ltb, res = osys.solve(x0, md, tgt)
    assert po.STATUS[res.status] == "CONVERGED"
    ctg = gridgen.sample_contingencies(net, 24, 0.25, 11)
    ref = osys.n1(xb, mb, tb, ltb, ctg)
    sa = GpuBatchedSecurityAnalysisSolver(N.Structure.from_network(net), net.reactive_limits(),
                                          force_mode=N.LU_ROBUST)
    r = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
    assert np.array_equal(r.status, ref["status"]), (r.status, ref["status"])
    assert np.array_equal(r.iterations, ref["iterations"])
    assert np.array_equal(r.total_iterations, ref["total_iterations"])
    conv = r.status == 0
    compare_states(r.x[conv], ref["x"][conv], "robust n1 grid2000")
    assert r.robust_attempted == 0


def test_resonant_base_case_is_retried_in_robust_mode(resonant):
    base, net, osys, (x0, tgt, md), (xo, mo, to, lto, ro) = resonant
    print("[robust] grid118 + 2 resonant paths, base case from flat start", flush=True)
    st = N.Structure.from_network(net)
    plain = GpuNewtonRaphson(st, net.reactive_limits(), robust_retry=False)
    r0 = plain.run(x0, tgt, md)
    assert r0.status == "LINEAR_SOLVE_FAILED", r0.status
    solver = GpuNewtonRaphson(st, net.reactive_limits())
    r = solver.run(x0, tgt, md)
    assert solver.robust_mode
    assert r.status == po.STATUS[ro.status] == "CONVERGED"
    assert r.iterations == ro.iterations and r.total_iterations == ro.total_iterations
    compare_states(r.x, xo, "resonant base")
    # the robust mode stays for the next runs of the same solver
    r2 = solver.run(x0, tgt, md)
    assert r2.status == "CONVERGED"
    compare_states(r2.x, xo, "resonant base, second run")


def test_resonant_n1_rejected_scenarios_solved_by_the_robust_retry(resonant):
    base, net, osys, _, (xb, mb, tb, ltb, _) = resonant
    e0 = base.n_branch   # the resonant paths' own branches: e0 + p (bus a side) and e0 + 2 + p (bus b side)
    ctg = np.concatenate([gridgen.sample_contingencies(net, 20, 0.1, 3), [e0, e0 + 1, e0 + 3]]).astype(np.int32)
    print(f"[robust] grid118 + 2 resonant paths, N-1 over {len(ctg)} contingencies", flush=True)
    ref = osys.n1(xb, mb, tb, ltb, ctg)
    assert (ref["status"] == 0).all()
    st = N.Structure.from_network(net)
    first = GpuBatchedSecurityAnalysisSolver(st, net.reactive_limits(), robust_retry=False).solve_batch_n1(
        xb, tb, mb, ltb, ctg)
    rejected = (first.status == N.REJECTED) | (first.status == N.LINEAR_SOLVE_FAILED)
    # every scenario keeps at least one resonant path whole (one outage each): the unpivoted factorization leaves
    # most of them (a zero pivot block, or steps too inaccurate to converge in the budget)
    assert rejected.sum() >= len(ctg) // 2, first.status
    assert np.array_equal(first.status[~rejected], ref["status"][~rejected])
    r = GpuBatchedSecurityAnalysisSolver(st, net.reactive_limits()).solve_batch_n1(xb, tb, mb, ltb, ctg)
    assert r.robust_attempted == rejected.sum()
    assert r.robust_recovered == rejected.sum()
    assert np.array_equal(r.status, ref["status"]), (r.status, ref["status"])
    # the robust solves' steps are accurate to 1e-10: the oracle's iteration counts
    assert np.array_equal(r.iterations[rejected], ref["iterations"][rejected])
    compare_states(r.x, ref["x"], "resonant n1")


def test_robust_context_rejects_several_lanes_per_tile(resonant):
    _, net, _, _, _ = resonant
    st = N.Structure.from_network(net)
    with pytest.raises(N.OlfHipError, match="one scenario per tile"):
        N.BatchContext(st, 4, N.LU_ROBUST, net.reactive_limits(), lanes_per_tile=2)


def test_resonant_9241_n1_robust_retry_matches_oracle(grid):
    """The same at full size: grid9241 with series-resonant paths beside its first two branches, 16 contingencies
    in the FP32 batch; every rejected scenario comes back from the robust retry with the oracle's status, iteration
    count and state."""
    base = grid("grid9241")
    net = base.with_series_resonance([0, 1], x=0.05)
    print("[robust] grid9241 + 2 resonant paths, N-1 over 16 contingencies", flush=True)
    osys = po.OracleSystem(net)
    x0, tgt, md, _ = net.system()
    xb, mb, tb, ltb, res = osys.solve(x0, md, tgt)
    assert po.STATUS[res.status] == "CONVERGED"
    ctg = gridgen.sample_contingencies(net, 16, 0.05, 4)
    ref = osys.n1(xb, mb, tb, ltb, ctg)
    st = N.Structure.from_network(net)
    sa = GpuBatchedSecurityAnalysisSolver(st, net.reactive_limits(),
                                          GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"))
    r = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
    assert r.robust_attempted >= 8 and r.robust_recovered == r.robust_attempted
    assert np.array_equal(r.status, ref["status"]), (r.status, ref["status"])
    conv = r.status == 0
    compare_states(r.x[conv], ref["x"][conv], "resonant n1 grid9241")


def test_resonant_robust_with_distributed_slack_and_line_search(resonant):
    """The robust retry under the outer loops: distributed slack and LINE_SEARCH scaling on the resonant network,
    base case (batch of one) and N-1, against the oracle with the same settings."""
    from olfgpu.solver import NewtonRaphsonParameters

    base, net, osys, (x0, tgt, md), _ = resonant
    print("[robust] grid118 + resonant paths, distributed slack + LINE_SEARCH", flush=True)
    ds = gridgen.distributed_slack(net)
    xo, mo, to, lto, ro = osys.solve(x0, md, tgt, prm=po.params(scaling_mode=1), ds=ds, ds_tol=1.0 / gridgen.SB)
    assert po.STATUS[ro.status] == "CONVERGED"
    st = N.Structure.from_network(net)
    nrp = NewtonRaphsonParameters(stateVectorScalingMode="LINE_SEARCH")
    solver = GpuNewtonRaphson(st, net.reactive_limits(), nr_parameters=nrp, distributed_slack=ds)
    r = solver.run(x0, tgt, md)
    assert solver.robust_mode and r.status == "CONVERGED"
    assert r.total_iterations == ro.total_iterations
    compare_states(r.x, xo, "resonant base, ds + line search")
    assert np.abs(r.target[2 * ds["bus"]] - to[2 * ds["bus"]]).max() <= V_TOL
    ctg = gridgen.sample_contingencies(net, 20, 0.1, 3)
    ref = osys.n1(xo, mo, to, lto, ctg, prm=po.params(batch_of_one=False, scaling_mode=1), ds=ds,
                  ds_tol=1.0 / gridgen.SB)
    sa = GpuBatchedSecurityAnalysisSolver(st, net.reactive_limits(), nr_parameters=nrp, distributed_slack=ds)
    g = sa.solve_batch_n1(r.x, r.target, r.row_mode, r.reactive_limit_types, ctg)
    assert g.robust_attempted > 0 and g.robust_recovered == g.robust_attempted
    assert np.array_equal(g.status, ref["status"]), (g.status, ref["status"])
    conv = g.status == 0
    compare_states(g.x[conv], ref["x"][conv], "resonant n1, ds + line search")


def test_failed_robust_pass_keeps_the_first_pass(resonant, monkeypatch):
    """The retry pass is optional: when it fails (forced from the host: its context cannot be created), the first
    pass's answers stand and the failure is reported."""
    base, net, osys, _, (xb, mb, tb, ltb, _) = resonant
    st = N.Structure.from_network(net)
    ctg = gridgen.sample_contingencies(net, 8, 0.0, 3)
    sa = GpuBatchedSecurityAnalysisSolver(st, net.reactive_limits())

    def refuse(chunk):
        raise N.OlfHipError(N.E_MEMORY, "forced: no memory for the robust context")
    monkeypatch.setattr(sa, "robust_context", refuse)
    r = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
    first = GpuBatchedSecurityAnalysisSolver(st, net.reactive_limits(), robust_retry=False).solve_batch_n1(
        xb, tb, mb, ltb, ctg)
    assert r.robust_attempted > 0 and r.robust_recovered == 0 and len(r.robust_errors) == 1
    assert np.array_equal(r.status, first.status)
