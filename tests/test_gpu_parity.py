"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle on the same seeded inputs.

Bar (from the workload definition): per scenario identical status; bus |V| within 1e-6 pu and angle within 1e-6 rad
of the oracle (both stop at ||F||inf < 1e-4 pu). The standalone LU matches the oracle's pivoted sparse solve to
1e-9 relative in FP64 and to 1e-3 relative in FP32.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import pyoracle as po
from olfgpu import _native as N
from olfgpu import gridgen
from olfgpu.solver import (GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters, GpuNewtonRaphson,
                           NewtonRaphsonParameters)

pytestmark = pytest.mark.gpu

V_TOL = 1e-6
ANGLE_TOL = 1e-6


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if N.device_count() == 0:
        pytest.fail("no GPU visible: the gpu-marked tests need an MI355X")


_BASE = {}


def base_case(net, name):
    """Converged base case from the oracle: (x, mode, target, limit types)."""
    if name not in _BASE:
        osys = po.OracleSystem(net)
        x0, tgt, mode, _ = net.system()
        xb, mb, tb, ltb, res = osys.solve(x0, mode, tgt)
        assert po.STATUS[res.status] == "CONVERGED"
        _BASE[name] = (osys, xb, mb, tb, ltb)
    return _BASE[name]


def compare_states(xg, xo, what=""):
    dv = np.abs(xg[..., 1::2] - xo[..., 1::2]).max()
    da = np.abs(xg[..., 0::2] - xo[..., 0::2]).max()
    assert dv <= V_TOL and da <= ANGLE_TOL, f"{what}: |dV|={dv:.3e} |dtheta|={da:.3e}"
    return dv, da


@pytest.mark.parametrize("name", ["grid118", "grid2000"])
@pytest.mark.parametrize("mode", [N.LU_FP64, N.LU_FP32_REFINED])
def test_standalone_lu_matches_oracle_solve(grid, name, mode):
    net = grid(name)
    print(f"[lu] {name} mode={mode}", flush=True)
    osys = po.OracleSystem(net)
    st = N.Structure.from_network(net)
    x0, tgt, md, _ = net.system()
    rng = np.random.default_rng(1)
    lanes = 12
    vals = np.zeros((lanes, osys.nnz))
    rhs = rng.normal(size=(lanes, net.n))
    ref = np.zeros((lanes, net.n))
    for l in range(lanes):
        x = x0.copy()
        x[0::2] = rng.normal(0, 0.1, net.n_bus)
        x[1::2] = 1 + rng.normal(0, 0.03, net.n_bus)
        _, _, J = osys.fill(x, md, tgt)
        vals[l] = J
        ref[l] = po.lu_solve(osys.rowptr, osys.colidx, J, rhs[l])
    sol, clamps = st.lu_factorize_and_solve(vals, rhs, mode=mode)
    assert (clamps == 0).all()
    if mode == N.LU_FP64:
        rel = np.abs(sol - ref).max(axis=1) / np.abs(ref).max(axis=1)
        assert rel.max() < 1e-9, rel.max()
    # normwise backward error ||A x - b|| / (||A|| ||x|| + ||b||): the property a stable factorization guarantees at
    # its working precision (forward error also scales with the condition number)
    for l in range(lanes):
        A = sp.csr_matrix((vals[l], osys.colidx, osys.rowptr), shape=(net.n, net.n))
        anorm = np.abs(A).sum(axis=1).max()
        berr = np.abs(A @ sol[l] - rhs[l]).max() / (anorm * np.abs(sol[l]).max() + np.abs(rhs[l]).max())
        assert berr < (1e-13 if mode == N.LU_FP64 else 1e-5), (l, berr)


def test_flat_start_118_matches_oracle(grid):
    net = grid("grid118")
    print("[base] grid118 flat start", flush=True)
    osys = po.OracleSystem(net)
    x0, tgt, md, _ = net.system()
    xo, mo, to, lto, ro = osys.solve(x0, md, tgt)
    solver = GpuNewtonRaphson(N.Structure.from_network(net), net.reactive_limits())
    r = solver.run(x0, tgt, md)
    assert r.status == po.STATUS[ro.status] == "CONVERGED"
    assert r.iterations == ro.iterations and r.total_iterations == ro.total_iterations
    compare_states(r.x, xo, "grid118")


@pytest.mark.parametrize("lu", ["BATCHED_LU", "BATCHED_LU_FP32_REFINED"])
def test_base_2000_with_reactive_limits(grid, lu):
    net = grid("grid2000")
    print(f"[base] grid2000 {lu}", flush=True)
    osys, xb, mb, tb, ltb = base_case(net, "grid2000")
    x0, tgt, md, _ = net.system()
    solver = GpuNewtonRaphson(N.Structure.from_network(net), net.reactive_limits(),
                              GpuLoadFlowParameters(linearSolver=lu))
    r = solver.run(x0, tgt, md)
    assert r.status == "CONVERGED"
    compare_states(r.x, xb, f"grid2000 {lu}")


def run_n1(net, name, count, bridge_fraction, seed, lu, lanes_per_tile=0, threads=0):
    osys, xb, mb, tb, ltb = base_case(net, name)
    ctg = gridgen.sample_contingencies(net, count, bridge_fraction, seed)
    ref = osys.n1(xb, mb, tb, ltb, ctg)
    sa = GpuBatchedSecurityAnalysisSolver(N.Structure.from_network(net), net.reactive_limits(),
                                          GpuLoadFlowParameters(linearSolver=lu), lanes_per_tile=lanes_per_tile,
                                          threads_per_tile=threads)
    res = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
    return ctg, ref, res


@pytest.mark.parametrize("lu", ["BATCHED_LU", "BATCHED_LU_FP32_REFINED"])
def test_n1_2000_matches_oracle(grid, lu):
    net = grid("grid2000")
    print(f"[n1] grid2000 {lu}", flush=True)
    ctg, ref, res = run_n1(net, "grid2000", 48, 0.25, 5, lu)
    assert np.array_equal(res.status, ref["status"]), (res.status, ref["status"])
    assert (res.islanded_count > 0).sum() >= 10
    conv = res.status == 0
    compare_states(res.x[conv], ref["x"][conv], f"n1 grid2000 {lu}")
    if lu == "BATCHED_LU":
        assert np.array_equal(res.iterations, ref["iterations"])
        assert np.array_equal(res.total_iterations, ref["total_iterations"])


@pytest.mark.parametrize("lu", ["BATCHED_LU", "BATCHED_LU_FP32_REFINED"])
def test_n1_9241_matches_oracle(grid, lu):
    net = grid("grid9241")
    print(f"[n1] grid9241 {lu}", flush=True)
    ctg, ref, res = run_n1(net, "grid9241", 40, 0.1, 2, lu)
    assert np.array_equal(res.status, ref["status"])
    conv = res.status == 0
    compare_states(res.x[conv], ref["x"][conv], f"n1 grid9241 {lu}")


def test_tile_shapes_agree(grid):
    """Tile shapes change the summation order only: the lane-tile width sets how many threads share a long
    contribution sum (capped at 64 / T), and the LDS a workgroup may use (which depends on the workgroup size) sets
    how many top nodes are factored as one dense block. Results agree to rounding; a rerun of one shape is bitwise
    identical (test_rerun_is_bitwise_identical)."""
    net = grid("grid2000")
    print("[tiles] grid2000", flush=True)
    outs = {}
    for T, NT in [(1, 256), (1, 1024), (4, 512), (4, 1024), (8, 1024), (16, 256), (32, 512)]:
        ctg, ref, res = run_n1(net, "grid2000", 24, 0.25, 6, "BATCHED_LU", T, NT)
        outs[(T, NT)] = res
    base = outs[(1, 256)]
    for key, r in outs.items():
        assert np.array_equal(r.status, base.status), key
        assert np.abs(r.x - base.x).max() < 1e-9, key


def test_rerun_is_bitwise_identical(grid):
    net = grid("grid2000")
    a = run_n1(net, "grid2000", 16, 0.25, 8, "BATCHED_LU_FP32_REFINED")[2]
    b = run_n1(net, "grid2000", 16, 0.25, 8, "BATCHED_LU_FP32_REFINED")[2]
    assert np.array_equal(a.x, b.x) and np.array_equal(a.status, b.status)


@pytest.mark.parametrize("lu", ["BATCHED_LU", "BATCHED_LU_FP32_REFINED"])
def test_a_lane_answer_does_not_depend_on_its_tile_partner(grid, lu):
    """Once one lane of a tile is left solving, the tile runs its passes in the work units of that lane alone (every
    thread on it); its arithmetic is the same as beside a partner, so each scenario's answer is bitwise the one it
    gets alone in a tile (the other lane padding, a lane alone from the first step)."""
    net = grid("grid2000")
    print(f"[partner] grid2000 {lu}", flush=True)
    osys, xb, mb, tb, ltb = base_case(net, "grid2000")
    ctg = gridgen.sample_contingencies(net, 12, 0.25, 9)
    sa = GpuBatchedSecurityAnalysisSolver(N.Structure.from_network(net), net.reactive_limits(),
                                          GpuLoadFlowParameters(linearSolver=lu), lanes_per_tile=2,
                                          threads_per_tile=1024)
    together = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
    pairs = together.total_iterations.reshape(-1, 2)
    assert (pairs[:, 0] != pairs[:, 1]).any(), "no tile ran a lane alone"
    for j in range(len(ctg)):
        alone = sa.solve_batch_n1(xb, tb, mb, ltb, ctg[j:j + 1])
        assert alone.status[0] == together.status[j], j
        assert alone.total_iterations[0] == together.total_iterations[j], j
        assert np.array_equal(alone.x[0], together.x[j]), j


def test_statuses_forced_from_the_host(grid):
    net = grid("grid118")
    print("[status] grid118", flush=True)
    st = N.Structure.from_network(net)
    x0, tgt, md, _ = net.system()
    ctx = N.BatchContext(st, 3, N.LU_FP64, net.reactive_limits())
    xp = np.tile(x0, (3, 1))
    xp[1, 7] = np.nan
    inp, keep = N.BatchContext.make_inputs(3, x0, tgt, md, x0_per_scenario=xp)
    r = ctx.solve(inp)
    assert [N.STATUS_NAMES[s] for s in r["status"]] == ["CONVERGED", "DIVERGED", "CONVERGED"]
    assert np.array_equal(r["x"][0], r["x"][2])
    inp, keep = N.BatchContext.make_inputs(3, x0, tgt, md, max_iter=1, skip_reactive_limits=True)
    r = ctx.solve(inp)
    assert all(N.STATUS_NAMES[s] == "MAX_ITERATION_REACHED" for s in r["status"])
    assert (r["iterations"] == 1).all()


def test_shared_and_per_scenario_seeds_agree(grid):
    """One shared row of reactive-limit seeds gives bitwise the same solves as that row repeated per scenario, and
    per-scenario rows still apply per scenario."""
    net = grid("grid2000")
    print("[seed] grid2000 shared vs per-scenario reactive-limit seeds", flush=True)
    osys, xb, mb, tb, ltb = base_case(net, "grid2000")
    st = N.Structure.from_network(net)
    ctg = gridgen.sample_contingencies(net, 12, 0.0, 21)
    ctx = N.BatchContext(st, len(ctg), N.LU_FP64, net.reactive_limits())
    lt = np.asarray(ltb, np.int8)
    res = []
    for seed in (lt, np.tile(lt, (len(ctg), 1))):
        inp, keep = N.BatchContext.make_inputs(len(ctg), xb, tb, mb, outaged_branches=ctg, reactive_limits_seed_state=seed)
        assert inp.reactive_limits_seed_shared == (1 if seed.ndim == 1 else 0)
        res.append(ctx.solve(inp))
    assert np.array_equal(res[0]["x"], res[1]["x"]) and np.array_equal(res[0]["status"], res[1]["status"])
    # a different seed row for one scenario changes only that scenario's start
    per = np.tile(lt, (len(ctg), 1))
    per[3] = 0
    inp, keep = N.BatchContext.make_inputs(len(ctg), xb, tb, mb, outaged_branches=ctg, reactive_limits_seed_state=per)
    r = ctx.solve(inp)
    others = np.arange(len(ctg)) != 3
    assert np.array_equal(r["x"][others], res[0]["x"][others])


def test_states_left_on_device_read_back(grid):
    net = grid("grid118")
    st = N.Structure.from_network(net)
    x0, tgt, md, _ = net.system()
    ctx = N.BatchContext(st, 2, N.LU_FP64, net.reactive_limits())
    inp, keep = N.BatchContext.make_inputs(2, x0, tgt, md, states_to_host=True)
    a = ctx.solve(inp)
    inp, keep = N.BatchContext.make_inputs(2, x0, tgt, md, states_to_host=False)
    b = ctx.solve(inp, want_states=False)
    assert np.array_equal(ctx.read_states(2), a["x"])
    assert np.array_equal(a["status"], b["status"])


def test_n1_with_max_voltage_change_scaling_matches_oracle(grid):
    """MAX_VOLTAGE_CHANGE step scaling (state_vector_scaling.cu:57-97) with limits small enough to act on most steps:
    statuses, iterations and states equal the oracle's with the same scaling."""
    net = grid("grid2000")
    print("[n1] grid2000 MAX_VOLTAGE_CHANGE", flush=True)
    osys, xb, mb, tb, ltb = base_case(net, "grid2000")
    ctg = gridgen.sample_contingencies(net, 32, 0.25, 21)
    dv, dphi = 0.004, 0.02
    ref = osys.n1(xb, mb, tb, ltb, ctg, prm=po.params(scaling_mode=2, max_dv=dv, max_dphi=dphi,
                                                       batch_of_one=False))
    nrp = NewtonRaphsonParameters(stateVectorScalingMode="MAX_VOLTAGE_CHANGE", maxVoltageChangeV=dv,
                                  maxVoltageChangePhi=dphi)
    sa = GpuBatchedSecurityAnalysisSolver(N.Structure.from_network(net), net.reactive_limits(),
                                          GpuLoadFlowParameters(linearSolver="BATCHED_LU"), nrp)
    res = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
    assert np.array_equal(res.status, ref["status"])
    assert np.array_equal(res.total_iterations, ref["total_iterations"])
    # the limits made the solves take more iterations than unscaled ones
    plain = osys.n1(xb, mb, tb, ltb, ctg)
    assert ref["total_iterations"].sum() > plain["total_iterations"].sum()
    conv = res.status == 0
    compare_states(res.x[conv], ref["x"][conv], "n1 grid2000 MAX_VOLTAGE_CHANGE")


def test_longest_first_order_leaves_every_result_unchanged(grid):
    """Solving in a cost order (longest_first) changes which scenarios share a tile and when they start, not what
    each lane computes: per-contingency results are bitwise those of the given order."""
    net = grid("grid2000")
    print("[parity] grid2000 cost order", flush=True)
    osys, xb, mb, tb, ltb = base_case(net, "grid2000")
    st = N.Structure.from_network(net)
    sa = GpuBatchedSecurityAnalysisSolver(st, net.reactive_limits(),
                                          GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"))
    ctg = gridgen.sample_contingencies(net, 64, 0.2, 11)
    a = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
    cost = np.random.default_rng(0).random(len(ctg))
    b = sa.solve_batch_n1(xb, tb, mb, ltb, ctg, cost=cost)
    assert np.array_equal(a.status, b.status) and np.array_equal(a.total_iterations, b.total_iterations)
    np.testing.assert_array_equal(a.x, b.x)


@pytest.mark.parametrize("lu", ["BATCHED_LU", "BATCHED_LU_FP32_REFINED"])
def test_n1_with_two_norm_stopping_criterion(grid, lu):
    """OLF's ||F||_2 < sqrt(eps^2 * equations) criterion (convEpsPerEq = 1e-4, OLF's default), with each
    contingency's equation count lowered by its islanded buses: statuses, FP64 iteration counts and states as the
    oracle's N-1 under the same rule."""
    net = grid("grid2000")
    print(f"[parity] grid2000 N-1 two-norm criterion {lu}", flush=True)
    osys, xb, mb, tb, ltb = base_case(net, "grid2000")
    ctg = gridgen.sample_contingencies(net, 48, 0.25, 13)
    ref = osys.n1(xb, mb, tb, ltb, ctg, prm=po.params(batch_of_one=False, conv_eps_per_eq=1e-4, tol=0.0))
    st = N.Structure.from_network(net)
    sa = GpuBatchedSecurityAnalysisSolver(st, net.reactive_limits(), GpuLoadFlowParameters(linearSolver=lu),
                                          NewtonRaphsonParameters(convEpsPerEq=1e-4))
    r = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
    assert np.array_equal(r.status, ref["status"])
    if lu == "BATCHED_LU":
        assert np.array_equal(r.total_iterations, ref["total_iterations"])
    ok = ref["status"] == 0
    compare_states(r.x[ok], ref["x"][ok], f"two-norm {lu}")


@pytest.mark.parametrize("lu", ["BATCHED_LU", "BATCHED_LU_FP32_REFINED"])
def test_bus_shunts_base_and_n1(grid, lu):
    """The bus-shunt family (the workload has none; Network.with_shunts puts reactors and capacitors on 5 % of the
    buses): base case and N-1 against the oracle's shunt fill (pinned in tests/test_oracle_shunts.py)."""
    net = grid("grid2000").with_shunts(0.05)
    print(f"[parity] grid2000 with shunts {lu}", flush=True)
    o = po.OracleSystem(net)
    x0, tgt, md, _ = net.system()
    xb, mb, tb, ltb, rb = o.solve(x0, md, tgt)
    assert rb.status == 0
    st = N.Structure.from_network(net)
    rl = net.reactive_limits()
    g = GpuNewtonRaphson(st, rl, GpuLoadFlowParameters(linearSolver=lu)).run(x0, tgt, md)
    assert g.status == "CONVERGED"
    compare_states(g.x, xb, f"shunt base {lu}")
    if lu == "BATCHED_LU":
        assert g.total_iterations == rb.total_iterations
    ctg = gridgen.sample_contingencies(net, 40, 0.2, 21)
    ref = o.n1(xb, mb, tb, ltb, ctg)
    sa = GpuBatchedSecurityAnalysisSolver(st, rl, GpuLoadFlowParameters(linearSolver=lu))
    r = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
    assert np.array_equal(r.status, ref["status"])
    ok = ref["status"] == 0
    compare_states(r.x[ok], ref["x"][ok], f"shunt N-1 {lu}")


@pytest.mark.parametrize("lu", ["BATCHED_LU", "BATCHED_LU_FP32_REFINED"])
@pytest.mark.parametrize("k_out", [2, 4])
def test_multi_branch_outages_match_oracle(grid, lu, k_out):
    """Scenarios with 2 and 4 outaged branches (the most the library models per scenario; some cut buses off):
    same statuses and states as the oracle, and in FP64 the same iteration counts."""
    net = grid("grid2000")
    print(f"[n-{k_out}] grid2000 {lu}", flush=True)
    osys, xb, mb, tb, ltb = base_case(net, "grid2000")
    rng = np.random.default_rng(30 + k_out)
    ctg = np.stack([rng.choice(net.n_branch, k_out, replace=False) for _ in range(24)]).astype(np.int32)
    ref = osys.n1(xb, mb, tb, ltb, ctg)
    sa = GpuBatchedSecurityAnalysisSolver(N.Structure.from_network(net), net.reactive_limits(),
                                          GpuLoadFlowParameters(linearSolver=lu))
    res = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
    assert np.array_equal(res.status, ref["status"]), (res.status, ref["status"])
    conv = res.status == 0
    assert conv.sum() >= 20
    compare_states(res.x[conv], ref["x"][conv], f"n-{k_out} grid2000 {lu}")
    if lu == "BATCHED_LU":
        assert np.array_equal(res.iterations, ref["iterations"])


@pytest.mark.parametrize("lu", ["BATCHED_LU", "BATCHED_LU_FP32_REFINED"])
def test_per_equation_type_criteria_match_oracle(grid, lu):
    """OLF's PerEquationTypeStoppingCriteria (the reference's per-row thresholds, setBatchRowCriteria) with the
    reference test's values (1e-2 MW, 1e-2 MVar, 1e-4 pu, 1e-4 rad): base case from the flat start and N-1, same
    statuses and states as the oracle under the same rows and the reactive-limit band the criterion sets
    (maxReactivePowerMismatch / base), FP64 iteration counts equal; the rule is in force (the uniform 1e-4 pu
    criterion stops some scenarios earlier)."""
    from olfgpu.solver import PerEquationTypeStoppingCriteria

    net = grid("grid2000")
    print(f"[criteria] grid2000 per-equation-type {lu}", flush=True)
    crit = PerEquationTypeStoppingCriteria(maxActivePowerMismatch=1e-2, maxReactivePowerMismatch=1e-2,
                                           maxVoltageMismatch=1e-4, maxAngleMismatch=1e-4, base_mva=gridgen.SB)
    thr = crit.row_thresholds(net.n)
    nrp = NewtonRaphsonParameters(stoppingCriteria=crit)
    band = nrp.reactive_limit_band()   # the criterion's maxReactivePowerMismatch / base, as OLF derives it
    assert band == pytest.approx(1e-2 / gridgen.SB)
    osys = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    xo, mo, to, lto, ro = osys.solve(x0, md, tg, prm=po.params(row_thresholds=thr, rl_tol=band))
    assert po.STATUS[ro.status] == "CONVERGED"
    st = N.Structure.from_network(net)
    g = GpuNewtonRaphson(st, net.reactive_limits(), GpuLoadFlowParameters(linearSolver=lu), nrp).run(x0, tg, md)
    assert g.status == "CONVERGED"
    compare_states(g.x, xo, f"per-row base {lu}")
    if lu == "BATCHED_LU":
        assert g.total_iterations == ro.total_iterations
    ctg = gridgen.sample_contingencies(net, 32, 0.25, 17)
    ref = osys.n1(xo, mo, to, lto, ctg, prm=po.params(batch_of_one=False, row_thresholds=thr, rl_tol=band))
    sa = GpuBatchedSecurityAnalysisSolver(st, net.reactive_limits(), GpuLoadFlowParameters(linearSolver=lu), nrp)
    res = sa.solve_batch_n1(xo, to, mo, lto, ctg)
    assert np.array_equal(res.status, ref["status"]), (res.status, ref["status"])
    conv = res.status == 0
    compare_states(res.x[conv], ref["x"][conv], f"per-row n1 {lu}")
    if lu == "BATCHED_LU":
        assert np.array_equal(res.total_iterations, ref["total_iterations"])
    uniform = osys.n1(xo, mo, to, lto, ctg, prm=po.params(batch_of_one=False, rl_tol=band))
    assert (uniform["total_iterations"] < ref["total_iterations"]).any()


def test_row_criteria_contract(grid):
    net = grid("grid118")
    st = N.Structure.from_network(net)
    ctx = N.BatchContext(st, 2, N.LU_FP64, net.reactive_limits())
    ones = np.ones(net.n)
    with pytest.raises(N.OlfHipError) as e:
        ctx.set_row_criteria((ones, -ones))
    assert e.value.code == N.E_ARGUMENT
    with pytest.raises(N.OlfHipError) as e:
        N.check(N.lib().olfhip_set_row_criteria(ctx.h, N.ptr(ones), None))
    assert e.value.code == N.E_ARGUMENT
    ctx.set_row_criteria((ones, ones))
    ctx.set_row_criteria(None)
