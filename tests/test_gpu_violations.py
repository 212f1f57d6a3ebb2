"""GPU tests of violation detection on N-1 states (device screen + exact host decisions, through the C ABI):

* on the device's own states, the detected list equals the oracle's full OLF scan of those states, and every
  reported limit was among the screen's candidates (the screen may only over-emit);
* a capacity too small for the candidates falls back to a full scan with the same answer;
* end to end, the violation lists of the batched N-1 equal those of the oracle's N-1 (its own FP64 states): same
  contingencies, same (kind, element, side, quantity, limit) in the same order, values within 1e-5 relative; any
  entry whose value is within 1e-6 of its threshold would be reported as ambiguous (states agree to 1e-6)."""
import numpy as np
import pytest

import pyoracle as po
from olfgpu import _native as N
from olfgpu import gridgen
from olfgpu.solver import GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters, GpuNewtonRaphson

pytestmark = pytest.mark.gpu


def _setup(net, linear_solver, count, bridge_fraction, seed):
    st = N.Structure.from_network(net)
    rl = net.reactive_limits()
    x0, tg, md, _ = net.system()
    base = GpuNewtonRaphson(st, rl).run(x0, tg, md)
    assert base.status == "CONVERGED"
    table = gridgen.limit_table(net, base.x)
    ctg = gridgen.sample_contingencies(net, count, bridge_fraction, seed)
    sa = GpuBatchedSecurityAnalysisSolver(st, rl, GpuLoadFlowParameters(linearSolver=linear_solver))
    return st, base, table, ctg, sa


def _by_scenario(v, k):
    out = [[] for _ in range(k)]
    for r in v:
        out[int(r["scenario"])].append((int(r["kind"]), int(r["element"]), int(r["side"]), int(r["quantity"]),
                                        int(r["limit"]), float(r["value"])))
    return out


@pytest.mark.parametrize("linear_solver", ["BATCHED_LU", "BATCHED_LU_FP32_REFINED"])
def test_detection_equals_the_oracle_scan_of_the_device_states(grid, linear_solver):
    net = grid("grid2000")
    print(f"[viol] grid2000 {linear_solver} device states", flush=True)
    st, base, table, ctg, sa = _setup(net, linear_solver, 48, 0.25, 6)
    lt = N.LimitTable(st, table)
    r = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg, limit_table=lt)
    got = _by_scenario(r.violations, len(ctg))
    counts, records = lt.screen(sa._ctx, len(ctg))
    o = po.OracleSystem(net)
    total = 0
    for j, e in enumerate(ctg):
        if r.status[j] != 0:
            assert got[j] == []
            continue
        isl = o.islands([e])
        ref = po.violations(o, r.x[j], table, outages=[e], islanded=isl)
        assert [g[:5] for g in got[j]] == [q[:5] for q in ref], (j, e)
        if ref:
            np.testing.assert_allclose([g[5] for g in got[j]], [q[5] for q in ref], rtol=1e-12)
        cand = set(records[j, :counts[j]].tolist())
        assert counts[j] <= records.shape[1]
        for g in got[j]:
            rec = g[4] if g[0] == 0 else lt.n_limits + 2 * g[1] + (g[0] - 1)
            assert rec in cand
        total += len(ref)
    assert total > 10
    # a capacity of one record: scenarios overflow and are scanned in full, with the same answer
    small = lt.detect(sa._ctx, capacity=1)
    assert _by_scenario(small, len(ctg)) == got


@pytest.mark.parametrize("name, count, bridges", [("grid2000", 48, 0.25), ("grid9241", 40, 0.1)])
def test_violation_lists_match_the_oracle_n1(grid, name, count, bridges):
    net = grid(name)
    print(f"[viol] {name} N-1 vs oracle", flush=True)
    st, base, table, ctg, sa = _setup(net, "BATCHED_LU_FP32_REFINED", count, bridges, 8)
    lt = N.LimitTable(st, table)
    r = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg, limit_table=lt)
    got = _by_scenario(r.violations, len(ctg))
    o = po.OracleSystem(net)
    ref = o.n1(base.x, base.row_mode, base.target, base.reactive_limit_types, ctg, threads=0)
    assert np.array_equal(ref["status"], r.status)
    total = ambiguous = 0
    for j, e in enumerate(ctg):
        if ref["status"][j] != 0:
            continue
        want = po.violations(o, ref["x"][j], table, outages=[e], islanded=o.islands([e]))
        near = [q for q in want if abs(abs(q[5]) - q[6]) <= 1e-6 * abs(q[6])]
        ambiguous += len(near)
        assert [g[:5] for g in got[j]] == [q[:5] for q in want], (j, e, near)
        if want:
            np.testing.assert_allclose([g[5] for g in got[j]], [q[5] for q in want], rtol=1e-5)
        total += len(want)
    print(f"[viol] {name}: {total} violations over {len(ctg)} contingencies, {ambiguous} within 1e-6 of a limit",
          flush=True)
    assert total > 10


def _records(rows, scenario):
    return np.array([(scenario, q[0], q[1], q[2], q[3], q[4], q[5], q[6]) for q in rows], N.VIOLATION_DTYPE)


def test_contingency_lists_against_the_base_case(grid):
    """OLF's post-contingency lists (its LimitViolationManager against the base case's violations): the base case's
    violations, detected on the device's base state, equal the oracle's scan of that state; solve_batch_n1 given them
    returns violations.manage of its full scan (in the caller's order, also when it solves longest first); and each
    contingency's list equals manage of the oracle's scan of the device's state against the oracle's base list, under
    the thresholds recommended in docs/increased-violations-thresholds.md:60-71 (0.5 kV and 0.1 % on voltages, 0.1 %
    on flows, 400 kV buses). Pre-existing violations that did not grow are dropped."""
    from olfgpu.violations import IncreasedViolationsParameters, manage

    net = grid("grid2000")
    print("[viol] grid2000 post-contingency lists against the base case", flush=True)
    st = N.Structure.from_network(net)
    rl = net.reactive_limits()
    x0, tg, md, _ = net.system()
    solver = GpuNewtonRaphson(st, rl)
    base = solver.run(x0, tg, md)
    assert base.status == "CONVERGED"
    table = dict(gridgen.limit_table(net, base.x), bus_nominal_kv=400.0)
    lt = N.LimitTable(st, table)
    o = po.OracleSystem(net)
    raw_base = solver.detect_violations(lt, managed=False)
    want_base = po.violations(o, base.x, table)
    assert len(want_base) > 0
    assert [g[:5] for g in _by_scenario(raw_base, 1)[0]] == [q[:5] for q in want_base]
    np.testing.assert_allclose(raw_base["value"], [q[5] for q in want_base], rtol=1e-12)
    bv = solver.detect_violations(lt)
    np.testing.assert_array_equal(bv, manage(raw_base, lt.n_branches, bus_nominal_kv=400.0))
    prm = IncreasedViolationsParameters(flowProportionalThreshold=0.001, lowVoltageAbsoluteThreshold=0.5,
                                        lowVoltageProportionalThreshold=0.001, highVoltageAbsoluteThreshold=0.5,
                                        highVoltageProportionalThreshold=0.001)
    ctg = gridgen.sample_contingencies(net, 48, 0.25, 6)
    sa = GpuBatchedSecurityAnalysisSolver(st, rl, GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"))
    raw = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg, limit_table=lt)
    got = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg, limit_table=lt,
                            base_violations=bv, increased_violations=prm)
    np.testing.assert_array_equal(got.violations, manage(raw.violations, lt.n_branches, bv, prm, 400.0))
    assert len(got.violations) < len(raw.violations)
    longest = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg, limit_table=lt,
                                base_violations=bv, increased_violations=prm, cost=np.arange(len(ctg)) % 5)
    assert _by_scenario(longest.violations, len(ctg)) == _by_scenario(got.violations, len(ctg))
    obase = manage(_records(want_base, 0), lt.n_branches, bus_nominal_kv=400.0)
    per = _by_scenario(got.violations, len(ctg))
    for j, e in enumerate(ctg):
        if got.status[j] != 0:
            assert per[j] == []
            continue
        scan = po.violations(o, got.x[j], table, outages=[e], islanded=o.islands([e]))
        want = manage(_records(scan, j), lt.n_branches, obase, prm, 400.0)
        assert [g[:5] for g in per[j]] == [tuple(int(w[f]) for f in ("kind", "element", "side", "quantity", "limit"))
                                           for w in want], (j, e)
    print(f"[viol] {len(raw.violations)} violations scanned, {len(got.violations)} kept against "
          f"{len(bv)} base violations", flush=True)


def test_voltage_angle_limits_match_the_oracle_scan(grid):
    """Voltage-angle limits (300 across branches of grid2000, bounds within 0.003 rad of the base-case difference):
    the detected lists, angle records last, equal the oracle's scan of the device's own N-1 states, values to 1e-12
    relative; islanded buses and non-converged scenarios report none."""
    net = grid("grid2000")
    print("[viol] grid2000 voltage-angle limits", flush=True)
    st, base, table, ctg, sa = _setup(net, "BATCHED_LU_FP32_REFINED", 48, 0.25, 6)
    table = gridgen.limit_table(net, base.x, angle_limits=300)
    lt = N.LimitTable(st, table)
    assert lt.n_angle_limits == 300
    r = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg, limit_table=lt)
    got = _by_scenario(r.violations, len(ctg))
    o = po.OracleSystem(net)
    n_angle = 0
    for j, e in enumerate(ctg):
        if r.status[j] != 0:
            assert got[j] == []
            continue
        ref = po.violations(o, r.x[j], table, outages=[e], islanded=o.islands([e]))
        assert [g[:5] for g in got[j]] == [q[:5] for q in ref], (j, e)
        if ref:
            np.testing.assert_allclose([g[5] for g in got[j]], [q[5] for q in ref], rtol=1e-12, atol=1e-15)
        n_angle += sum(1 for q in ref if q[0] >= 3)
    print(f"[viol] {n_angle} voltage-angle violations over {len(ctg)} contingencies", flush=True)
    assert n_angle > 0


@pytest.mark.parametrize("name, lanes", [("grid9241", 2), ("grid2000", 1)])
def test_staged_ambiguity_flags_equal_the_gathering_kernel(grid, name, lanes, monkeypatch):
    """The ambiguity check with a workgroup per scenario and its states in LDS (nearLimitsStagedKernel) flags exactly
    the scenarios the one-thread-per-item kernel (nearLimitsKernel, OLFHIP_NEAR_GATHER=1) flags, with islanded buses
    (bridges) and PV buses in the batch, at bands from 1e-6 to 0.1."""
    net = grid(name)
    st, base, table, ctg, _ = _setup(net, "BATCHED_LU_FP32_REFINED", 200, 0.1, 3)
    sa = GpuBatchedSecurityAnalysisSolver(st, net.reactive_limits(),
                                          GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"),
                                          lanes_per_tile=lanes)
    r = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg, states_to_host=False)
    assert (r.status == N.CONVERGED).sum() > 150
    lt = N.LimitTable(st, table)
    counts = []
    for band in (1e-6, 1e-4, 1e-2, 0.1):
        monkeypatch.delenv("OLFHIP_NEAR_GATHER", raising=False)
        staged = lt.ambiguous(sa._ctx, band)
        monkeypatch.setenv("OLFHIP_NEAR_GATHER", "1")
        gathered = lt.ambiguous(sa._ctx, band)
        monkeypatch.delenv("OLFHIP_NEAR_GATHER", raising=False)
        assert np.array_equal(staged, gathered), band
        counts.append(int(staged.sum()))
    print(f"[viol] {name}: ambiguous scenarios at bands 1e-6..0.1: {counts}")
    assert counts == sorted(counts) and counts[3] > 0
