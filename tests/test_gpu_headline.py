"""GPU parity at the workloads' full sizes on the 9,241-bus grid (load x1.1), through the C ABI, against the FP64
oracle on the same inputs:

* 500 single-branch outages drawn uniformly from the non-bridge branches (seed 1), FP64 block LU;
* 2,000 single-branch outages of which 100 bridges (seed 2), FP32 block LU + one FP64-residual refinement;
* the same 2,000 outages from the learned warm start (a 3-layer MLP, hidden 512, trained on 20,000 post-contingency
  states the solver computed at load scalings U(0.9, 1.1), seed 3), solved longest first by the training solves'
  iteration counts, with the base-start fallback.

Bar: per contingency the same status; |V| within 1e-6 pu and angle within 1e-6 rad of the oracle started from the
same state; in FP64 the same iteration counts; and the post-contingency violation lists (OLF's manager against the
base case's violations) equal to those of the oracle's base-start N-1, entry for entry.
"""
import numpy as np
import pytest

import parity
import pyoracle as po
from olfgpu import _native as N
from olfgpu import gridgen
from olfgpu.solver import (GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters, GpuNewtonRaphson,
                           outage_flow_magnitude)
from olfgpu.violations import IncreasedViolationsParameters, manage

pytestmark = pytest.mark.gpu

RECOMMENDED = IncreasedViolationsParameters(
    flowProportionalThreshold=0.001, lowVoltageAbsoluteThreshold=0.5, lowVoltageProportionalThreshold=0.001,
    highVoltageAbsoluteThreshold=0.5, highVoltageProportionalThreshold=0.001)


class Setup:
    pass


@pytest.fixture(scope="module")
def h(grid):
    if N.device_count() == 0:
        pytest.fail("no GPU visible: the gpu-marked tests need an MI355X")
    s = Setup()
    s.net = grid("grid9241")
    s.st = N.Structure.from_network(s.net)
    s.rl = s.net.reactive_limits()
    x0, tg, md, _ = s.net.system()
    s.base_solver = GpuNewtonRaphson(s.st, s.rl)
    s.base = s.base_solver.run(x0, tg, md)
    assert s.base.status == "CONVERGED"
    s.osys = po.OracleSystem(s.net)
    s.table = gridgen.limit_table(s.net, s.base.x)
    s.ltab = N.LimitTable(s.st, s.table)
    s.bv = s.base_solver.detect_violations(s.ltab)
    # the oracle's base list: its own scan of the same base state
    s.obase = manage(parity.scan_records(s.osys, s.base.x, s.table, [], 0, N.VIOLATION_DTYPE), s.ltab.n_branches,
                     bus_nominal_kv=400.0)
    assert np.array_equal(s.bv[["kind", "element", "side", "quantity", "limit"]],
                          manage(parity.scan_records(s.osys, s.base.x, s.table, [], 0, N.VIOLATION_DTYPE),
                                 s.ltab.n_branches)[["kind", "element", "side", "quantity", "limit"]])
    return s


def base_args(h):
    b = h.base
    return b.x, b.row_mode, b.target, b.reactive_limit_types


def oracle_lists(h, ref, ctg):
    scans = parity.n1_scans(h.osys, ref["x"], ref["status"], ctg, h.table, N.VIOLATION_DTYPE)
    return manage(scans, h.ltab.n_branches, h.obase, RECOMMENDED, 400.0), scans


def solve(h, lu, ctg, **kw):
    sa = GpuBatchedSecurityAnalysisSolver(h.st, h.rl, GpuLoadFlowParameters(linearSolver=lu))
    b = h.base
    return sa.solve_batch_n1(b.x, b.target, b.row_mode, b.reactive_limit_types, ctg, limit_table=h.ltab,
                             base_violations=h.bv, increased_violations=RECOMMENDED, **kw)


def check_against_base_start_oracle(h, r, ctg, ref, what, same_iterations):
    st = parity.compare_states(r.x, r.status, ref)
    want, scans = oracle_lists(h, ref, ctg)
    viol = parity.compare_violations(r.violations, want, len(ctg))
    print(f"[{what}] states {st}; violations {viol}; oracle scan entries within 1e-6 of a limit "
          f"{parity.near_limit(scans, 1e-6)}", flush=True)
    assert st["status_mismatches"] == 0, st
    assert st["within_1e-6"] == st["compared"], st
    if same_iterations:
        assert np.array_equal(r.iterations, ref["iterations"])
        assert np.array_equal(r.total_iterations, ref["total_iterations"])
    assert viol["contingencies_differing"] == 0, viol
    assert viol["device_entries"] > 1000
    return st, viol


def test_500_non_bridge_outages_fp64_match_oracle(h):
    ctg = gridgen.sample_contingencies(h.net, 500, 0.0, 1)
    print("[full-K] grid9241 500 non-bridge outages, FP64", flush=True)
    r = solve(h, "BATCHED_LU", ctg)
    ref = h.osys.n1(*base_args(h), ctg)
    check_against_base_start_oracle(h, r, ctg, ref, "500 FP64", same_iterations=True)


def test_2000_outages_fp32_refined_match_oracle(h):
    ctg = gridgen.sample_contingencies(h.net, 2000, 0.05, 2)
    print("[full-K] grid9241 2000 outages (100 bridges), FP32 + refinement", flush=True)
    r = solve(h, "BATCHED_LU_FP32_REFINED", ctg)
    assert (r.islanded_count > 0).sum() == 100
    ref = h.osys.n1(*base_args(h), ctg)
    check_against_base_start_oracle(h, r, ctg, ref, "2000 FP32", same_iterations=False)


@pytest.fixture(scope="module")
def headline(h):
    """The headline workload solved through the product path: the warm start trained on 20,000 solved contingencies,
    the batch solved longest first, violation lists managed against the base case."""
    from olfgpu import warmstart as W

    net = h.net
    ctg = gridgen.sample_contingencies(net, 2000, 0.05, 2)
    print("[headline] training the warm start on 20,000 solved contingencies", flush=True)
    mlp = W.WarmStartMLP(net, hidden=512, seed=0)
    data = W.make_training_set(net, h.st, mlp.hoods, n_samples=20000, levels=20, seed=3)
    assert len(data["branch"]) >= 19000
    hist = W.train(mlp, data, epochs=150, batch=128, lr=2e-3, seed=0)
    print(f"[headline] {len(data['branch'])} samples, loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    b = h.base
    feats = W.endpoint_features(net, b.x, ctg, 1.0, gridgen.bridges(net))
    cost = W.iteration_costs(data, net.n_branch)[ctg]
    tie = outage_flow_magnitude(net, b.x, ctg)
    r = solve(h, "BATCHED_LU_FP32_REFINED", ctg, warm_start=mlp, warm_start_features=feats, cost=cost, cost_tie=tie)
    x0 = mlp.warm_starts_host(b.x, ctg, feats)
    ref = h.osys.n1(*base_args(h), ctg)
    return ctg, r, x0, ref


def test_2000_outages_learned_warm_start_match_oracle_from_the_same_start(h, headline):
    """Each warm-started answer equals the oracle's Newton solve from the same start (the device's warm start, read
    back); each answer that fell back to the base start equals the oracle's base-start solve; statuses equal those
    of the oracle's base-start N-1; and the warm start cuts the Newton iterations of the base-case start."""
    ctg, r, x0, ref = headline
    warm = np.nonzero(r.warm_started)[0]
    print(f"[headline] {warm.size} warm-started answers, {r.warm_start_fallbacks} base-start fallbacks", flush=True)
    ref_same = h.osys.n1(*base_args(h), ctg, x0=x0)
    same = parity.compare_states(r.x, r.status, ref_same, warm)
    print(f"[headline] same start: {same}", flush=True)
    assert same["status_mismatches"] == 0 and same["within_1e-6"] == same["compared"] == warm.size, same
    fb = np.nonzero(~r.warm_started)[0]
    if fb.size:
        back = parity.compare_states(r.x, r.status, ref, fb)
        assert back["status_mismatches"] == 0 and back["within_1e-6"] == back["compared"], back
    st = parity.compare_states(r.x, r.status, ref)
    print(f"[headline] against the base-start oracle: states {st}", flush=True)
    assert st["status_mismatches"] == 0, st
    cold_total = ref["total_iterations"].sum()
    print(f"[headline] Newton iterations: warm {r.total_iterations.sum()} (base-start confirmations of {r.confirmed} "
          f"included) vs base start {cold_total}", flush=True)
    assert r.total_iterations.sum() < 0.9 * cold_total


def test_2000_outages_learned_warm_start_violation_lists_match_oracle(h, headline):
    """The post-contingency lists equal the base-start oracle's entry for entry. A warm-started answer is another
    iterate inside the 1e-4 pu tolerance, so the product solves again from the base start every contingency whose
    warm-started answer has a monitored quantity within 1e-4 (relative) of a threshold (solve_batch_n1's
    confirm_band); those answers then equal the oracle's base-start ones to 1e-6."""
    ctg, r, x0, ref = headline
    want, scans = oracle_lists(h, ref, ctg)
    viol = parity.compare_violations(r.violations, want, len(ctg))
    print(f"[headline] {r.confirmed} answers confirmed from the base start; violations against the base-start "
          f"oracle: {viol}", flush=True)
    assert 0 < r.confirmed < 0.2 * len(ctg)
    assert viol["contingencies_differing"] == 0, viol
    confirmed = np.nonzero(~r.warm_started & (r.status == 0))[0]
    st = parity.compare_states(r.x, r.status, ref, confirmed)
    assert st["within_1e-6"] == st["compared"] == confirmed.size, st
