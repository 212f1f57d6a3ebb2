"""Benchmark: N-1 contingencies/s at 1e-4 pu on the 9,241-bus synthetic grid (load x1.1).

One step = one batched N-1 solve of the whole contingency set (2,000 single-branch outages, 5 % of them bridges
that island buses) from the converged base case, FP32 block LU + one FP64-residual refinement step, reactive-limit
outer loop on, states left on the device (inputs resident; the per-call upload is a few hundred KB). The plain FP64
one-case-at-a-time GPU Newton-Raphson (the same library at batch 1) is timed on a sample beside it, and the CPU
oracle (FP64, OpenMP over contingencies) on a bounded sample.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "olf-gpu_amd"))

from olfgpu import _native as N  # noqa: E402
from olfgpu import gridgen  # noqa: E402
from olfgpu.violations import IncreasedViolationsParameters  # noqa: E402
from olfgpu.solver import (GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters,  # noqa: E402
                           GpuNewtonRaphson, NewtonRaphsonParameters)

METRIC = "N-1 contingencies/sec @1e-4pu, 9241-bus: batched FP32-LU+warm-start vs plain FP64"
HBM_PEAK_GBS = 8000.0


def device_synchronize():
    """hipDeviceSynchronize (what torch.cuda.synchronize() does), without importing torch."""
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def algorithmic_bytes(stats, n, mode, lane_iterations, solves):
    """Algorithmic HBM bytes of the Newton steps: per lane-iteration (fill, norm, factorization, solves, update and,
    in the FP32 mode, the FP64-residual refinement), per the per-unit model of the hot path, with L = LU nonzeros of
    the symbolic plan and m = pattern nonzeros; plus one mismatch fill per ended solve (its convergence test)."""
    m = stats["pattern_nnz"]
    L = stats["lu_nnz"]
    w = 4 if mode == N.LU_FP32_REFINED else 8
    fill = 8 * n + w * m + 8 * n
    norm = 8 * n
    factor = w * m + w * L
    solve = w * L + 32 * n
    update = 24 * n
    refine = (8 * m + 24 * n + w * L + 32 * n + 16 * n) if mode == N.LU_FP32_REFINED else 0
    per_iter = fill + norm + factor + solve + update + refine
    return lane_iterations * per_iter + solves * (16 * n), per_iter


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--contingencies", type=int, default=2000)
    ap.add_argument("--plain-sample", type=int, default=20)
    ap.add_argument("--cpu-sample", type=int, default=2000)
    ap.add_argument("--cpu-sample-1t", type=int, default=60, help="contingencies for the one-thread CPU figure")
    ap.add_argument("--no-violations", action="store_true", help="skip the violation-detection measurement")
    ap.add_argument("--robust-sample", type=int, default=64,
                    help="contingencies of the robust-retry measurement on the grid with series-resonant paths (0: off)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-verify", action="store_true", help="skip the per-contingency comparison with the oracle")
    ap.add_argument("--lanes-per-tile", type=int, default=0)
    ap.add_argument("--threads-per-tile", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true", help="skip the tile-shape autotuner (use the default or "
                    "the shape given)")
    ap.add_argument("--no-configs", action="store_true", help="skip the same-run lines of the other workloads")
    ap.add_argument("--no-warm-start", action="store_true", help="base-case start only (no learned warm start)")
    ap.add_argument("--train-samples", type=int, default=20000)
    ap.add_argument("--train-epochs", type=int, default=150)
    ap.add_argument("--dump-schedule", default=None, help="write the solve order's costs and the timed solve's "
                    "per-contingency iterations to this .npz")
    ap.add_argument("--order", choices=("warm", "cost", "given"), default="warm",
                    help="solve order: longest first by the iterations the training set's outages take from the "
                         "trained network's warm start (ties: from the base start, then base flow), by the base-start "
                         "iterations alone, or as sampled")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the N-1 batch shards by contingency across GPUs "
                                   "(see the sharding design); only --gpus 1 is measured"}))
        sys.exit(2)
    # the copy program runs first, as a child process, before this process touches the GPU
    copy_gbs = measured_copy_bandwidth()
    if N.device_count() == 0:
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)

    t_setup = time.perf_counter()
    net = gridgen.workload("grid9241")
    structure = N.Structure.from_network(net)
    stats = structure.statistics()
    rl = net.reactive_limits()
    x0, tgt, mode, _ = net.system()
    # base case on the device (batch of one, FP64)
    base_solver = GpuNewtonRaphson(structure, rl)
    base = base_solver.run(x0, tgt, mode)
    if base.status != "CONVERGED":
        print(json.dumps({"error": f"base case {base.status}"}))
        sys.exit(1)
    # contingencies start from the base case's converged state, row modes, targets and reactive-limit types
    xb, mb, tb, lt = base.x, base.row_mode, base.target, base.reactive_limit_types

    ctg_given = gridgen.sample_contingencies(net, args.contingencies, 0.05, 2)
    nrp = NewtonRaphsonParameters()
    sa = GpuBatchedSecurityAnalysisSolver(structure, rl, GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"),
                                          nrp, lanes_per_tile=args.lanes_per_tile,
                                          threads_per_tile=args.threads_per_tile)

    # learned warm start: trained on post-contingency states the solver itself computes (load scalings U(0.9, 1.1)),
    # before the timed region
    mlp = None
    warm_info = None
    data = None
    if not args.no_warm_start:
        from olfgpu import warmstart as W
        tw = time.perf_counter()
        mlp = W.WarmStartMLP(net, hidden=512, seed=0)
        data = W.make_training_set(net, structure, mlp.hoods, n_samples=args.train_samples, levels=20, seed=3)
        tgen = time.perf_counter() - tw
        hist = W.train(mlp, data, epochs=args.train_epochs, batch=128, lr=2e-3, seed=0)
        warm_info = {"training_samples": int(len(data["branch"])), "generation_s": round(tgen, 1),
                     "training_s": round(time.perf_counter() - tw - tgen, 1), "epochs": args.train_epochs,
                     "loss_first_last": [round(hist[0], 4), round(hist[-1], 4)], "hidden": 512,
                     "neighbourhood_buses": mlp.n_local}
    # solve order: longest first by the Newton iterations each outage took in the training solves (at other load
    # scalings); results stay per contingency, the order only decides which scenarios share a tile and start first
    order = np.arange(len(ctg_given))
    warm_cost = None
    if data is not None and args.order in ("cost", "warm"):
        from olfgpu import warmstart as W
        from olfgpu.solver import longest_first, outage_flow_magnitude
        base_cost = W.iteration_costs(data, net.n_branch)[ctg_given]
        flow = outage_flow_magnitude(net, xb, ctg_given)
        sched_cost = base_cost
        if args.order == "warm":
            tc = time.perf_counter()
            warm_cost = W.warm_iteration_costs(net, structure, mlp, data)[ctg_given]
            warm_info["warm_cost_s"] = round(time.perf_counter() - tc, 1)
            # warm-start iterations (means of one or two solves: multiples of 0.5), the base-start ones only breaking
            # ties, then the flow the outage redistributes
            sched_cost = warm_cost + 1e-3 * base_cost
        order = longest_first(sched_cost, flow)
    ctg = ctg_given[order]
    K = len(ctg)
    rme, tix, tvl, isl = sa.plan_islands(ctg[:, None])
    inputs, keep = N.BatchContext.make_inputs(
        K, xb, tb, mb, row_mode_entries=rme, target_overrides=(tix, tvl), outaged_branches=ctg[:, None],
        reactive_limits_seed_state=lt, max_iter=nrp.maxIterations, tol=nrp.maxMismatch,
        reactive_limit_tol=nrp.reactive_limit_band(), states_to_host=False)
    feats = None
    if mlp is not None:
        from olfgpu import warmstart as W
        feats = W.endpoint_features(net, xb, ctg, 1.0, gridgen.bridges(net))
    # the project's autotuner picks the tile shape on this very workload (the warm starts included), before the
    # timed region, unless a shape is given
    tuned = None
    if not args.no_autotune and not (args.lanes_per_tile or args.threads_per_tile):
        if mlp is not None:
            inputs.x0_per_scenario_device = mlp.predict_warm_start(xb, ctg, feats)
        ta = time.perf_counter()
        tuned = N.autotune(structure, inputs, N.LU_FP32_REFINED, rl)
        tuned["seconds"] = round(time.perf_counter() - ta, 2)
        print(json.dumps({"autotune": tuned}), file=sys.stderr, flush=True)
        sa.lanes_per_tile = tuned["lanes_per_tile"]
        sa.threads_per_tile = tuned["threads_per_tile"]
    ctx = sa.context(K)
    setup_s = time.perf_counter() - t_setup

    def one_step():
        """One batched N-1 of the whole set: warm starts (network inference on the device), the solve, and a
        base-start solve of any warm-started scenario that did not converge."""
        ms_total, launches = 0.0, 0
        if mlp is not None:
            inputs.x0_per_scenario_device = mlp.predict_warm_start(xb, ctg, feats)
        r = ctx.solve(inputs, want_states=False)
        ms, st = ctx.timing()
        ms_total += ms
        launches += st
        timed_iterations = int(r["total_iterations"].sum())
        fallbacks = 0
        if mlp is not None:
            redo = np.nonzero(r["status"] != 0)[0]
            if redo.size:
                fallbacks = int(redo.size)
                rr = sa.solve_batch_n1(xb, tb, mb, lt, ctg[redo], states_to_host=False)
                r["status"][redo] = rr.status
                r["total_iterations"][redo] += rr.total_iterations
        return r, ms_total, launches, fallbacks, timed_iterations

    for _ in range(args.warmup):
        one_step()
    device_synchronize()
    t0 = time.perf_counter()
    step_ms = 0.0
    steps_launched = 0
    last = None
    fallbacks = 0
    timed_iters = 0
    for _ in range(args.steps):
        last, ms, st, fb, ti = one_step()
        step_ms += ms
        steps_launched += st
        fallbacks += fb
        timed_iters += ti
    device_synchronize()
    wall = time.perf_counter() - t0
    value = K * args.steps / wall
    status = last["status"]
    lane_iters = int(last["total_iterations"].sum())
    if args.dump_schedule:
        np.savez(args.dump_schedule, ctg=ctg, total_iterations=last["total_iterations"],
                 cost=(W.iteration_costs(data, net.n_branch)[ctg] if data is not None else np.zeros(len(ctg))))
    if warm_info is not None:
        warm_info["fallbacks_per_step"] = fallbacks / args.steps

    # plain FP64 one-case-at-a-time GPU NR on a sample (the same library at batch 1, FP64, base-case start)
    plain = GpuBatchedSecurityAnalysisSolver(structure, rl, GpuLoadFlowParameters(linearSolver="BATCHED_LU"), nrp)
    sample = ctg_given[: args.plain_sample]
    prme, ptix, ptvl, _ = plain.plan_islands(sample[:, None])
    pctx = plain.context(1)
    plain_inputs = []
    for s in range(len(sample)):
        sel = prme[:, 0] == s
        r1 = prme[sel].copy()
        r1[:, 0] = 0
        tsel = (ptix // net.n) == s
        pin, pk = N.BatchContext.make_inputs(
            1, xb, tb, mb, row_mode_entries=r1, target_overrides=(ptix[tsel] - s * net.n, ptvl[tsel]),
            outaged_branches=sample[s:s + 1, None], reactive_limits_seed_state=lt[None, :],
            max_iter=nrp.maxIterations, tol=nrp.maxMismatch, reactive_limit_tol=nrp.reactive_limit_band(),
            states_to_host=False)
        plain_inputs.append((pin, pk))
    pctx.solve(plain_inputs[0][0], want_states=False)   # warm-up
    device_synchronize()
    tp = time.perf_counter()
    plain_status = []
    for pin, _ in plain_inputs:
        plain_status.append(int(pctx.solve(pin, want_states=False)["status"][0]))
    device_synchronize()
    plain_s = time.perf_counter() - tp
    plain_value = len(sample) / plain_s
    status_given = np.empty_like(status)
    status_given[order] = status
    plain_match = bool(np.array_equal(np.array(plain_status), status_given[: len(sample)]))

    # violation detection on the last step's states (reported beside the value, not inside it: the metric is the
    # Newton solve, as for the plain FP64 arm)
    viol_info = None
    table = gridgen.limit_table(net, xb)
    ltab = N.LimitTable(structure, table)
    # OLF's post-contingency lists: its manager against the base case's violations, under the increased-violation
    # thresholds of docs/increased-violations-thresholds.md:60-71 (buses taken at 400 kV: the workload gives no
    # nominal voltages)
    bv = base_solver.detect_violations(ltab)
    recommended = IncreasedViolationsParameters(
        flowProportionalThreshold=0.001, lowVoltageAbsoluteThreshold=0.5, lowVoltageProportionalThreshold=0.001,
        highVoltageAbsoluteThreshold=0.5, highVoltageProportionalThreshold=0.001)
    cost_kw = {}
    feats_given = None
    if data is not None and args.order in ("cost", "warm"):
        cost_kw = dict(cost=sched_cost, cost_tie=flow)
    if mlp is not None:
        feats_given = W.endpoint_features(net, xb, ctg_given, 1.0, gridgen.bridges(net))

    def n1_with_violations(states_to_host):
        """The product's whole N-1 with its post-contingency violation lists (solve_batch_n1: warm starts, the
        solve, base-start solves of the warm-started answers that did not converge or whose limit decisions sit
        within the confirmation band of a threshold, device screen + exact host decisions, OLF's manager against
        the base case)."""
        return sa.solve_batch_n1(xb, tb, mb, lt, ctg_given, states_to_host=states_to_host, warm_start=mlp,
                                 warm_start_features=feats_given, limit_table=ltab, base_violations=bv,
                                 increased_violations=recommended, **cost_kw)

    if not args.no_violations:
        # the detection alone on the timed step's states: screen + exact host decisions
        if mlp is not None:
            inputs.x0_per_scenario_device = mlp.predict_warm_start(xb, ctg, feats)
        ctx.solve(inputs, want_states=False)
        v = ltab.detect(ctx)
        device_synchronize()
        tv = time.perf_counter()
        for _ in range(3):
            v = ltab.detect(ctx)
        dv = (time.perf_counter() - tv) / 3
        # the whole N-1 with its lists
        n1_with_violations(False)
        device_synchronize()
        tn = time.perf_counter()
        for _ in range(3):
            full = n1_with_violations(False)
        device_synchronize()
        dn = (time.perf_counter() - tn) / 3
        kept = full.violations
        viol_info = {"detection_ms": round(dv * 1e3, 2), "violations_scanned": int(len(v)),
                     "by_kind_scanned": {N.VIOLATION_KINDS[k]: int((v["kind"] == k).sum()) for k in N.VIOLATION_KINDS},
                     "monitored_groups": int(len(table["group_side"])), "screen_band": 1e-7,
                     "n1_with_lists": {"ms_per_step": round(dn * 1e3, 2), "value": round(K / dn, 2),
                                       "unit": "contingencies/s",
                                       "confirmed_from_base_start": int(full.confirmed),
                                       "confirm_band": 1e-4,
                                       "base_violations": int(len(bv)), "kept": int(len(kept)),
                                       "by_kind": {N.VIOLATION_KINDS[k]: int((kept["kind"] == k).sum())
                                                   for k in N.VIOLATION_KINDS},
                                       "thresholds": "flow 0.1 %, voltage min(0.5 kV, 0.1 %), buses at 400 kV"}}

    robust_info = robust_retry_measure(net, nrp, args.robust_sample) if args.robust_sample > 0 else None
    configs_info = None if args.no_configs else other_configs(net, structure, rl, xb, mb, tb, lt, nrp, ctg_given,
                                                               plain_value)

    # roofline of the Newton-step kernel: the event-timed launches of the batched solve against the lane iterations
    # they ran (the fallback re-solves run in another context and are in neither figure)
    ab, per_iter = algorithmic_bytes(stats, net.n, N.LU_FP32_REFINED, timed_iters / args.steps, K)
    step_ms_per_solve = step_ms / args.steps
    launches_per_solve = steps_launched / args.steps
    achieved = ab / (step_ms_per_solve * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "nrStepKernel<T,float>", "launches_per_step": launches_per_solve,
                "avg_launch_ms": round(step_ms_per_solve / max(launches_per_solve, 1), 4),
                "algorithmic_bytes_per_launch": int(ab / max(launches_per_solve, 1)),
                "bytes_per_lane_iteration": int(per_iter),
                "lane_iterations_per_launch_group": round(timed_iters / args.steps, 1)}
    if copy_gbs:
        roofline["measured_copy_gbs"] = copy_gbs
        roofline["frac_of_measured_copy"] = round(achieved / copy_gbs, 4)
    traffic = committed_traffic()
    if traffic:
        roofline["traffic"] = traffic["bytes_per_launch"]
        roofline["traffic_source"] = traffic["source"]

    # the product's answer for every contingency, outside the timed region: the same batch through
    # solve_batch_n1 (same order, tiles, warm starts and fallbacks as the timed steps), states and post-contingency
    # violation lists to the host, for the per-contingency comparison with the oracle
    verify = None
    if not args.no_verify:
        ver = n1_with_violations(True)
        same = np.ones(K, bool) if ver.warm_started is None else ver.warm_started
        tot_given = np.empty_like(last["total_iterations"])
        tot_given[order] = last["total_iterations"]
        verify = {"result": ver, "table": table, "recommended": recommended,
                  "x0": None if mlp is None else mlp.warm_starts_host(xb, ctg_given, feats_given),
                  "matches_timed_step": bool(np.array_equal(ver.status, status_given)
                                             and np.array_equal(ver.total_iterations[same], tot_given[same])),
                  "confirmed_from_base_start": int(ver.confirmed)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(net, xb, mb, tb, lt, ctg_given[: args.cpu_sample], ctg_given[: args.cpu_sample_1t],
                           ctg_given, verify)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "contingencies/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32-lu+fp64-residual", "data": "synthetic",
        "config": {"workload": "grid9241 x1.1 load, 2000 single-branch outages (5% bridges, seed 2), FP32 block LU + "
                               "1 FP64 refinement, reactive limits on, " +
                               ("base-case start" if mlp is None else "learned warm start (base-start fallback)") +
                               ({"cost": ", solved longest first by the training solves' iteration counts (ties: "
                                         "larger base flow of the outaged branch first)",
                                 "warm": ", solved longest first by the iterations the training set's outages take "
                                         "from the trained warm start (ties: base-start iterations, then larger base "
                                         "flow of the outaged branch)"}.get(args.order, "")
                                if data is not None else ""),
                   "buses": net.n_bus, "branches": net.n_branch, "contingencies": K, "tol_pu": nrp.maxMismatch,
                   "tile": ctx_lanes(ctx), "lu_levels": stats["levels"], "lu_nnz": stats["lu_nnz"]},
        "statuses": {N.STATUS_NAMES[int(s)]: int((status == s).sum()) for s in np.unique(status)},
        "newton_lane_iterations_per_step": lane_iters,
        "warm_start": warm_info,
        "plain_fp64_gpu": {"value": round(plain_value, 2), "unit": "contingencies/s", "sample": len(sample),
                           "statuses_match_batched": plain_match},
        "speedup_vs_plain_fp64_gpu": round(value / plain_value, 2),
        "autotune": tuned,
        "graph": dict(zip(("instantiations", "in_place_updates"), ctx.graph_statistics())),
        "other_configs": configs_info,
        "violation_detection": viol_info,
        "robust_retry": robust_info,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "oracle_error": None if cpu is None else cpu.pop("oracle_error", None),
        "setup_s": round(setup_s, 2),
    }
    print(json.dumps(out), flush=True)


def other_configs(net, structure, rl, xb, mb, tb, lt, nrp, ctg_c4, plain_value, steps=5):
    """Same-run lines of the workload's other configurations, each timed over `steps` solves after one warm-up:
    the batching-only step (500 single-branch outages drawn from the non-bridge branches, seed 1, FP64 block LU,
    base-case start), the FP32 step without the warm start (the headline's 2,000 outages, FP32 LU + one FP64
    refinement, base-case start, longest first is not available without the training solves: given order), and the
    batch-of-one load flow of the 2,000-bus grid from a flat start with the reactive-limit loop (the AcSolver
    drop-in)."""
    out = {}

    def rate(sa, ctg):
        sa.solve_batch_n1(xb, tb, mb, lt, ctg, states_to_host=False)
        device_synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            r = sa.solve_batch_n1(xb, tb, mb, lt, ctg, states_to_host=False)
        device_synchronize()
        dt = (time.perf_counter() - t) / steps
        return r, dt

    c3 = gridgen.sample_contingencies(net, 500, 0.0, 1)
    sa = GpuBatchedSecurityAnalysisSolver(structure, rl, GpuLoadFlowParameters(linearSolver="BATCHED_LU"), nrp)
    r, dt = rate(sa, c3)
    out["batched_fp64_500_non_bridge"] = {
        "value": round(len(c3) / dt, 2), "unit": "contingencies/s", "ms_per_step": round(dt * 1e3, 3),
        "speedup_vs_plain_fp64_gpu": round(len(c3) / dt / plain_value, 2),
        "statuses": {N.STATUS_NAMES[int(s)]: int((r.status == s).sum()) for s in np.unique(r.status)},
        "lane_iterations": int(r.total_iterations.sum())}
    sa = GpuBatchedSecurityAnalysisSolver(structure, rl, GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"),
                                          nrp)
    r, dt = rate(sa, ctg_c4)
    out["batched_fp32_refined_2000_base_start"] = {
        "value": round(len(ctg_c4) / dt, 2), "unit": "contingencies/s", "ms_per_step": round(dt * 1e3, 3),
        "speedup_vs_plain_fp64_gpu": round(len(ctg_c4) / dt / plain_value, 2),
        "statuses": {N.STATUS_NAMES[int(s)]: int((r.status == s).sum()) for s in np.unique(r.status)},
        "lane_iterations": int(r.total_iterations.sum())}
    g2 = gridgen.workload("grid2000")
    st2 = N.Structure.from_network(g2)
    x0, tg, md, _ = g2.system()
    nr = GpuNewtonRaphson(st2, g2.reactive_limits())
    nr.run(x0, tg, md)
    device_synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        r1 = nr.run(x0, tg, md)
    dt = (time.perf_counter() - t) / steps
    out["acsolver_base_case_2000_bus"] = {"ms_per_solve": round(dt * 1e3, 3), "status": r1.status,
                                          "iterations": r1.total_iterations,
                                          "note": "flat start, reactive-limit loop, FP64, batch of one, warm context"}
    return out


def robust_retry_measure(net, nrp, k):
    """The retry of rejected scenarios in the robust mode, on the same grid with two series-resonant paths beside
    two of its branches (a load-free bus between reactances x and -x, whose 2x2 Jacobian block is singular): the
    batch (FP32 block LU, no pivoting across buses) rejects the scenarios it cannot factor, and the robust mode
    (FP64 static pivoting + GMRES, one scenario per workgroup) solves them again. Reported beside the value: the
    metric's workload needs no pivoting."""
    deg = np.bincount(np.concatenate([net.bus1, net.bus2]), minlength=net.n_bus)
    slack = set(np.nonzero(net.bus_type == gridgen.SLACK)[0].tolist())
    beside = [e for e in range(net.n_branch) if deg[net.bus1[e]] >= 3 and deg[net.bus2[e]] >= 3
              and int(net.bus1[e]) not in slack and int(net.bus2[e]) not in slack][:2]
    rnet = net.with_series_resonance(beside, x=0.05)
    rst = N.Structure.from_network(rnet)
    rrl = rnet.reactive_limits()
    x0, tgt, md, _ = rnet.system()
    nr1 = GpuNewtonRaphson(rst, rrl)
    b = nr1.run(x0, tgt, md)
    info = {"network": f"grid9241 + series-resonant paths beside branches {beside}", "contingencies": k,
            "base_case": b.status, "base_case_in_robust_mode": nr1.robust_mode}
    if b.status != "CONVERGED":
        return info
    ctg = gridgen.sample_contingencies(rnet, k, 0.05, 4)
    lu = GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED")
    first = GpuBatchedSecurityAnalysisSolver(rst, rrl, lu, nrp, robust_retry=False)
    full = GpuBatchedSecurityAnalysisSolver(rst, rrl, lu, nrp)
    args = (b.x, b.target, b.row_mode, b.reactive_limit_types, ctg)
    first.solve_batch_n1(*args, states_to_host=False)   # warm-up (contexts, code objects)
    full.solve_batch_n1(*args, states_to_host=False)
    device_synchronize()
    t0 = time.perf_counter()
    r1 = first.solve_batch_n1(*args, states_to_host=False)
    device_synchronize()
    t1 = time.perf_counter()
    r2 = full.solve_batch_n1(*args, states_to_host=False)
    device_synchronize()
    t2 = time.perf_counter()
    rejected = int(np.isin(r1.status, (N.REJECTED, N.LINEAR_SOLVE_FAILED)).sum())
    retry_ms = ((t2 - t1) - (t1 - t0)) * 1e3
    info.update({"rejected_by_batch": rejected, "retried": r2.robust_attempted, "recovered": r2.robust_recovered,
                 "statuses": {N.STATUS_NAMES[int(s)]: int((r2.status == s).sum()) for s in np.unique(r2.status)},
                 "batch_ms": round((t1 - t0) * 1e3, 2), "batch_with_retry_ms": round((t2 - t1) * 1e3, 2),
                 "retry_ms_per_scenario": round(retry_ms / max(r2.robust_attempted, 1), 3)})
    return info


def ctx_lanes(ctx):
    return {"lanes_per_tile": ctx._tuning.lanes_per_tile or "default", "threads_per_tile":
            ctx._tuning.threads_per_tile or "default"}


def cpu_baseline(net, xb, mb, tb, lt, sample, sample_1t, ctg_all, verify):
    """The FP64 CPU oracle (own sparse LU, OpenMP over contingencies) on a bounded sample, all host cores, and on a
    smaller sample with one thread. With the device's verification solve (verify), the oracle is also the checker of
    every contingency's answer (oracle_error)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as po

    osys = po.OracleSystem(net)
    t = time.perf_counter()
    r = osys.n1(xb, mb, tb, lt, sample)
    dt = time.perf_counter() - t
    t = time.perf_counter()
    osys.n1(xb, mb, tb, lt, sample_1t, threads=1)
    d1 = time.perf_counter() - t
    out = {"value": round(len(sample) / dt, 2), "unit": "contingencies/s", "cores": int(r["threads"]),
           "cpu_model": host_cpu_model(), "kind": "port", "label": "CPU restatement of OLF NR (not OLF)",
           "sample": f"first {len(sample)} of the 2000 contingencies, same base state and rules, "
                     f"{dt:.1f} s wall on {int(r['threads'])} threads",
           "one_thread_value": round(len(sample_1t) / d1, 2),
           "one_thread_sample": f"first {len(sample_1t)} contingencies, {d1:.1f} s on 1 thread"}
    if verify is not None:
        out["oracle_error"] = oracle_error(po, osys, net, xb, mb, tb, lt, ctg_all,
                                           r if len(sample) == len(ctg_all) else None, verify)
    return out


def oracle_error(po, osys, net, xb, mb, tb, lt, ctg, ref_base, verify):
    """Every contingency's device answer against the FP64 oracle: (1) the oracle from the same start (the warm start
    the device used, or the base state where the answer is a base-start solve): statuses, |V| and angle;
    (2) the oracle from the base-case start, the reference semantics: statuses, states, and the post-contingency
    violation lists (the oracle's own scan of its own states through OLF's manager against the oracle's scan of the
    base state), entry for entry."""
    import parity
    from olfgpu.violations import manage

    ver = verify["result"]
    k = len(ctg)
    if ref_base is None:
        ref_base = osys.n1(xb, mb, tb, lt, ctg)
    out = {"contingencies": k, "verification_solve_matches_timed_step": verify["matches_timed_step"],
           "confirmed_from_base_start": verify["confirmed_from_base_start"]}
    if verify["x0"] is not None:
        t = time.perf_counter()
        ref_same = osys.n1(xb, mb, tb, lt, ctg, x0=verify["x0"])
        warm = np.asarray(ver.warm_started)
        same = parity.compare_states(ver.x, ver.status, ref_same, np.nonzero(warm)[0])
        same["oracle_s"] = round(time.perf_counter() - t, 2)
        out["same_start"] = same
        out["same_start"]["warm_started_answers"] = int(warm.sum())
        base_rows = np.nonzero(~warm)[0]
        if base_rows.size:
            out["base_start_fallbacks"] = parity.compare_states(ver.x, ver.status, ref_base, base_rows)
    out["vs_base_start_oracle"] = parity.compare_states(ver.x, ver.status, ref_base)
    table = dict(verify["table"])
    obase = manage(parity.scan_records(osys, xb, table, [], 0, N.VIOLATION_DTYPE), len(table["branch"]),
                   bus_nominal_kv=400.0)
    scans = parity.n1_scans(osys, ref_base["x"], ref_base["status"], ctg, table, N.VIOLATION_DTYPE)
    want = manage(scans, len(table["branch"]), obase, verify["recommended"], 400.0)
    viol = parity.compare_violations(ver.violations, want, k)
    viol["oracle_scan_entries"] = int(len(scans))
    viol["oracle_scan_entries_within_1e-6_of_limit"] = parity.near_limit(scans, 1e-6)
    out["violation_lists_vs_base_start_oracle"] = viol
    return out


def host_cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def measured_copy_bandwidth():
    """Best device-to-device copy bandwidth of the plain copy program built beside the library (2 x 1 GiB moved per
    dispatch), GB/s; None when it is not built."""
    import subprocess

    exe = os.path.join(REPO, "olf-gpu_amd", "hbm_copy")
    if not os.path.exists(exe):
        return None
    res = subprocess.run([exe, "1024", "5"], capture_output=True, text=True, timeout=120)
    if res.returncode != 0:
        return None
    try:
        return float(res.stdout.strip().split(",")[-1].split()[0])
    except (ValueError, IndexError):
        return None


def committed_traffic():
    """HBM bytes per Newton-step launch from the committed rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE of the
    step kernel on this workload, scaled by the counters' reading of a copy of known size), if present."""
    path = os.path.join(REPO, "profiles", "nrstep_grid9241_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return {"bytes_per_launch": t["bytes_per_launch"], "source": "profiles/nrstep_grid9241_traffic.json (" +
            t.get("command_summary", "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE") + ")"}


if __name__ == "__main__":
    main()
