"""Host-side mirror of the reference's Java surface for this path, over the C ABI.

* ``GpuLoadFlowParameters``  — the parameter extension (names and defaults of
  src/main/java/com/powsybl/openloadflow/ac/solver/gpu/GpuLoadFlowParameters.java:37-63);
* ``GpuNewtonRaphson``       — the GPU_NEWTON_RAPHSON AcSolver's batch-of-one load flow
  (GpuNewtonRaphson.java:510-602, factory GpuNewtonRaphsonFactory.java:35-86);
* ``GpuBatchedSecurityAnalysisSolver`` — the batched N-1 (GpuBatchedSecurityAnalysisSolver.java:107-373 with
  chunking as GpuBatchChunkSolver.java:101-160 and islands as batch/plan/IslandedBuses.java:50-60).

The numerics all run in libolfhip.so; this layer plans scenarios and marshals arrays, as the Java layer does.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

NAME = "GPU_NEWTON_RAPHSON"          # AcSolverFactory name (GpuNewtonRaphsonFactory.java:35-38)
PROVIDER_NAME = "OpenLoadFlowGpu"    # LoadFlowProvider / SecurityAnalysisProvider name


@dataclass
class GpuLoadFlowParameters:
    """The ``gpu-load-flow-parameters`` extension (GpuLoadFlowParameters.java:39-63): same names and defaults.

    linearSolver selects the device linear solve: BATCHED_LU (FP64 block LU) or BATCHED_LU_FP32_REFINED (FP32 factors
    and solves plus one FP64-residual refinement step). CUDSS, the reference's closed-library alternative, is refused.
    The cuDSS-specific knobs (reorderingAlgorithm, batchedLuOrdering, batchedLuSplitLongRows, matchingAlgorithm,
    pivotType, pivotThreshold) are accepted and carried but have no effect: the block LU always uses its own
    minimum-degree order and static 2x2-block pivots. iterativeRefinementBaseSteps is the number of FP64-residual
    refinement steps of the FP32 mode (1 supported)."""

    NAME = "gpu-load-flow-parameters"
    MODULE = "open-loadflow-gpu-default-parameters"

    linearSolver: str = "BATCHED_LU"
    reorderingAlgorithm: str = "AUTO"
    batchedLuOrdering: str = "CUDSS"
    batchedLuSplitLongRows: bool = True
    matchingAlgorithm: str = "AUTO"
    pivotType: str = "BY_ORDERING"
    pivotThreshold: float = float("nan")
    iterativeRefinementBaseSteps: int = 1
    iterativeRefinementMaxSteps: int = 4
    iterativeRefinementTolerance: float = 1e-12

    LINEAR_SOLVERS = ("BATCHED_LU", "BATCHED_LU_FP32_REFINED")

    def __post_init__(self):
        self.validate()

    def validate(self):
        if self.linearSolver == "CUDSS":
            raise ValueError("linearSolver CUDSS is not available: the MI355X build has its own batched LU")
        if self.linearSolver not in self.LINEAR_SOLVERS:
            raise ValueError(f"unknown linearSolver {self.linearSolver!r}")
        if self.iterativeRefinementBaseSteps not in (0, 1):
            raise ValueError("iterativeRefinementBaseSteps: the FP32 mode runs exactly one refinement step")

    @property
    def lu_mode(self) -> int:
        return N.LU_FP32_REFINED if self.linearSolver == "BATCHED_LU_FP32_REFINED" else N.LU_FP64

    @classmethod
    def load(cls, values: dict) -> "GpuLoadFlowParameters":
        """From a name -> value map (the reference's PlatformConfig / parameter-map route)."""
        p = cls()
        for k, v in values.items():
            if not hasattr(p, k) or k.isupper():
                raise ValueError(f"unknown parameter {k!r}")
            cur = getattr(p, k)
            if isinstance(cur, bool):
                v = v if isinstance(v, bool) else str(v).lower() == "true"
            elif isinstance(cur, int):
                v = int(v)
            elif isinstance(cur, float):
                v = float(v)
            setattr(p, k, v)
        p.validate()
        return p

    def to_json(self) -> str:
        d = {k: getattr(self, k) for k in self.__dataclass_fields__}
        d["pivotThreshold"] = None if math.isnan(self.pivotThreshold) else self.pivotThreshold
        return json.dumps(d)


@dataclass
class PerEquationTypeStoppingCriteria:
    """OLF's PerEquationTypeStoppingCriteria (NewtonRaphsonStoppingCriteriaType.PER_EQUATION_TYPE_CRITERIA), as the
    reference's GpuAcActivationExtractor.extractRowCriteria turns it into per-row thresholds
    (GpuAcActivationExtractor.java:280-325): on a P row maxAngleMismatch when the row is an identity (slack angle)
    and maxActivePowerMismatch / base when it is an equation; on a Q row maxVoltageMismatch (PV voltage) and
    maxReactivePowerMismatch / base. Powers in MW / MVar; base_mva is the system's per-unit base (OLF's is 100 MVA,
    the synthetic grids' gridgen.SB)."""

    maxActivePowerMismatch: float = 1e-2
    maxReactivePowerMismatch: float = 1e-2
    maxVoltageMismatch: float = 1e-4
    maxAngleMismatch: float = 1e-5
    base_mva: float = 400.0

    def row_thresholds(self, n: int):
        """(fixed [n], equation [n]) on the stable rows (2b: P / angle, 2b + 1: Q / voltage)."""
        fixed = np.empty(n)
        eq = np.empty(n)
        fixed[0::2] = self.maxAngleMismatch
        fixed[1::2] = self.maxVoltageMismatch
        eq[0::2] = self.maxActivePowerMismatch / self.base_mva
        eq[1::2] = self.maxReactivePowerMismatch / self.base_mva
        return fixed, eq


@dataclass
class NewtonRaphsonParameters:
    """The OLF Newton-Raphson parameters the AcSolver copies (GpuNewtonRaphsonFactory.java:62-79)."""

    maxIterations: int = 15
    maxMismatch: float = 1e-4                 # ||F||inf stopping threshold, pu; <= 0 (with convEpsPerEq <= 0 and no
                                              # per-equation-type criterion): exactly maxIterations steps, no test
    convEpsPerEq: float = 0.0                 # > 0: OLF's ||F||_2 < sqrt(convEpsPerEq^2 * equations) instead
    stateVectorScalingMode: str = "NONE"      # NONE, LINE_SEARCH or MAX_VOLTAGE_CHANGE
    maxVoltageChangeV: float = 0.1
    maxVoltageChangePhi: float = 0.174533
    lineSearchStateVectorScalingMaxIteration: int = 10
    lineSearchStateVectorScalingStepFold: float = 4.0 / 3.0
    reactiveLimits: bool = True
    maxReactivePowerMismatch: float = 1e-4    # pu: the PV->PQ switch band when no other criterion sets it
    maxOuterLoopIterations: int = 20
    maxPqPvSwitch: int = 3
    # distributed slack (OLF LoadFlowParameters.distributedSlack, OpenLoadFlowParameters.slackBusPMaxMismatch in MW,
    # slackDistributionFailureBehavior); it runs when a solver is given a distributed-slack description
    # (gridgen.distributed_slack) and distributedSlack is set
    distributedSlack: bool = True
    slackBusPMaxMismatch: float = 1.0
    slackDistributionFailureBehavior: str = "LEAVE_ON_SLACK_BUS"
    # OLF's stopping criterion object: None for the uniform ones (maxMismatch, or convEpsPerEq), or a
    # PerEquationTypeStoppingCriteria (per-row thresholds installed on the device context)
    stoppingCriteria: PerEquationTypeStoppingCriteria | None = None
    # the linear-solve guard's bound on ||J dx + F||inf / ||F||inf; 0: the reference's 1e-2
    linearSolveFailRatio: float = 0.0

    def reactive_limit_band(self) -> float:
        """The reactive-limit loop's switch band (pu), derived from the stopping criterion as OLF's
        AbstractAcOuterLoopConfig.createReactiveLimitsOuterLoop does and the reference mirrors
        (GpuSaRunScope.java:339-350, GpuSolveSettings.java:35-41): the per-equation-type criterion's
        maxReactivePowerMismatch / base; under the uniform 2-norm criterion (convEpsPerEq > 0) convEpsPerEq;
        under the plain ||F||inf criterion maxReactivePowerMismatch."""
        sc = self.stoppingCriteria
        if sc is not None:
            return sc.maxReactivePowerMismatch / sc.base_mva
        if self.convEpsPerEq > 0:
            return self.convEpsPerEq
        return self.maxReactivePowerMismatch

    def install_stopping_criteria(self, ctx: N.BatchContext) -> None:
        """GpuNewtonRaphson.installRowCriteria (GpuNewtonRaphson.java:446-459): the per-row thresholds, or none."""
        sc = self.stoppingCriteria
        ctx.set_row_criteria(None if sc is None else sc.row_thresholds(ctx.structure.n))

    def slack_inputs(self, ds: dict | None) -> dict:
        """make_inputs keywords of the distributed-slack loop for a description ds (None: off)."""
        if ds is None or not self.distributedSlack:
            return dict(distributed_slack=False)
        if self.slackDistributionFailureBehavior not in ("LEAVE_ON_SLACK_BUS", "FAIL"):
            raise ValueError(f"slackDistributionFailureBehavior {self.slackDistributionFailureBehavior!r} is not "
                             "modelled (LEAVE_ON_SLACK_BUS or FAIL)")
        return dict(distributed_slack=True, slack_dist_tol=self.slackBusPMaxMismatch / float(ds["base_mva"]),
                    reject_slack_distribution_failure=self.slackDistributionFailureBehavior == "FAIL")

    def scaling_code(self) -> int:
        if self.stateVectorScalingMode == "NONE":
            return N.SCALING_NONE
        if self.stateVectorScalingMode == "MAX_VOLTAGE_CHANGE":
            return N.SCALING_MAX_VOLTAGE_CHANGE
        if self.stateVectorScalingMode == "LINE_SEARCH":
            return N.SCALING_LINE_SEARCH
        raise ValueError(f"state vector scaling {self.stateVectorScalingMode!r} is not modelled")


@dataclass
class AcSolverResult:
    status: str
    iterations: int
    total_iterations: int
    norm: float
    x: np.ndarray
    row_mode: np.ndarray = field(default=None, repr=False)
    target: np.ndarray = field(default=None, repr=False)
    reactive_limit_types: np.ndarray = field(default=None, repr=False)


class GpuNewtonRaphson:
    """Batch-of-one device load flow on a cached structure and context (GpuNewtonRaphson.run/runImpl). A solve whose
    linear solve failed (LINEAR_SOLVE_FAILED: the unpivoted factorization met a singular pivot block, or the guard
    rejected the step) is solved again on the device in the robust mode (N.LU_ROBUST), which the solver then keeps
    for the rest of its life (retryOrRejectOnDevice, GpuNewtonRaphson.java:751-775); robust_retry=False turns that
    off."""

    def __init__(self, structure: N.Structure, reactive_limits: dict | None = None,
                 gpu_parameters: GpuLoadFlowParameters | None = None,
                 nr_parameters: NewtonRaphsonParameters | None = None, distributed_slack: dict | None = None,
                 robust_retry: bool = True):
        self.structure = structure
        self.gpu_parameters = gpu_parameters or GpuLoadFlowParameters()
        self.nr = nr_parameters or NewtonRaphsonParameters()
        self.reactive_limits = reactive_limits
        self.distributed_slack = distributed_slack
        self.robust_retry = robust_retry
        self.robust_mode = False
        self.context = self._make_context(self.gpu_parameters.lu_mode)

    def _make_context(self, mode):
        ctx = N.BatchContext(self.structure, 1, mode, self.reactive_limits)
        if self.distributed_slack is not None:
            ctx.set_distributed_slack(self.distributed_slack)
        self.nr.install_stopping_criteria(ctx)
        return ctx

    def run(self, x0, target, row_mode, seed_limit_types=None) -> AcSolverResult:
        r = self._run(x0, target, row_mode, seed_limit_types)
        if r.status == "LINEAR_SOLVE_FAILED" and self.robust_retry and not self.robust_mode:
            self.context.close()
            self.context = self._make_context(N.LU_ROBUST)
            self.robust_mode = True
            r = self._run(x0, target, row_mode, seed_limit_types)
        return r

    def _run(self, x0, target, row_mode, seed_limit_types=None) -> AcSolverResult:
        nr = self.nr
        inp, keep = N.BatchContext.make_inputs(
            1, x0, target, row_mode, reactive_limits_seed_state=None if seed_limit_types is None else
            np.asarray(seed_limit_types, np.int8)[None, :], max_iter=nr.maxIterations, tol=nr.maxMismatch,
            reactive_limit_tol=nr.reactive_limit_band(), state_vector_scaling_mode=nr.scaling_code(),
            max_dv=nr.maxVoltageChangeV, max_dphi=nr.maxVoltageChangePhi,
            max_outer_iterations=nr.maxOuterLoopIterations, max_pq_pv_switch=nr.maxPqPvSwitch,
            skip_reactive_limits=not nr.reactiveLimits, **nr.slack_inputs(self.distributed_slack),
            line_search_max_iteration=nr.lineSearchStateVectorScalingMaxIteration,
            line_search_step_fold=nr.lineSearchStateVectorScalingStepFold, conv_eps_per_eq=nr.convEpsPerEq,
            linear_solve_fail_ratio=nr.linearSolveFailRatio)
        r = self.context.solve(inp)
        n_rl = 0 if self.reactive_limits is None else len(self.reactive_limits["voltage_row"])
        m, t, lt = self.context.read_outer_loop_state(1, n_rl)
        self._last_status = int(r["status"][0])
        return AcSolverResult(N.STATUS_NAMES[int(r["status"][0])], int(r["iterations"][0]),
                              int(r["total_iterations"][0]), float(r["norm"][0]), r["x"][0].copy(), m[0], t[0],
                              lt[0])


    def detect_violations(self, limit_table: N.LimitTable, band=1e-7, managed=True) -> np.ndarray:
        """The violations of the last run's converged state (N.VIOLATION_DTYPE, scenario 0): the base case's, against
        which solve_batch_n1 compares the contingencies' (the reference's pre-contingency detection,
        GpuBaseViolations.java:164-190). managed: as OLF's manager keeps them (one per key, the last one fed);
        otherwise every violation the scan feeds it."""
        if getattr(self, "_last_status", None) != N.CONVERGED:
            raise RuntimeError("violations are detected on a converged run's state")
        v = limit_table.detect(self.context, band=band)
        if not managed:
            return v
        from .violations import manage
        return manage(v, limit_table.n_branches, bus_nominal_kv=limit_table.bus_nominal_kv)


def longest_first(cost, tie=None) -> np.ndarray:
    """The order in which a batch solves its scenarios, given an expected cost per scenario: most expensive first;
    ties broken by `tie`, largest first (e.g. the outaged branch's base-case flow magnitude, outage_flow_magnitude:
    a larger flow to redistribute is a harder start), then in the given order. Consecutive scenarios share a tile and
    tiles start in order as workgroups free up, so this pairs scenarios of similar length in a tile and starts the
    long ones early (longest-processing-time scheduling); results are reported in the caller's order."""
    c = -np.asarray(cost, dtype=np.float64)
    if tie is None:
        return np.argsort(c, kind="stable")
    return np.lexsort((-np.asarray(tie, dtype=np.float64), c))


def outage_flow_magnitude(net, x_base, branches) -> np.ndarray:
    """|S| at side 1 of each outaged branch in the base-case state (pu): the flow the outage redistributes (for
    scenarios of several branches, the largest)."""
    from . import gridgen
    s1 = np.abs(gridgen.branch_flows(net, x_base)[0])
    br = np.asarray(branches)
    if br.ndim == 1:
        return s1[br]
    return np.where(br >= 0, s1[np.maximum(br, 0)], 0.0).max(axis=1)


@dataclass
class BatchScenarioResult:
    """Per-contingency result (the reference's BatchScenarioResult)."""
    status: np.ndarray
    iterations: np.ndarray
    total_iterations: np.ndarray
    norm: np.ndarray
    x: np.ndarray | None
    islanded_count: np.ndarray
    chunks: int
    device_seconds: float
    warm_start_fallbacks: int = 0
    violations: np.ndarray | None = None   # N.VIOLATION_DTYPE, scenario = contingency index (with a limit table)
    robust_attempted: int = 0              # scenarios the robust retry pass solved again
    robust_recovered: int = 0              # of which it kept the answer (CONVERGED or MAX_ITERATION_REACHED)
    robust_errors: list = field(default_factory=list)   # failures of the retry pass (the first pass's answers kept)
    warm_started: np.ndarray | None = None  # [K] bool: the answer is the warm-started solve's (not a base-start redo)
    confirmed: int = 0                      # converged warm-started answers solved again for their limit decisions


class GpuBatchedSecurityAnalysisSolver:
    """Batched N-1 from a converged base case: islands frozen at 1<0 in their scenario, chunks sized from device
    memory, every chunk one batched solve. Scenarios the batch rejects (REJECTED or LINEAR_SOLVE_FAILED: a singular
    pivot block in the unpivoted factorization, a bad linear solve at the iteration budget) are solved again on the
    device in the robust mode (N.LU_ROBUST), whose answer is kept when it is CONVERGED or MAX_ITERATION_REACHED
    (retryRejectedOnDevice, GpuBatchPrecomputer.java:635-830); robust_retry=False turns that pass off. force_mode
    builds every context in one N.LU_* mode instead of the parameters' (GpuBatchContextCache.forceModeForTests,
    GpuBatchContextCache.java:144)."""

    def __init__(self, structure: N.Structure, reactive_limits: dict | None = None,
                 gpu_parameters: GpuLoadFlowParameters | None = None,
                 nr_parameters: NewtonRaphsonParameters | None = None, slack_bus: int = 0, max_chunk: int = 4096,
                 lanes_per_tile: int = 0, threads_per_tile: int = 0, distributed_slack: dict | None = None,
                 robust_retry: bool = True, force_mode: int | None = None):
        self.structure = structure
        self.robust_retry = robust_retry
        self.force_mode = force_mode
        self._robust_ctx = None
        self._redo_ctx = [None, None]
        self.distributed_slack = distributed_slack
        self.reactive_limits = reactive_limits
        self.gpu_parameters = gpu_parameters or GpuLoadFlowParameters()
        self.nr = nr_parameters or NewtonRaphsonParameters()
        self.slack_bus = slack_bus
        self.max_chunk = max_chunk
        self.lanes_per_tile = lanes_per_tile
        self.threads_per_tile = threads_per_tile
        self._ctx = None
        self._islands = {}    # outage set -> islanded buses (they depend on the structure only)

    def plan_islands(self, contingencies: np.ndarray):
        """(row_mode_entries, target_override_index, target_override_value, islanded_count) for the chunk."""
        n = self.structure.n
        k = len(contingencies)
        c = np.asarray(contingencies, dtype=np.int64).reshape(k, -1)
        # one island analysis per distinct outage set (sorted, unused places -1 first)
        uniq, inv = np.unique(np.sort(c, axis=1), axis=0, return_inverse=True)
        inv = np.asarray(inv).reshape(-1)
        per_set = []
        for row in uniq:
            key = tuple(int(o) for o in row if o >= 0)
            buses = self._islands.get(key)
            if buses is None:
                buses = np.nonzero(self.structure.islanded_buses(list(key), self.slack_bus))[0]
                self._islands[key] = buses
            per_set.append(buses)
        counts = np.array([b.size for b in per_set], np.int32)[inv] if k else np.zeros(0, np.int32)
        hit = np.nonzero(counts)[0]
        if hit.size == 0:
            return np.zeros((0, 3), np.int32), np.zeros(0, np.int64), np.zeros(0, np.float64), counts
        # per scenario with islanded buses, in scenario order: for each bus its angle row (0) then its voltage row (1)
        buses = np.concatenate([per_set[inv[s]] for s in hit]).astype(np.int64)
        scen = np.repeat(hit, counts[hit]).astype(np.int64)
        rows = np.stack([2 * buses, 2 * buses + 1], axis=1).reshape(-1)
        scen2 = np.repeat(scen, 2)
        rme = np.stack([scen2, rows, np.zeros_like(rows)], axis=1).astype(np.int32)   # FIXED_VARIABLE
        tix = scen2 * n + rows
        tvl = np.tile(np.array([0.0, 1.0]), buses.size)
        return rme, tix.astype(np.int64), tvl, counts

    def context(self, chunk: int) -> N.BatchContext:
        if self._ctx is None or self._ctx.capacity < chunk:
            if self._ctx is not None:
                self._ctx.close()
            mode = self.lu_mode
            self._ctx = N.BatchContext(self.structure, chunk, mode, self.reactive_limits,
                                       1 if mode == N.LU_ROBUST else self.lanes_per_tile, self.threads_per_tile)
            if self.distributed_slack is not None:
                self._ctx.set_distributed_slack(self.distributed_slack)
            self.nr.install_stopping_criteria(self._ctx)
        return self._ctx

    @property
    def lu_mode(self) -> int:
        return self.gpu_parameters.lu_mode if self.force_mode is None else self.force_mode

    def robust_context(self, chunk: int) -> N.BatchContext:
        if self._robust_ctx is None or self._robust_ctx.capacity < chunk:
            if self._robust_ctx is not None:
                self._robust_ctx.close()
            self._robust_ctx = N.BatchContext(self.structure, chunk, N.LU_ROBUST, self.reactive_limits, 1,
                                              self.threads_per_tile)
            if self.distributed_slack is not None:
                self._robust_ctx.set_distributed_slack(self.distributed_slack)
            self.nr.install_stopping_criteria(self._robust_ctx)
        return self._robust_ctx

    def redo_context(self, chunk: int, which: int = 0) -> N.BatchContext:
        """The contexts of the base-start solves of warm-started scenarios (fallbacks and confirmations): one
        scenario per tile, so that each of these few scenarios runs its Newton steps on a whole workgroup (a step's
        time is set by the levels' load chains, shorter with more threads per scenario). `which` picks one of two
        such contexts (each with its own stream), so that two sets of them can be solved at once."""
        ctx = self._redo_ctx[which]
        if ctx is None or ctx.capacity < chunk:
            if ctx is not None:
                ctx.close()
            # capacities in powers of two (at least 16), so that a context is rarely made again: making one waits
            # for the device, including a solve running on the other context
            cap = max(16, 1 << (int(chunk) - 1).bit_length())
            ctx = N.BatchContext(self.structure, cap, self.lu_mode, self.reactive_limits, 1, self.threads_per_tile)
            if self.distributed_slack is not None:
                ctx.set_distributed_slack(self.distributed_slack)
            self.nr.install_stopping_criteria(ctx)
            self._redo_ctx[which] = ctx
        return ctx

    def solve_batch_n1(self, x_base, target_base, row_mode_base, limit_types_base, contingencies,
                       x0_per_scenario=None, states_to_host=True, warm_start=None,
                       warm_start_features=None, limit_table: N.LimitTable | None = None,
                       violation_band=1e-7, cost=None, cost_tie=None, base_violations=None,
                       increased_violations=None, keep_voltage_angle_violations=False,
                       confirm_band=1e-4) -> BatchScenarioResult:
        """solveBatchN1: contingencies is [K] or [K, m] closed-branch ids (-1 padded). With a warm_start network
        (olfgpu.warmstart.WarmStartMLP, single outages) every scenario starts from its prediction, and any that does
        not converge is solved again from the base-case start. With a limit_table the violations of every converged
        contingency are detected on its final state (device screen, exact host decisions); given the base case's
        violations as well (GpuNewtonRaphson.detect_violations), each contingency's list is the one OLF's
        LimitViolationManager keeps against them (olfgpu.violations.manage: one violation per key, those weaker than
        or equivalent to the base's dropped under increased_violations, default powsybl-core's thresholds; the
        reference's post-contingency detection, GpuBaseViolations.java:27-41). With a cost per
        contingency (e.g. warmstart.iteration_costs of its branch) the batch is solved longest first
        (longest_first, ties broken by cost_tie when given); the results are in the contingencies' order either
        way. Tables with voltage-angle limits need keep_voltage_angle_violations to be filtered against the base case
        (violations.manage).

        With a warm start and a limit table, a converged warm-started answer is another Newton iterate inside the same
        stopping tolerance as the base-start one (the reference's semantics), so a limit decision on a quantity that
        close to its threshold can differ: every contingency whose warm-started answer has a monitored quantity within
        relative confirm_band of a threshold (voltages held at a setpoint excluded; LimitTable.ambiguous) is solved
        again from the base-case start, whose answer and violations stand (result.confirmed counts them; 0 turns the
        rule off)."""
        import threading
        import time

        from .violations import near_base_boundary

        if cost is not None:
            order = longest_first(cost, cost_tie)
            inv = np.empty_like(order)
            inv[order] = np.arange(len(order))

            def part(a):
                return None if a is None else np.asarray(a)[order]
            r = self.solve_batch_n1(x_base, target_base, row_mode_base, limit_types_base,
                                    np.asarray(contingencies)[order], part(x0_per_scenario), states_to_host,
                                    warm_start, part(warm_start_features), limit_table, violation_band,
                                    base_violations=base_violations, increased_violations=increased_violations,
                                    keep_voltage_angle_violations=keep_voltage_angle_violations,
                                    confirm_band=confirm_band)
            r.status, r.iterations, r.total_iterations, r.norm = (r.status[inv], r.iterations[inv],
                                                                  r.total_iterations[inv], r.norm[inv])
            r.islanded_count = r.islanded_count[inv]
            if r.warm_started is not None:
                r.warm_started = r.warm_started[inv]
            if r.x is not None:
                r.x = r.x[inv]
            if r.violations is not None:
                r.violations["scenario"] = order[r.violations["scenario"]]
                r.violations = r.violations[np.argsort(r.violations["scenario"], kind="stable")]
            return r

        ctg = np.asarray(contingencies, dtype=np.int32)
        if ctg.ndim == 1:
            ctg = ctg[:, None]
        k = ctg.shape[0]
        chunk = min(k, self.max_chunk, self.structure.recommend_batch_size(self.max_chunk, self.lu_mode))
        nr = self.nr
        n = self.structure.n
        st = np.zeros(k, np.int32)
        it = np.zeros(k, np.int32)
        tot = np.zeros(k, np.int32)
        nm = np.zeros(k)
        xs = np.zeros((k, n)) if states_to_host else None
        islanded = np.zeros(k, np.int32)
        seed = None
        lt = None if limit_types_base is None else np.asarray(limit_types_base, np.int8)
        device_s = 0.0
        chunks = 0
        fallback = 0
        robust_attempted = 0
        robust_recovered = 0
        robust_errors = []
        from_warm = np.zeros(k, bool)

        def solve_part(part, x0_host=None, x0_dev=None, robust=False, redo=False, which=0):
            nonlocal device_s, chunks, seed
            rme, tix, tvl, cnt = self.plan_islands(part)
            if lt is not None and lt.size:
                seed = lt   # one row, shared by every scenario of the chunk
            inp, keep = N.BatchContext.make_inputs(
                len(part), x_base, target_base, row_mode_base, row_mode_entries=rme, target_overrides=(tix, tvl),
                outaged_branches=part, x0_per_scenario=x0_host, reactive_limits_seed_state=seed,
                max_iter=nr.maxIterations, tol=nr.maxMismatch, reactive_limit_tol=nr.reactive_limit_band(),
                state_vector_scaling_mode=nr.scaling_code(), max_dv=nr.maxVoltageChangeV,
                max_dphi=nr.maxVoltageChangePhi, max_outer_iterations=nr.maxOuterLoopIterations,
                max_pq_pv_switch=nr.maxPqPvSwitch, skip_reactive_limits=not nr.reactiveLimits,
                states_to_host=states_to_host, x0_per_scenario_device=x0_dev,
                **nr.slack_inputs(self.distributed_slack),
                line_search_max_iteration=nr.lineSearchStateVectorScalingMaxIteration,
                line_search_step_fold=nr.lineSearchStateVectorScalingStepFold, conv_eps_per_eq=nr.convEpsPerEq,
                active_row_counts=(n - 2 * cnt).astype(np.int32), linear_solve_fail_ratio=nr.linearSolveFailRatio)
            ctx = (self.robust_context(len(part)) if robust else
                   self.redo_context(len(part), which) if redo and self.lu_mode != N.LU_ROBUST else
                   self.context(len(part)))
            if not redo:
                last_ctx[0] = ctx
            t0 = time.perf_counter()
            r = ctx.solve(inp, want_states=states_to_host)
            with counters:
                device_s += time.perf_counter() - t0
                chunks += 1
            return r, cnt, ctx

        counters = threading.Lock()
        viol = []
        managed = []   # lists already through OLF's manager (scenarios disjoint from those in viol)
        confirmed = 0

        def manage_kept(v):
            from .violations import manage
            return manage(v, limit_table.n_branches, base_violations, increased_violations,
                          limit_table.bus_nominal_kv, keep_voltage_angle_violations)

        last_ctx = [None]

        def detect(rows, ctx=None, keep=None, v=None):
            """violations of the last solve (or v), scenario indices mapped through rows (those of scenarios not in
            keep dropped)"""
            if limit_table is None:
                return
            if v is None:
                v = limit_table.detect(ctx or last_ctx[0], band=violation_band)
            if keep is not None:
                v = v[keep[v["scenario"]]]
            v["scenario"] = rows[v["scenario"]]
            viol.append(v)

        for c0 in range(0, k, chunk):
            c1 = min(k, c0 + chunk)
            part = ctg[c0:c1]
            x0_dev = None
            if warm_start is not None:
                x0_dev = warm_start.predict_warm_start(x_base, part[:, 0], warm_start_features[c0:c1])
            r, cnt, _ = solve_part(part, None if x0_per_scenario is None else x0_per_scenario[c0:c1], x0_dev)
            islanded[c0:c1] = cnt
            st[c0:c1] = r["status"]
            it[c0:c1] = r["iterations"]
            tot[c0:c1] = r["total_iterations"]
            nm[c0:c1] = r["norm"]
            if states_to_host:
                xs[c0:c1] = r["x"]
            if warm_start is None:
                detect(np.arange(c0, c1))
            else:
                from_warm[c0:c1] = True
                # a warm start that did not converge, or whose answer has a limit decision within confirm_band of
                # its threshold (or of the increased-violation comparison with the base case's record of the same
                # limit): solve again from the base-case start, whose answer stands
                again = r["status"] != 0
                redo_runs = []   # (rows of the chunk, thread, its outcome) of the base-start solves under way

                def start_redo(rows, which):
                    self.redo_context(len(rows), which)   # made (or grown) on this thread, before its solve starts
                    out = {}

                    def run():
                        try:
                            out["r"] = solve_part(part[rows], redo=True, which=which)
                        except BaseException as e:   # re-raised on the caller's thread
                            out["e"] = e
                    th = threading.Thread(target=run)
                    th.start()
                    redo_runs.append((rows, th, out))

                v = None
                if limit_table is not None and confirm_band > 0:
                    near = limit_table.ambiguous(self._ctx, confirm_band) & (r["status"] == 0)
                    again |= near
                    if base_violations is not None:
                        # these solves start now, on a context of their own; the increased-violation boundary is
                        # checked on the detected records meanwhile, and the few scenarios it adds start on a
                        # second context
                        first = np.nonzero(again)[0]
                        if first.size:
                            start_redo(first, 0)
                        v = limit_table.detect(self._ctx, band=violation_band)
                        nb = near_base_boundary(v, limit_table.n_branches, base_violations, increased_violations,
                                                limit_table.bus_nominal_kv, confirm_band)
                        extra = np.zeros(len(part), bool)
                        extra[v["scenario"][nb]] = True
                        extra &= (r["status"] == 0) & ~again
                        near |= extra
                        again |= extra
                        if extra.any():
                            start_redo(np.nonzero(extra)[0], 1)
                    confirmed += int(near.sum())
                redo = np.nonzero(again)[0]
                from_warm[c0 + redo] = False
                fallback += int((r["status"] != 0).sum()) if redo.size else 0
                try:
                    detect(np.arange(c0, c1), keep=~again, v=v)
                    if redo_runs and limit_table is not None and base_violations is not None:
                        # the manager's pass over the warm-started answers kept (scenarios of their own) runs on
                        # the host while the device solves the others again
                        managed.append(manage_kept(viol.pop()))
                finally:
                    for _, th, _ in redo_runs:
                        th.join()
                for _, _, out in redo_runs:
                    if "e" in out:
                        raise out["e"]
                if redo.size and not redo_runs:
                    redo_runs.append((redo, None, {"r": solve_part(part[redo], redo=True)}))
                for rows, _, out in redo_runs:
                    r2, _, ctx2 = out["r"]
                    detect(c0 + rows, ctx2)
                    st[c0 + rows] = r2["status"]
                    it[c0 + rows] = r2["iterations"]
                    tot[c0 + rows] += r2["total_iterations"]
                    nm[c0 + rows] = r2["norm"]
                    if states_to_host:
                        xs[c0 + rows] = r2["x"]
            if self.robust_retry and self.lu_mode != N.LU_ROBUST:
                # the scenarios the batch rejected, once more in the robust mode from the base-case start
                redo = np.nonzero((st[c0:c1] == N.REJECTED) | (st[c0:c1] == N.LINEAR_SOLVE_FAILED))[0]
                if redo.size:
                    robust_attempted += redo.size
                    try:
                        r3, _, _ = solve_part(part[redo], robust=True)
                    except N.OlfHipError as e:
                        # an optional pass: its failure (e.g. no device memory for its FP64 Krylov basis) keeps the
                        # first pass's answers (GpuBatchPrecomputer.java:780-786)
                        robust_errors.append(str(e))
                        continue
                    keep = np.isin(r3["status"], (N.CONVERGED, N.MAX_ITERATION_REACHED))
                    rows = c0 + redo[keep]
                    from_warm[rows] = False
                    robust_recovered += rows.size
                    detect(c0 + redo, self._robust_ctx)   # converged lanes only, all of them kept
                    st[rows] = r3["status"][keep]
                    it[rows] = r3["iterations"][keep]
                    tot[rows] += r3["total_iterations"][keep]
                    nm[rows] = r3["norm"][keep]
                    if states_to_host:
                        xs[rows] = r3["x"][keep]
        violations = None
        if limit_table is not None:
            violations = np.concatenate(viol) if viol else np.zeros(0, N.VIOLATION_DTYPE)
            violations = violations[np.argsort(violations["scenario"], kind="stable")]
            if base_violations is not None:
                violations = manage_kept(violations)
                if managed:
                    # the manager works scenario by scenario: lists of disjoint scenarios merge by scenario
                    violations = np.concatenate([violations] + managed)
                    violations = violations[np.argsort(violations["scenario"], kind="stable")]
        return BatchScenarioResult(st, it, tot, nm, xs, islanded, chunks, device_s, fallback, violations,
                                   robust_attempted, robust_recovered, robust_errors,
                                   from_warm if warm_start is not None else None, confirmed)
