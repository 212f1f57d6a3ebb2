"""ctypes binding of libolfhip.so (the C ABI declared in include/olfhip.h).

The library is required: every entry point raises if it is missing — there is no CPU fallback on this path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# OLFHIP_LIBRARY: another build of the same library (a build with other compile-time settings)
LIB_PATH = os.environ.get("OLFHIP_LIBRARY") or os.path.join(os.path.dirname(HERE), "libolfhip.so")

OK, E_ARGUMENT, E_DEVICE, E_MEMORY, E_UNSUPPORTED = 0, -1, -2, -3, -4
CONVERGED, MAX_ITERATION_REACHED, DIVERGED, REJECTED, LINEAR_SOLVE_FAILED = 0, 1, -1, -2, -3
STATUS_NAMES = {0: "CONVERGED", 1: "MAX_ITERATION_REACHED", -1: "DIVERGED", -2: "REJECTED", -3: "LINEAR_SOLVE_FAILED"}
LU_FP64, LU_FP32_REFINED, LU_ROBUST = 0, 1, 2   # LU_ROBUST: one scenario per tile, static pivoting + GMRES
SCALING_NONE, SCALING_LINE_SEARCH, SCALING_MAX_VOLTAGE_CHANGE = 0, 1, 2

# every entry point include/olfhip.h declares
EXPORTS = [
    "olfhip_last_error", "olfhip_device_count", "olfhip_create_structure", "olfhip_destroy_structure",
    "olfhip_structure_statistics", "olfhip_structure_order", "olfhip_recommend_batch_size",
    "olfhip_create_batch_context", "olfhip_destroy_batch_context", "olfhip_solve_batch_context",
    "olfhip_read_states", "olfhip_islanded_buses", "olfhip_lu_factorize_and_solve", "olfhip_structure_pattern",
    "olfhip_autotune", "olfhip_context_timing", "olfhip_read_outer_loop_state", "olfhip_structure_schedule",
    "olfhip_neighbourhoods", "olfhip_mlp_create", "olfhip_mlp_destroy", "olfhip_mlp_set_normalization",
    "olfhip_mlp_set_base_states", "olfhip_mlp_train_step", "olfhip_mlp_forward", "olfhip_mlp_gradients",
    "olfhip_mlp_parameters", "olfhip_mlp_predict_warm_start", "olfhip_mlp_gemm_test",
    "olfhip_create_limit_table", "olfhip_destroy_limit_table", "olfhip_screen_violations", "olfhip_detect_violations",
    "olfhip_set_distributed_slack",
    "olfhip_set_row_criteria", "olfhip_screen_violations_no_alloc", "olfhip_ambiguous_scenarios",
    "olfhip_context_graph_statistics", "olfhip_set_x0_seeds", "olfhip_begin_batch_install",
    "olfhip_read_handed_back_slots",
]
NO_ALLOC_SCRATCH_BYTES = 32 << 20   # OLFHIP_NO_ALLOC_SCRATCH_BYTES


class OlfHipError(RuntimeError):
    """A failed C-ABI call: `code` is the OLFHIP_E_* code, the message is olfhip_last_error()."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class ReactiveLimits(C.Structure):
    _fields_ = [("n_entries", C.c_int), ("voltage_row", C.c_void_p), ("min_q", C.c_void_p), ("max_q", C.c_void_p),
                ("load_q", C.c_void_p), ("target_v", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("lanes_per_tile", C.c_int), ("threads_per_tile", C.c_int)]


class SolveInputs(C.Structure):
    _fields_ = [("n_scenarios", C.c_int), ("x0_base", C.c_void_p), ("target_base", C.c_void_p),
                ("base_row_mode", C.c_void_p), ("n_row_mode_entries", C.c_int), ("row_mode_entries", C.c_void_p),
                ("n_target_overrides", C.c_int), ("target_override_index", C.c_void_p),
                ("target_override_value", C.c_void_p), ("max_outages", C.c_int), ("outaged_branches", C.c_void_p),
                ("x0_per_scenario", C.c_void_p), ("reactive_limits_seed_state", C.c_void_p),
                ("max_iter", C.c_int), ("tol", C.c_double), ("reactive_limit_tol", C.c_double),
                ("state_vector_scaling_mode", C.c_int), ("max_dv", C.c_double), ("max_dphi", C.c_double),
                ("max_outer_iterations", C.c_int), ("max_pq_pv_switch", C.c_int), ("skip_reactive_limits", C.c_int),
                ("states_to_host", C.c_int), ("x0_per_scenario_device", C.c_void_p),
                ("distributed_slack", C.c_int), ("slack_dist_tol", C.c_double),
                ("reject_slack_distribution_failure", C.c_int), ("line_search_max_iteration", C.c_int),
                ("line_search_step_fold", C.c_double), ("conv_eps_per_eq", C.c_double),
                ("active_row_counts", C.c_void_p), ("reactive_limits_seed_shared", C.c_int),
                ("linear_solve_fail_ratio", C.c_double), ("pause_at_ended", C.c_int)]


class DistributedSlackDesc(C.Structure):
    _fields_ = [("n_buses", C.c_int), ("bus", C.c_void_p), ("factor", C.c_void_p), ("base_target_p", C.c_void_p),
                ("max_delta", C.c_void_p), ("min_delta", C.c_void_p), ("slack_bus", C.c_int),
                ("slack_target_p", C.c_double)]


class SolveOutputs(C.Structure):
    _fields_ = [("results", C.c_void_p), ("status", C.c_void_p), ("iterations", C.c_void_p),
                ("total_iterations", C.c_void_p), ("norm", C.c_void_p), ("n_handed_back", C.c_void_p)]


_LIB = None


class LimitTableDesc(C.Structure):
    _fields_ = [("n_branches", C.c_int), ("branch", C.c_void_p), ("group_start", C.c_void_p),
                ("group_side", C.c_void_p), ("group_quantity", C.c_void_p), ("limit_start", C.c_void_p),
                ("limit_threshold", C.c_void_p), ("n_buses", C.c_int), ("bus_high", C.c_void_p),
                ("bus_low", C.c_void_p), ("n_angle_limits", C.c_int), ("angle_from", C.c_void_p),
                ("angle_to", C.c_void_p), ("angle_high", C.c_void_p), ("angle_low", C.c_void_p)]


# olfhip_violation as a numpy record
VIOLATION_DTYPE = np.dtype([("scenario", np.int32), ("kind", np.int32), ("element", np.int32), ("side", np.int32),
                            ("quantity", np.int32), ("limit", np.int32), ("value", np.float64),
                            ("threshold", np.float64)])
VIOLATION_KINDS = {0: "BRANCH", 1: "HIGH_VOLTAGE", 2: "LOW_VOLTAGE", 3: "HIGH_VOLTAGE_ANGLE", 4: "LOW_VOLTAGE_ANGLE"}


def lib():
    """The loaded library; raises if it was not built (no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        L.olfhip_last_error.restype = C.c_char_p
        L.olfhip_device_count.restype = i
        L.olfhip_create_structure.argtypes = [i, i, vp, vp, i, vp, vp, C.POINTER(vp)]
        L.olfhip_destroy_structure.argtypes = [vp]
        L.olfhip_destroy_structure.restype = None
        L.olfhip_structure_statistics.argtypes = [vp, vp]
        L.olfhip_structure_order.argtypes = [vp, vp]
        L.olfhip_structure_pattern.argtypes = [vp, C.POINTER(i), vp, vp]
        L.olfhip_structure_schedule.argtypes = [vp, i, C.POINTER(i), vp]
        L.olfhip_recommend_batch_size.argtypes = [vp, i, i, C.POINTER(i)]
        L.olfhip_create_batch_context.argtypes = [vp, i, i, C.POINTER(ReactiveLimits), C.POINTER(Tuning),
                                                  C.POINTER(vp)]
        L.olfhip_destroy_batch_context.argtypes = [vp]
        L.olfhip_destroy_batch_context.restype = None
        L.olfhip_solve_batch_context.argtypes = [vp, C.POINTER(SolveInputs), C.POINTER(SolveOutputs)]
        L.olfhip_read_states.argtypes = [vp, i, vp]
        L.olfhip_context_timing.argtypes = [vp, C.POINTER(d), C.POINTER(i)]
        L.olfhip_context_graph_statistics.argtypes = [vp, C.POINTER(i), C.POINTER(i)]
        L.olfhip_set_x0_seeds.argtypes = [vp, i, vp, vp]
        L.olfhip_begin_batch_install.argtypes = [vp, i, vp]
        L.olfhip_read_handed_back_slots.argtypes = [vp, C.POINTER(i), vp]
        L.olfhip_set_distributed_slack.argtypes = [vp, C.c_void_p]
        L.olfhip_set_row_criteria.argtypes = [vp, C.c_void_p, C.c_void_p]
        L.olfhip_read_outer_loop_state.argtypes = [vp, i, vp, vp, vp]
        L.olfhip_islanded_buses.argtypes = [vp, i, i, vp, vp, C.POINTER(i)]
        L.olfhip_lu_factorize_and_solve.argtypes = [vp, i, i, vp, vp, i, vp, vp, vp]
        L.olfhip_autotune.argtypes = [vp, i, C.POINTER(ReactiveLimits), C.POINTER(SolveInputs), C.POINTER(Tuning),
                                      C.POINTER(d)]
        L.olfhip_create_limit_table.argtypes = [vp, C.POINTER(LimitTableDesc), C.POINTER(vp)]
        L.olfhip_destroy_limit_table.argtypes = [vp]
        L.olfhip_destroy_limit_table.restype = None
        L.olfhip_screen_violations.argtypes = [vp, vp, d, i, vp, vp]
        L.olfhip_screen_violations_no_alloc.argtypes = [vp, vp, d, i, i, vp, vp]
        L.olfhip_ambiguous_scenarios.argtypes = [vp, vp, d, i, vp]
        L.olfhip_detect_violations.argtypes = [vp, vp, d, i, i, vp, C.POINTER(i)]
        _LIB = L
    return _LIB


def check(rc):
    if rc != OK:
        raise OlfHipError(rc, lib().olfhip_last_error().decode())


def ptr(a):
    return None if a is None else a.ctypes.data


def device_to_host(device_ptr: int, shape, dtype=np.float64) -> np.ndarray:
    """A synchronous copy of device memory (e.g. olfhip_mlp_predict_warm_start's starts) into a new host array."""
    out = np.zeros(shape, dtype)
    hip = C.CDLL("libamdhip64.so")
    rc = hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(device_ptr), C.c_size_t(out.nbytes),
                       C.c_int(2))   # hipMemcpyDeviceToHost
    if rc != 0:
        raise OlfHipError(E_DEVICE, f"hipMemcpy device to host failed: {rc}")
    return out


def device_count() -> int:
    return lib().olfhip_device_count()


class Structure:
    """Owner of an olfhip_structure (StructureNative.createStructure / destroyStructure)."""

    def __init__(self, n, closed_branch_parameters, closed_branch_indices, shunt_parameters=None, shunt_indices=None):
        self.n = int(n)
        self._p = np.ascontiguousarray(closed_branch_parameters, dtype=np.float64).reshape(-1)
        self._i = np.ascontiguousarray(closed_branch_indices, dtype=np.int32).reshape(-1)
        ns = 0 if shunt_parameters is None else len(shunt_parameters)
        self._sp = None if ns == 0 else np.ascontiguousarray(shunt_parameters, dtype=np.float64).reshape(-1)
        self._si = None if ns == 0 else np.ascontiguousarray(shunt_indices, dtype=np.int32).reshape(-1)
        h = C.c_void_p()
        check(lib().olfhip_create_structure(self.n, self._i.size // 12, ptr(self._p), ptr(self._i), ns, ptr(self._sp),
                                            ptr(self._si), C.byref(h)))
        self.h = h

    @classmethod
    def from_network(cls, net):
        sp_, si_ = net.shunts() if hasattr(net, "shunts") else (None, None)
        return cls(net.n, net.pi_model(), net.branch_indices(), sp_, si_)

    def close(self):
        if getattr(self, "h", None):
            lib().olfhip_destroy_structure(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def statistics(self) -> dict:
        s = np.zeros(6, np.int64)
        check(lib().olfhip_structure_statistics(self.h, ptr(s)))
        return dict(buses=int(s[0]), lu_slots=int(s[1]), levels=int(s[2]), factor_block_mads=int(s[3]),
                    pattern_nnz=int(s[4]), lu_nnz=int(s[5]))

    def order(self) -> np.ndarray:
        o = np.zeros(self.n // 2, np.int32)
        check(lib().olfhip_structure_order(self.h, ptr(o)))
        return o

    def pattern(self):
        nnz = C.c_int()
        check(lib().olfhip_structure_pattern(self.h, C.byref(nnz), None, None))
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        check(lib().olfhip_structure_pattern(self.h, C.byref(nnz), ptr(rp), ptr(ci)))
        return rp, ci

    SCHEDULE = ("item", "bucket", "level_bucket_ptr", "bwd", "bwd_bucket", "level_bwd_bucket_ptr", "level_item_ptr",
                "level_node_ptr", "diag_item_of", "contrib", "level_nodes", "s_members")

    def schedule(self):
        """The LU level schedule (olfhip_structure_schedule), records reshaped to [.., 4]."""
        out = {}
        for which, name in enumerate(self.SCHEDULE):
            cnt = C.c_int()
            check(lib().olfhip_structure_schedule(self.h, which, C.byref(cnt), None))
            v = np.zeros(cnt.value, np.int32)
            check(lib().olfhip_structure_schedule(self.h, which, C.byref(cnt), ptr(v)))
            out[name] = v.reshape(-1, 4) if name in ("item", "bucket", "bwd", "bwd_bucket", "contrib") else v
        return out

    def islanded_buses(self, outaged, slack_bus=0) -> np.ndarray:
        o = np.ascontiguousarray(np.asarray(outaged, dtype=np.int32).reshape(-1))
        isl = np.zeros(self.n // 2, np.uint8)
        cnt = C.c_int()
        check(lib().olfhip_islanded_buses(self.h, slack_bus, o.size, ptr(o), ptr(isl), C.byref(cnt)))
        return isl.astype(bool)

    def recommend_batch_size(self, max_cap, mode=LU_FP64) -> int:
        out = C.c_int()
        check(lib().olfhip_recommend_batch_size(self.h, max_cap, mode, C.byref(out)))
        return out.value

    def lu_factorize_and_solve(self, values, rhs, lane_map=None, mode=LU_FP64):
        """BatchedLuNative.factorizeAndSolve: values [capacity, nnz] on pattern(), rhs [capacity, n]."""
        values = np.ascontiguousarray(values, dtype=np.float64)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        cap = values.shape[0]
        lm = np.arange(cap, dtype=np.int32) if lane_map is None else np.ascontiguousarray(lane_map, np.int32)
        sol = np.zeros((cap, self.n))
        clamps = np.zeros(lm.size, np.int32)
        check(lib().olfhip_lu_factorize_and_solve(self.h, mode, cap, ptr(values), ptr(rhs), lm.size, ptr(lm),
                                                  ptr(sol), ptr(clamps)))
        return sol, clamps


def reactive_limits_struct(rl: dict | None):
    """(ctypes struct, keep-alive arrays) from a dict with voltage_row, min_q, max_q, load_q, target_v."""
    if rl is None or len(rl["voltage_row"]) == 0:
        return ReactiveLimits(0, None, None, None, None, None), []
    keep = [np.ascontiguousarray(rl["voltage_row"], np.int32)] + [
        np.ascontiguousarray(rl[k], np.float64) for k in ("min_q", "max_q", "load_q", "target_v")]
    return ReactiveLimits(len(keep[0]), *[ptr(a) for a in keep]), keep


class BatchContext:
    """Owner of an olfhip_context (BatchSolveNative.createBatchContext / solveBatchContext / destroyBatchContext)."""

    def __init__(self, structure: Structure, n_scenarios: int, mode=LU_FP64, reactive_limits=None,
                 lanes_per_tile=0, threads_per_tile=0):
        self.structure = structure
        self.capacity = n_scenarios
        self.mode = mode
        self._rl, self._rl_keep = reactive_limits_struct(reactive_limits)
        self._tuning = Tuning(lanes_per_tile, threads_per_tile)
        h = C.c_void_p()
        check(lib().olfhip_create_batch_context(structure.h, n_scenarios, mode, C.byref(self._rl),
                                                C.byref(self._tuning), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().olfhip_destroy_batch_context(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def make_inputs(n_scenarios, x0_base, target_base, base_row_mode, row_mode_entries=None, target_overrides=None,
                    outaged_branches=None, x0_per_scenario=None, reactive_limits_seed_state=None, max_iter=15,
                    tol=1e-4, reactive_limit_tol=1e-4, state_vector_scaling_mode=SCALING_NONE, max_dv=0.1,
                    max_dphi=0.174533, max_outer_iterations=20, max_pq_pv_switch=3, skip_reactive_limits=False,
                    states_to_host=True, x0_per_scenario_device=None, distributed_slack=False, slack_dist_tol=0.0,
                    reject_slack_distribution_failure=False, line_search_max_iteration=10,
                    line_search_step_fold=4.0 / 3.0, conv_eps_per_eq=0.0, active_row_counts=None,
                    linear_solve_fail_ratio=0.0, pause_at_ended=0):
        keep = []

        def arr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        x0 = arr(x0_base, np.float64)
        tg = arr(target_base, np.float64)
        md = arr(base_row_mode, np.int32)
        rme = arr(None if row_mode_entries is None or len(row_mode_entries) == 0 else
                  np.asarray(row_mode_entries).reshape(-1, 3), np.int32)
        if target_overrides is not None and len(target_overrides[0]) > 0:
            ti = arr(target_overrides[0], np.int64)
            tv = arr(target_overrides[1], np.float64)
        else:
            ti = tv = None
        ob = None
        mo = 0
        if outaged_branches is not None:
            ob = np.asarray(outaged_branches, dtype=np.int32)
            if ob.ndim == 1:
                ob = ob[:, None]
            mo = ob.shape[1]
            ob = arr(ob, np.int32)
        xp = arr(x0_per_scenario, np.float64)
        # one row of seeds (1-d, or [1][n_entries] with more than one scenario): shared by every scenario
        shared = 0
        if reactive_limits_seed_state is not None:
            rs = np.asarray(reactive_limits_seed_state)
            if rs.ndim == 1 or (rs.shape[0] == 1 and n_scenarios > 1):
                shared = 1
                reactive_limits_seed_state = rs.reshape(-1)
        sd = arr(reactive_limits_seed_state, np.int8)
        ac = arr(active_row_counts, np.int32)
        si = SolveInputs(n_scenarios, ptr(x0), ptr(tg), ptr(md), 0 if rme is None else rme.shape[0], ptr(rme),
                         0 if ti is None else ti.size, ptr(ti), ptr(tv), mo, ptr(ob), ptr(xp), ptr(sd), max_iter, tol,
                         reactive_limit_tol, state_vector_scaling_mode, max_dv, max_dphi, max_outer_iterations,
                         max_pq_pv_switch, int(skip_reactive_limits), int(states_to_host), x0_per_scenario_device,
                         int(distributed_slack), float(slack_dist_tol), int(reject_slack_distribution_failure),
                         int(line_search_max_iteration), float(line_search_step_fold), float(conv_eps_per_eq), ptr(ac),
                         shared, float(linear_solve_fail_ratio), int(pause_at_ended))
        return si, keep

    def set_row_criteria(self, thresholds):
        """olfhip_set_row_criteria: thresholds = (fixed [n], equation [n]) per-row stopping thresholds (pu), or None
        to clear."""
        if thresholds is None:
            check(lib().olfhip_set_row_criteria(self.h, None, None))
            self._rows_keep = None
            return
        keep = tuple(np.ascontiguousarray(a, np.float64) for a in thresholds)
        if any(a.size != self.structure.n for a in keep):
            raise ValueError("row thresholds need n values each")
        check(lib().olfhip_set_row_criteria(self.h, ptr(keep[0]), ptr(keep[1])))
        self._rows_keep = keep

    def set_distributed_slack(self, ds):
        """olfhip_set_distributed_slack: ds is a dict (bus, factor, base_target_p, max_delta, min_delta, slack_bus,
        slack_target_p; pu), or None to clear."""
        if ds is None:
            check(lib().olfhip_set_distributed_slack(self.h, None))
            self._ds_keep = None
            return
        keep = [np.ascontiguousarray(ds["bus"], np.int32)] + [
            np.ascontiguousarray(ds[k], np.float64) for k in ("factor", "base_target_p", "max_delta", "min_delta")]
        d = DistributedSlackDesc(len(keep[0]), *(ptr(a) for a in keep), int(ds["slack_bus"]),
                                 float(ds["slack_target_p"]))
        check(lib().olfhip_set_distributed_slack(self.h, C.byref(d)))
        self._ds_keep = keep

    def begin_install(self, bank_slots=None):
        """olfhip_begin_batch_install: the next solve's scenario j goes into bank slot bank_slots[j], the other slots
        keeping what they hold (None: the whole bank, scenario j into slot j)."""
        if bank_slots is None or len(bank_slots) == 0:
            check(lib().olfhip_begin_batch_install(self.h, 0, None))
            self._bank = False
            return
        a = np.ascontiguousarray(bank_slots, np.int32)
        check(lib().olfhip_begin_batch_install(self.h, a.size, ptr(a)))
        self._bank = True

    def read_handed_back_slots(self):
        """olfhip_read_handed_back_slots: the bank slots of the last solve's rows."""
        n = C.c_int()
        out = np.zeros(self.capacity, np.int32)
        check(lib().olfhip_read_handed_back_slots(self.h, C.byref(n), ptr(out)))
        return out[:n.value].copy()

    def solve(self, inputs: SolveInputs, want_states=True):
        """One solve. A bank solve (after begin_install with slots) or a paused one (pause_at_ended > 0) returns the
        scenarios it handed back, in install order, with their bank slots under "slots"; the readbacks after it index
        scenarios by slot (n_scenarios = capacity)."""
        k = inputs.n_scenarios
        n = self.structure.n
        by_slot = getattr(self, "_bank", False) or inputs.pause_at_ended > 0
        rows = self.capacity if by_slot else k
        res = np.zeros((rows, n + 3)) if want_states else None
        st = np.zeros(rows, np.int32)
        it = np.zeros(rows, np.int32)
        tot = np.zeros(rows, np.int32)
        nm = np.zeros(rows)
        nh = np.zeros(1, np.int32)
        out = SolveOutputs(ptr(res), ptr(st), ptr(it), ptr(tot), ptr(nm), ptr(nh))
        try:
            check(lib().olfhip_solve_batch_context(self.h, C.byref(inputs), C.byref(out)))
        finally:
            self._bank = False   # the chosen slots hold for one solve
        h = int(nh[0])
        self.last_n_scenarios = self.capacity if by_slot else k
        r = dict(x=None if res is None else res[:h, :n], results=None if res is None else res[:h], status=st[:h],
                 iterations=it[:h], total_iterations=tot[:h], norm=nm[:h])
        if by_slot:
            r["slots"] = self.read_handed_back_slots()
        return r

    def timing(self):
        """(event-timed Newton-step milliseconds, step launches) of the last solve."""
        ms = C.c_double()
        steps = C.c_int()
        check(lib().olfhip_context_timing(self.h, C.byref(ms), C.byref(steps)))
        return ms.value, steps.value

    def set_x0_seeds(self, flat_index=None, value=None):
        """olfhip_set_x0_seeds: start values scenario * n + row -> value for the next solves (None clears)."""
        if flat_index is None or len(flat_index) == 0:
            check(lib().olfhip_set_x0_seeds(self.h, 0, None, None))
            self._seed_keep = None
            return
        keep = (np.ascontiguousarray(flat_index, np.int64), np.ascontiguousarray(value, np.float64))
        if keep[0].shape != keep[1].shape:
            raise ValueError("one value per seed index")
        check(lib().olfhip_set_x0_seeds(self.h, keep[0].size, ptr(keep[0]), ptr(keep[1])))
        self._seed_keep = keep

    def graph_statistics(self):
        """(instantiations, in-place updates) of the context's solve graphs so far."""
        a, b = C.c_int(), C.c_int()
        check(lib().olfhip_context_graph_statistics(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def read_outer_loop_state(self, k, n_entries):
        """(row modes [k, n], targets [k, n], reactive-limit types [k, n_entries]) left by the last solve."""
        n = self.structure.n
        m = np.zeros((k, n), np.int32)
        t = np.zeros((k, n))
        lt = np.zeros((k, max(n_entries, 1)), np.int8)
        check(lib().olfhip_read_outer_loop_state(self.h, k, ptr(m), ptr(t), ptr(lt)))
        return m, t, lt[:, :n_entries]

    def read_states(self, k):
        x = np.zeros((k, self.structure.n))
        check(lib().olfhip_read_states(self.h, k, ptr(x)))
        return x


class LimitTable:
    """Owner of an olfhip_limit_table (the reference's ViolationScreenNative.uploadViolationTable) built from the
    arrays of a table dict (gridgen.limit_table layout). An optional "bus_nominal_kv" entry (scalar or per bus, the
    reference's GpuLimitTable.busScale) puts the voltage comparisons of olfgpu.violations.manage in kV."""

    def __init__(self, structure: Structure, table: dict):
        self.structure = structure
        self.n_branches = len(table["branch"])
        kv = table.get("bus_nominal_kv")
        self.bus_nominal_kv = None if kv is None else np.asarray(kv, np.float64)
        self._keep = {k: np.ascontiguousarray(table[k], np.int32) for k in
                      ("branch", "group_start", "group_side", "group_quantity", "limit_start")}
        self._keep.update({k: np.ascontiguousarray(table[k], np.float64) for k in ("threshold", "bus_high",
                                                                                   "bus_low")})
        # optional voltage-angle limits: bus pairs (angle_from, angle_to), rad bounds (NaN: none)
        self._keep.update({k: np.ascontiguousarray(table.get(k, ()), np.int32) for k in ("angle_from", "angle_to")})
        self._keep.update({k: np.ascontiguousarray(table.get(k, ()), np.float64) for k in ("angle_high",
                                                                                           "angle_low")})
        k = self._keep
        self.n_limits = len(k["threshold"])
        self.n_buses = len(k["bus_high"])
        self.n_angle_limits = len(k["angle_from"])
        if not (len(k["angle_to"]) == len(k["angle_high"]) == len(k["angle_low"]) == self.n_angle_limits):
            raise ValueError("voltage-angle limit arrays must have one length")
        desc = LimitTableDesc(len(k["branch"]), ptr(k["branch"]), ptr(k["group_start"]), ptr(k["group_side"]),
                              ptr(k["group_quantity"]), ptr(k["limit_start"]), ptr(k["threshold"]),
                              len(k["bus_high"]), ptr(k["bus_high"]), ptr(k["bus_low"]), self.n_angle_limits,
                              ptr(k["angle_from"]), ptr(k["angle_to"]), ptr(k["angle_high"]), ptr(k["angle_low"]))
        h = C.c_void_p()
        check(lib().olfhip_create_limit_table(structure.h, C.byref(desc), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().olfhip_destroy_limit_table(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def screen(self, context: "BatchContext", k: int, band=1e-7, capacity=1024):
        """The device screen's candidates for the last solve of `context` (k scenarios, its solve's count): (counts
        [k], records [k][capacity])."""
        if k != getattr(context, "last_n_scenarios", None):
            raise ValueError(f"k = {k}: the context's last solve had {getattr(context, 'last_n_scenarios', None)} "
                             "scenarios (the screen fills one row per scenario of that solve)")
        counts = np.zeros(k, np.int32)
        records = np.zeros((k, capacity), np.int32)
        check(lib().olfhip_screen_violations(context.h, self.h, band, capacity, ptr(counts), ptr(records)))
        return counts, records

    def ambiguous(self, context: "BatchContext", band=1e-4) -> np.ndarray:
        """olfhip_ambiguous_scenarios: [k] bool, the converged scenarios of the context's last solve with a monitored
        quantity within relative `band` of a threshold (voltages held at a setpoint excluded)."""
        k = context.last_n_scenarios
        flags = np.zeros(k, np.int32)
        check(lib().olfhip_ambiguous_scenarios(context.h, self.h, band, k, ptr(flags)))
        return flags.astype(bool)

    def screen_no_alloc(self, context: "BatchContext", k: int, band=1e-7, capacity=1024):
        """olfhip_screen_violations_no_alloc: the screen's candidates with no device allocation (its working memory
        in the library's per-device scratch array); same layout as screen()."""
        counts = np.zeros(k, np.int32)
        records = np.zeros((k, capacity), np.int32)
        check(lib().olfhip_screen_violations_no_alloc(context.h, self.h, band, k, capacity, ptr(counts),
                                                       ptr(records)))
        return counts, records

    def detect(self, context: "BatchContext", band=1e-7, capacity=1024, max_violations=1 << 20):
        """Screen + exact host decisions: a VIOLATION_DTYPE array ordered by scenario, branches, buses (a view of
        the first records of a buffer of max_violations: only the pages the records fill are touched)."""
        out = np.empty(max_violations, VIOLATION_DTYPE)
        n = C.c_int()
        check(lib().olfhip_detect_violations(context.h, self.h, band, capacity, max_violations, ptr(out),
                                             C.byref(n)))
        return out[:n.value]


def autotune(structure: Structure, inputs: SolveInputs, mode=LU_FP64, reactive_limits=None):
    rl, keep = reactive_limits_struct(reactive_limits)
    best = Tuning()
    ms = C.c_double()
    check(lib().olfhip_autotune(structure.h, mode, C.byref(rl), C.byref(inputs), C.byref(best), C.byref(ms)))
    return dict(lanes_per_tile=best.lanes_per_tile, threads_per_tile=best.threads_per_tile, ms=ms.value)
