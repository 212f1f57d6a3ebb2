"""ctypes view of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Imported by tests/, by the smoke check and by bench.py's CPU-baseline leg, where the oracle is the checker
(or the timed CPU baseline), never the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
_ubp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

STATUS = {0: "CONVERGED", 1: "MAX_ITERATION_REACHED", -1: "DIVERGED", -2: "REJECTED", -3: "LINEAR_SOLVE_FAILED"}


class NrParams(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("tol", C.c_double), ("reactive_limits", C.c_int), ("rl_tol", C.c_double),
                ("max_outer", C.c_int), ("max_switch", C.c_int), ("scaling_mode", C.c_int),
                ("max_dv", C.c_double), ("max_dphi", C.c_double), ("batch_of_one", C.c_int),
                ("ls_max_iteration", C.c_int), ("ls_step_fold", C.c_double), ("conv_eps_per_eq", C.c_double),
                ("n_active", C.c_int), ("row_thr_fixed", C.c_void_p), ("row_thr_eq", C.c_void_p),
                ("lin_fail_ratio", C.c_double)]


class Rl(C.Structure):
    _fields_ = [("n", C.c_int), ("voltage_row", C.c_void_p), ("bus", C.c_void_p), ("min_q", C.c_void_p),
                ("max_q", C.c_void_p), ("load_q", C.c_void_p), ("target_v", C.c_void_p)]


class Limits(C.Structure):
    _fields_ = [("n_branches", C.c_int), ("branch", C.c_void_p), ("group_start", C.c_void_p),
                ("group_side", C.c_void_p), ("group_quantity", C.c_void_p), ("limit_start", C.c_void_p),
                ("threshold", C.c_void_p), ("n_buses", C.c_int), ("bus_high", C.c_void_p), ("bus_low", C.c_void_p),
                ("n_angle", C.c_int), ("angle_from", C.c_void_p), ("angle_to", C.c_void_p),
                ("angle_high", C.c_void_p), ("angle_low", C.c_void_p)]


class Violation(C.Structure):
    _fields_ = [("kind", C.c_int), ("element", C.c_int), ("side", C.c_int), ("quantity", C.c_int),
                ("limit", C.c_int), ("value", C.c_double), ("threshold", C.c_double)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("total_iterations", C.c_int), ("solves", C.c_int),
                ("outer_switches", C.c_int), ("norm", C.c_double), ("slack_distribution_failed", C.c_int)]


class Ds(C.Structure):
    _fields_ = [("n", C.c_int), ("bus", C.c_void_p), ("factor", C.c_void_p), ("base_target_p", C.c_void_p),
                ("max_delta", C.c_void_p), ("min_delta", C.c_void_p), ("slack_bus", C.c_int),
                ("slack_target_p", C.c_double), ("tol", C.c_double), ("reject_failure", C.c_int)]


def ds_struct(ds, tol, reject_failure=False):
    """An oracle_ds over a distributed-slack dict (gridgen.distributed_slack); returns (struct, kept arrays)."""
    keep = [np.ascontiguousarray(ds["bus"], np.int32)] + [
        np.ascontiguousarray(ds[k], np.float64) for k in ("factor", "base_target_p", "max_delta", "min_delta")]
    st = Ds(len(keep[0]), *(a.ctypes.data for a in keep), int(ds["slack_bus"]), float(ds["slack_target_p"]),
            float(tol), int(reject_failure))
    return st, keep


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.oracle_closed_branch.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.oracle_sys_create.restype = C.c_void_p
        L.oracle_sys_create.argtypes = [C.c_int, C.c_int, _dp, _ip]
        L.oracle_sys_free.argtypes = [C.c_void_p]
        L.oracle_sys_set_shunts.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_sys_nnz.argtypes = [C.c_void_p]
        L.oracle_sys_pattern.argtypes = [C.c_void_p, _ip, _ip]
        L.oracle_fill.argtypes = [C.c_void_p, _dp, _ip, _dp, _ip, C.c_int, _dp, _dp, C.c_void_p]
        L.oracle_lu_solve.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp]
        L.oracle_min_degree.argtypes = [C.c_int, _ip, _ip, _ip]
        L.oracle_nr_solve.argtypes = [C.c_void_p, C.POINTER(NrParams), C.POINTER(Rl), C.c_void_p, _ip, C.c_int,
                                      C.c_void_p, _dp, _ip, _dp, _bp, _bp, C.POINTER(Result)]
        L.oracle_bus_p.restype = C.c_double
        L.oracle_bus_p.argtypes = [C.c_void_p, _dp, _ip, C.c_int, C.c_int]
        L.oracle_islands.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int, _ubp]
        L.oracle_violations.restype = C.c_int
        L.oracle_violations.argtypes = [C.c_void_p, _dp, _ip, C.c_int, C.c_void_p, C.POINTER(Limits),
                                        C.POINTER(Violation), C.c_int]
        L.oracle_n1.restype = C.c_int
        L.oracle_n1.argtypes = [C.c_void_p, C.POINTER(NrParams), C.POINTER(Rl), C.c_void_p, C.c_int, _dp, _ip, _dp,
                                _bp, C.c_int, C.c_int, _ip, C.c_int, _dp, _ip, _ip, _ip, _dp, C.c_void_p,
                                C.c_void_p]
        _LIB = L
    return _LIB


def closed_branch(p8, v1, ph1, v2, ph2) -> np.ndarray:
    out = np.zeros(28)
    lib().oracle_closed_branch(np.ascontiguousarray(p8, dtype=np.float64), v1, ph1, v2, ph2, out)
    return out


def lu_solve(rowptr, colidx, vals, b):
    n = len(rowptr) - 1
    x = np.zeros(n)
    rc = lib().oracle_lu_solve(n, np.ascontiguousarray(rowptr, np.int32), np.ascontiguousarray(colidx, np.int32),
                               np.ascontiguousarray(vals, np.float64), np.ascontiguousarray(b, np.float64), x)
    if rc != 0:
        raise np.linalg.LinAlgError("singular")
    return x


def min_degree(adjptr, adj):
    n = len(adjptr) - 1
    perm = np.zeros(n, np.int32)
    lib().oracle_min_degree(n, np.ascontiguousarray(adjptr, np.int32), np.ascontiguousarray(adj, np.int32), perm)
    return perm


def params(max_iter=15, tol=1e-4, reactive_limits=True, rl_tol=1e-4, max_outer=20, max_switch=3,
           scaling_mode=0, max_dv=0.1, max_dphi=0.174533, batch_of_one=True, ls_max_iteration=10,
           ls_step_fold=4.0 / 3.0, conv_eps_per_eq=0.0, n_active=0, row_thresholds=None,
           lin_fail_ratio=0.0) -> NrParams:
    """scaling_mode: 0 NONE, 1 LINE_SEARCH, 2 MAX_VOLTAGE_CHANGE; conv_eps_per_eq > 0: the 2-norm criterion;
    row_thresholds = (fixed [n], equation [n]): the per-row criterion (every |F_i| below its mode's threshold);
    tol <= 0 without either: the fixed-iteration mode; lin_fail_ratio > 0: the linear-solve guard's bound"""
    p = NrParams(max_iter, tol, int(reactive_limits), rl_tol, max_outer, max_switch, scaling_mode, max_dv,
                 max_dphi, int(batch_of_one), ls_max_iteration, ls_step_fold, conv_eps_per_eq, n_active, None, None,
                 lin_fail_ratio)
    if row_thresholds is not None:
        p._rows = tuple(np.ascontiguousarray(a, np.float64) for a in row_thresholds)
        p.row_thr_fixed, p.row_thr_eq = p._rows[0].ctypes.data, p._rows[1].ctypes.data
    return p


class OracleSystem:
    """The oracle's stable-row system for one grid (gridgen.Network)."""

    def __init__(self, net):
        self.net = net
        self.n = net.n
        self.p8 = np.ascontiguousarray(net.pi_model())
        self.idx = np.ascontiguousarray(net.branch_indices())
        self.h = lib().oracle_sys_create(net.n_bus, net.n_branch, self.p8, self.idx)
        if getattr(net, "shunt_g", None) is not None:
            self._shunts = (np.ascontiguousarray(net.shunt_g, np.float64), np.ascontiguousarray(net.shunt_b, np.float64))
            lib().oracle_sys_set_shunts(self.h, self._shunts[0].ctypes.data, self._shunts[1].ctypes.data)
        self.nnz = lib().oracle_sys_nnz(self.h)
        self.rowptr = np.zeros(self.n + 1, np.int32)
        self.colidx = np.zeros(self.nnz, np.int32)
        lib().oracle_sys_pattern(self.h, self.rowptr, self.colidx)
        rl = net.reactive_limits()
        self._rl_arrays = {k: np.ascontiguousarray(v) for k, v in rl.items()}
        a = self._rl_arrays
        self.rl = Rl(len(a["bus"]), a["voltage_row"].ctypes.data, a["bus"].ctypes.data, a["min_q"].ctypes.data,
                     a["max_q"].ctypes.data, a["load_q"].ctypes.data, a["target_v"].ctypes.data)

    def __del__(self):
        try:
            if self.h:
                lib().oracle_sys_free(self.h)
                self.h = None
        except Exception:
            pass

    def fill(self, x, mode, target, outages=(), jacobian=True):
        F = np.zeros(self.n)
        q = np.zeros(self.net.n_bus)
        J = np.zeros(self.nnz) if jacobian else None
        out = np.ascontiguousarray(np.asarray(outages, dtype=np.int32).reshape(-1))
        lib().oracle_fill(self.h, np.ascontiguousarray(x, np.float64), np.ascontiguousarray(mode, np.int32),
                          np.ascontiguousarray(target, np.float64), out, len(out), F, q,
                          J.ctypes.data if J is not None else None)
        return F, q, J

    def bus_p(self, x, bus, outages=()):
        """The computed active injection of one bus (its in-service branches' p at its side), pu."""
        out = np.ascontiguousarray(np.asarray(outages, dtype=np.int32).reshape(-1))
        return lib().oracle_bus_p(self.h, np.ascontiguousarray(x, np.float64), out, len(out), int(bus))

    def solve(self, x0, mode, target, limit_type=None, prm: NrParams | None = None, outages=(), islanded=None,
              ds=None, ds_tol=0.0, ds_reject=False, switch_counts=False):
        """One load flow (distributed slack when ds is a gridgen.distributed_slack dict); returns
        (x, mode, target, limit_type, Result), and the reactive-limit entries' PV -> PQ switch counts as a sixth
        item when switch_counts is set."""
        prm = prm or params()
        x = np.ascontiguousarray(x0, np.float64).copy()
        mode = np.ascontiguousarray(mode, np.int32).copy()
        target = np.ascontiguousarray(target, np.float64).copy()
        nrl = self.rl.n
        lt = np.zeros(max(nrl, 1), np.int8) if limit_type is None else np.ascontiguousarray(limit_type, np.int8).copy()
        sc = np.zeros(max(nrl, 1), np.int8)
        out = np.ascontiguousarray(np.asarray(outages, dtype=np.int32).reshape(-1))
        isl = None if islanded is None else np.ascontiguousarray(islanded, np.uint8)
        res = Result()
        dst, dkeep = ds_struct(ds, ds_tol, ds_reject) if ds is not None else (None, None)
        lib().oracle_nr_solve(self.h, C.byref(prm), C.byref(self.rl), C.byref(dst) if dst is not None else None,
                              out, len(out), isl.ctypes.data if isl is not None else None, x, mode, target, lt, sc,
                              C.byref(res))
        if switch_counts:
            return x, mode, target, lt[:nrl], res, sc[:nrl].copy()
        return x, mode, target, lt[:nrl], res

    def islands(self, outages, slack=0):
        out = np.ascontiguousarray(np.asarray(outages, dtype=np.int32).reshape(-1))
        isl = np.zeros(self.net.n_bus, np.uint8)
        lib().oracle_islands(self.h, slack, out, len(out), isl)
        return isl.astype(bool)

    def n1(self, x_base, mode_base, target_base, limit_type_base, contingencies, prm: NrParams | None = None,
           threads=0, slack=0, ds=None, ds_tol=0.0, ds_reject=False, want_targets=False, x0=None):
        """Batched N-1 on the CPU. contingencies: [K] or [K, max_out] branch ids (-1 pad). With ds (a
        gridgen.distributed_slack dict) the distributed-slack loop runs first; want_targets returns each
        contingency's final targets; x0 [K][n] starts each contingency from its own state instead of x_base."""
        prm = prm or params(batch_of_one=False)
        ctg = np.asarray(contingencies, dtype=np.int32)
        if ctg.ndim == 1:
            ctg = ctg[:, None]
        ctg = np.ascontiguousarray(ctg)
        k, mo = ctg.shape
        x = np.zeros((k, self.n))
        st = np.zeros(k, np.int32)
        it = np.zeros(k, np.int32)
        tot = np.zeros(k, np.int32)
        nm = np.zeros(k)
        lt = np.ascontiguousarray(np.asarray(limit_type_base, dtype=np.int8).reshape(-1))
        if lt.size == 0:
            lt = np.zeros(1, np.int8)
        tg = np.zeros((k, self.n)) if want_targets else None
        dst, dkeep = ds_struct(ds, ds_tol, ds_reject) if ds is not None else (None, None)
        x0p = None
        if x0 is not None:
            x0p = np.ascontiguousarray(x0, np.float64)
            if x0p.shape != (k, self.n):
                raise ValueError(f"x0 must be [{k}][{self.n}]")
        used = lib().oracle_n1(self.h, C.byref(prm), C.byref(self.rl), C.byref(dst) if dst is not None else None,
                               slack, np.ascontiguousarray(x_base, np.float64),
                               np.ascontiguousarray(mode_base, np.int32), np.ascontiguousarray(target_base, np.float64),
                               lt, k, mo, ctg, threads, x, st, it, tot, nm, tg.ctypes.data if tg is not None else None,
                               x0p.ctypes.data if x0p is not None else None)
        out = dict(x=x, status=st, iterations=it, total_iterations=tot, norm=nm, threads=used)
        if want_targets:
            out["target"] = tg
        return out


def limits_struct(table):
    """An oracle_limits over a limit-table dict (gridgen.limit_table); returns (struct, kept arrays)."""
    keep = {k: np.ascontiguousarray(table[k], np.int32) for k in
            ("branch", "group_start", "group_side", "group_quantity", "limit_start")}
    keep.update({k: np.ascontiguousarray(table[k], np.float64) for k in ("threshold", "bus_high", "bus_low")})
    keep.update({k: np.ascontiguousarray(table.get(k, ()), np.int32) for k in ("angle_from", "angle_to")})
    keep.update({k: np.ascontiguousarray(table.get(k, ()), np.float64) for k in ("angle_high", "angle_low")})
    st = Limits(len(keep["branch"]), keep["branch"].ctypes.data, keep["group_start"].ctypes.data,
                keep["group_side"].ctypes.data, keep["group_quantity"].ctypes.data, keep["limit_start"].ctypes.data,
                keep["threshold"].ctypes.data, len(keep["bus_high"]), keep["bus_high"].ctypes.data,
                keep["bus_low"].ctypes.data, len(keep["angle_from"]), keep["angle_from"].ctypes.data,
                keep["angle_to"].ctypes.data, keep["angle_high"].ctypes.data, keep["angle_low"].ctypes.data)
    return st, keep


def violations(system, x, table, outages=(), islanded=None, max_out=100000):
    """OLF's detection scan on one state: list of (kind, element, side, quantity, limit, value, threshold)."""
    st, keep = limits_struct(table)
    out = (Violation * max_out)()
    ob = np.ascontiguousarray(np.asarray(outages, dtype=np.int32).reshape(-1))
    isl = None if islanded is None else np.ascontiguousarray(islanded, np.uint8)
    n = lib().oracle_violations(system.h, np.ascontiguousarray(x, np.float64), ob, len(ob),
                                isl.ctypes.data if isl is not None else None, C.byref(st), out, max_out)
    if n > max_out:
        raise RuntimeError("more violations than max_out")
    return [(v.kind, v.element, v.side, v.quantity, v.limit, v.value, v.threshold) for v in out[:n]]
