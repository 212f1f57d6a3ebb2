"""Comparison of a device N-1 result with the oracle's — TEST INFRASTRUCTURE ONLY.

Used by tests/ and by bench.py's CPU-baseline leg, as the checker. The bar of the path: per contingency the same
status; |V| within 1e-6 pu and angle within 1e-6 rad; the same violation list (kind, element, side, quantity, limit
in the same order). Nothing here is on the product path.
"""
from __future__ import annotations

import numpy as np

import pyoracle as po

V_TOL = 1e-6
ANGLE_TOL = 1e-6
_FIELDS = ("kind", "element", "side", "quantity", "limit")


def scan_records(osys, x, table, outages, scenario, dtype):
    """The oracle's OLF detection scan of one state, as records of `dtype` (olfgpu's VIOLATION_DTYPE)."""
    outs = [int(o) for o in np.atleast_1d(outages) if o >= 0]
    isl = osys.islands(outs) if outs else None
    rows = po.violations(osys, x, table, outages=outs, islanded=isl)
    return np.array([(scenario, q[0], q[1], q[2], q[3], q[4], q[5], q[6]) for q in rows], dtype)


def n1_scans(osys, xs, status, contingencies, table, dtype):
    """Scan records of every converged contingency (scenario = its index), in contingency order."""
    parts = [scan_records(osys, xs[j], table, contingencies[j], j, dtype)
             for j in range(len(status)) if status[j] == 0]
    return np.concatenate(parts) if parts else np.zeros(0, dtype)


def _lists(v, k):
    out = [[] for _ in range(k)]
    for r in v:
        out[int(r["scenario"])].append(tuple(int(r[f]) for f in _FIELDS))
    return out


def compare_violations(got, want, k):
    """Per-contingency comparison of two violation record arrays: contingencies whose lists differ (entries or
    order), entries in one list and not the other, and the totals."""
    g, w = _lists(got, k), _lists(want, k)
    differ = [j for j in range(k) if g[j] != w[j]]
    extra = sum(len(set(g[j]) ^ set(w[j])) for j in differ)
    return {"contingencies_differing": len(differ), "entries_differing": int(extra), "device_entries": int(len(got)),
            "oracle_entries": int(len(want)), "first_differing": differ[:5]}


def near_limit(records, rel):
    """Records whose value lies within `rel` (relative) of their threshold: a state difference of that order can
    flip them."""
    if len(records) == 0:
        return 0
    v = np.abs(records["value"])
    t = np.abs(records["threshold"])
    return int((np.abs(v - t) <= rel * t).sum())


def compare_states(x_dev, status_dev, ref, rows=None):
    """Statuses, states and iteration counts of device results against an oracle n1 result, over `rows` (all by
    default): mismatching statuses, max |dV| and |dtheta| over contingencies both converged, and how many of them
    are within the 1e-6 bar."""
    k = len(status_dev)
    rows = np.arange(k) if rows is None else np.asarray(rows)
    st_d = np.asarray(status_dev)[rows]
    st_o = ref["status"][rows]
    both = rows[(st_d == 0) & (st_o == 0)]
    out = {"contingencies": int(len(rows)), "status_mismatches": int((st_d != st_o).sum())}
    if both.size:
        dv = np.abs(x_dev[both, 1::2] - ref["x"][both, 1::2]).max(axis=1)
        da = np.abs(x_dev[both, 0::2] - ref["x"][both, 0::2]).max(axis=1)
        out.update({"max_abs_dv_pu": float(dv.max()), "max_abs_dtheta_rad": float(da.max()),
                    "within_1e-6": int(((dv <= V_TOL) & (da <= ANGLE_TOL)).sum()), "compared": int(both.size)})
    return out
