"""Pins of the oracle's violation scan (OLF's detection rules, GpuLimitScanner.java:95-131,394-410): its flows against
the textbook pi-branch S = V conj(I), hand cases of each rule, and a brute-force rescan in plain Python on solved
N-1 states."""
import math

import numpy as np

import pyoracle as po
from olfgpu import gridgen


def _solved(net):
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    xb, mb, tb, lt, res = o.solve(x0, md, tg)
    assert po.STATUS[res.status] == "CONVERGED"
    return o, xb, mb, tb, lt


def _zero_table(net, quantity, side):
    """Every branch one group of one zero threshold (or -1 for active power compared as |p|)."""
    nb = net.n_branch
    return dict(branch=np.arange(nb, dtype=np.int32), group_start=np.arange(nb + 1, dtype=np.int32),
                group_side=np.full(nb, side, np.int32), group_quantity=np.full(nb, quantity, np.int32),
                limit_start=np.arange(nb + 1, dtype=np.int32), threshold=np.full(nb, -1.0),
                bus_high=np.zeros(0), bus_low=np.zeros(0))


def test_flows_are_the_textbook_pi_branch_flows(grid):
    net = grid("grid118")
    o, xb, *_ = _solved(net)
    S1, S2, v1, v2 = gridgen.branch_flows(net, xb)
    for side, S, v in ((1, S1, v1), (2, S2, v2)):
        for qty, ref in ((0, np.abs(S) / v), (1, S.real), (2, np.abs(S))):
            got = po.violations(o, xb, _zero_table(net, qty, side))
            assert len(got) == net.n_branch
            val = np.array([g[5] for g in got])
            np.testing.assert_allclose(val, ref, rtol=1e-10, atol=1e-12)


def test_detection_rules_by_hand(grid):
    net = grid("grid118")
    o, xb, *_ = _solved(net)
    S1, S2, v1, v2 = gridgen.branch_flows(net, xb)
    i1 = abs(S1[0]) / v1[0]
    p1 = S1[3].real
    t = dict(branch=np.array([0, 3, 5], np.int32), group_start=np.array([0, 1, 3, 3], np.int32),
             group_side=np.array([1, 1, 2], np.int32), group_quantity=np.array([0, 1, 0], np.int32),
             limit_start=np.array([0, 3, 4, 5], np.int32),
             # group 0 (current of branch 0): most severe first; the 2nd is the first exceeded
             threshold=np.array([2 * i1, 0.9 * i1, 0.5 * i1,
                                 # group 1 (active power of branch 3): |p| compared, signed value reported
                                 0.5 * abs(p1),
                                 # group 2: never exceeded
                                 1e9]),
             bus_high=np.array([np.nan, xb[3] - 1e-3, np.nan]), bus_low=np.array([np.nan, np.nan, xb[5] + 1e-3]))
    got = po.violations(o, xb, t)
    assert [(g[0], g[1], g[2], g[3], g[4]) for g in got] == [(0, 0, 1, 0, 1), (0, 1, 1, 1, 3), (1, 1, 0, 0, -1),
                                                             (2, 2, 0, 0, -1)]
    assert math.isclose(got[0][5], i1, rel_tol=1e-12)
    assert math.isclose(got[1][5], p1, rel_tol=1e-12)            # signed
    # an outaged branch and an islanded bus are out of service
    isl = np.zeros(net.n_bus, np.uint8)
    isl[2] = 1
    got = po.violations(o, xb, t, outages=[0], islanded=isl)
    assert [(g[0], g[1]) for g in got] == [(0, 1), (1, 1)]


def _rescan(net, x, t, outages, islanded):
    S1, S2, v1, v2 = gridgen.branch_flows(net, x)
    out = []
    for b, e in enumerate(t["branch"]):
        if e in outages or islanded[net.bus1[e]] or islanded[net.bus2[e]]:
            continue
        for g in range(t["group_start"][b], t["group_start"][b + 1]):
            S, v = (S1[e], v1[e]) if t["group_side"][g] == 1 else (S2[e], v2[e])
            q = t["group_quantity"][g]
            value = abs(S) / v if q == 0 else (S.real if q == 1 else abs(S))
            compared = abs(value) if q == 1 else value
            for lim in range(t["limit_start"][g], t["limit_start"][g + 1]):
                if compared > t["threshold"][lim]:
                    out.append((0, b, int(t["group_side"][g]), int(q), lim, value))
                    break
    for i in range(len(t["bus_high"])):
        if islanded[i]:
            continue
        v = x[2 * i + 1]
        if v > t["bus_high"][i]:
            out.append((1, i, 0, 0, -1, v))
        if v < t["bus_low"][i]:
            out.append((2, i, 0, 0, -1, v))
    return out


def test_scan_matches_a_brute_force_rescan_on_n1_states(grid):
    net = grid("grid2000")
    o, xb, mb, tb, lt = _solved(net)
    t = gridgen.limit_table(net, xb)
    ctg = gridgen.sample_contingencies(net, 24, 0.25, 13)
    r = o.n1(xb, mb, tb, lt, ctg, threads=0)
    total = 0
    for j, e in enumerate(ctg):
        if r["status"][j] != 0:
            continue
        isl = o.islands([e])
        got = po.violations(o, r["x"][j], t, outages=[e], islanded=isl)
        ref = _rescan(net, r["x"][j], t, {int(e)}, isl)
        assert [g[:5] for g in got] == [q[:5] for q in ref], e
        np.testing.assert_allclose([g[5] for g in got], [q[5] for q in ref], rtol=1e-9)
        total += len(got)
    assert total > 20


def test_voltage_angle_limits_by_hand(grid):
    """The oracle's voltage-angle check (GpuLimitScanner.java:239-267): on a state with chosen angles, the difference
    to - from against high then low, NaN bounds never firing, a limit with an islanded bus skipped, after every bus;
    a swapped from / to or a strict comparison turned inclusive fails a case."""
    net = grid("grid118")
    x = net.system()[0].copy()
    x[0::2] = 0.0
    x[1::2] = 1.0                         # no bus limit fires (1.0 within every bus's band)
    x[2 * 3] = 0.30
    x[2 * 7] = -0.10
    table = _zero_table(net, 0, 1)
    table.update(branch=np.zeros(0, np.int32), group_start=np.zeros(1, np.int32), group_side=np.zeros(0, np.int32),
                 group_quantity=np.zeros(0, np.int32), limit_start=np.zeros(1, np.int32), threshold=np.zeros(0),
                 bus_high=np.full(net.n_bus, 1.1), bus_low=np.full(net.n_bus, 0.9),
                 angle_from=np.array([7, 3, 7, 3, 7, 7], np.int32), angle_to=np.array([3, 7, 3, 7, 3, 5], np.int32),
                 # to - from: 0.4, -0.4, 0.4, -0.4, 0.4 (exactly at the high bound), 0.1 (bus 5 islanded)
                 angle_high=np.array([0.35, np.nan, np.nan, 0.5, 0.4, 0.0]),
                 angle_low=np.array([np.nan, -0.3, 0.5, -0.45, 0.0, 0.2]))
    o = po.OracleSystem(net)
    isl = np.zeros(net.n_bus, np.uint8)
    isl[5] = 1
    got = po.violations(o, x, table, islanded=isl)
    want = [(3, 0, 0, 0, -1, 0.4, 0.35), (4, 1, 0, 0, -1, -0.4, -0.3), (4, 2, 0, 0, -1, 0.4, 0.5)]
    assert [g[:5] for g in got] == [w[:5] for w in want]
    np.testing.assert_allclose([g[5] for g in got], [w[5] for w in want], rtol=0, atol=1e-15)
    assert [g[6] for g in got] == [w[6] for w in want]
    # the same limits after a bus violation: buses first
    x[2 * 9 + 1] = 1.2
    got = po.violations(o, x, table, islanded=isl)
    assert got[0][:2] == (1, 9) and [g[0] for g in got[1:]] == [3, 4, 4]
