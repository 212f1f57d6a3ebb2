"""Pins of the oracle's reactive-limit outer loop beyond the PV -> PQ side (reference
native/cuda/outerloop/reactive_limits.cu:321-371, 410-460; OLF's ReactiveLimitsOuterLoop):

* the PQ -> PV switch-back: a bus held at its maximum reactive output returns to its voltage setpoint when its voltage
  rises above the setpoint (at its minimum: falls below), until it has switched PV -> PQ maxPqPvSwitch times. At a
  converged, stable outer loop every bus at a limit is therefore on the side of its setpoint that keeps it there, or
  at the switch cap;
* the forced-PV election: when a check would leave no PV bus, the strongest of the buses it switched (highest
  active-power target; equal nominal voltages here) stays PV, and only then.

An inverted switch-back comparison, a cap compared the wrong way or an election by the weakest bus fails one of
these."""
import numpy as np

import pyoracle as po
from olfgpu import gridgen

MIN, MAX = 1, 2


def four_bus(pgen1, pgen2, qmax1_mvar=1.0):
    """Slack 0 at 1.0 pu, PV buses 1 and 2 at 1.05 pu with a +-1 MVar reactive range (bus 1's upper limit
    qmax1_mvar), a 200 MW load on bus 3; holding 1.05 pu needs far more than 1 MVar, so a check switches both PV
    buses unless bus 1's range is wide."""
    return gridgen.Network(
        4, np.array([0, 1, 2, 3, 0], np.int32), np.array([1, 2, 3, 0, 2], np.int32), np.full(5, 0.01),
        np.full(5, 0.05), np.zeros(5), np.ones(5), np.array([gridgen.SLACK, gridgen.PV, gridgen.PV, gridgen.PQ], np.int8),
        np.array([1.0, 1.05, 1.05, 1.0]), np.array([0.0, -1.0, -1.0, 0.0]), np.array([0.0, qmax1_mvar, 1.0, 0.0]),
        np.array([0.0, 0.0, 0.0, 200.0]), np.array([0.0, 0.0, 0.0, 50.0]), np.array([0.0, pgen1, pgen2, 0.0]))


def solve(net, **kw):
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    x, mode, target, lt, res, sc = o.solve(x0, md, tg, prm=po.params(**kw), switch_counts=True)
    assert po.STATUS[res.status] == "CONVERGED"
    return x, mode, target, lt, sc


def test_forced_pv_keeps_the_strongest_bus():
    # both PV buses exceed their maximum at the first check: the one with the larger P target stays PV
    x, mode, _, lt, sc = solve(four_bus(50.0, 80.0))
    assert list(lt) == [MAX, 0]
    assert mode[3] == gridgen.EQUATION and mode[5] == gridgen.FIXED_VARIABLE
    assert abs(x[5] - 1.05) < 1e-12
    x, mode, _, lt, sc = solve(four_bus(80.0, 50.0))
    assert list(lt) == [0, MAX]
    assert mode[3] == gridgen.FIXED_VARIABLE and mode[5] == gridgen.EQUATION
    assert abs(x[3] - 1.05) < 1e-12


def test_no_election_while_a_pv_bus_remains():
    # bus 1 can hold its setpoint: the check leaves a PV bus, so the stronger bus 2 is switched like any other
    x, mode, _, lt, sc = solve(four_bus(50.0, 80.0, qmax1_mvar=4000.0))
    assert list(lt) == [0, MAX]
    assert mode[3] == gridgen.FIXED_VARIABLE and mode[5] == gridgen.EQUATION


def _seeded_at_limits(o, net):
    """The converged base case with a fifth of the buses that ended PV restarted switched to PQ at their maximum
    reactive output and a tenth at their minimum (entries already at a limit keep theirs): each of them then sits on
    the wrong side of its setpoint for staying there."""
    x0, tg, md, _ = net.system()
    xb, mb, tb, ltb, res = o.solve(x0, md, tg)
    assert po.STATUS[res.status] == "CONVERGED"
    rl = net.reactive_limits()
    md = mb.copy()
    tg = tb.copy()
    lt = ltb.copy()
    vr = rl["voltage_row"]
    u = np.random.default_rng(1).random(len(lt))
    for kind, sel, q in ((MAX, (lt == 0) & (u < 0.2), rl["max_q"]), (MIN, (lt == 0) & (u >= 0.2) & (u < 0.3),
                                                                    rl["min_q"])):
        md[vr[sel]] = gridgen.EQUATION
        tg[vr[sel]] = q[sel] - rl["load_q"][sel]
        lt[sel] = kind
    return xb, tg, md, lt, rl, lt != ltb


def test_switch_back_invariants_at_convergence(grid):
    net = grid("grid2000")
    o = po.OracleSystem(net)
    x0, tg, md, lt0, rl, seeded = _seeded_at_limits(o, net)
    x, mode, target, lt, res, sc = o.solve(x0, md, tg, lt0, prm=po.params(max_switch=3), switch_counts=True)
    assert po.STATUS[res.status] == "CONVERGED"
    v = x[rl["voltage_row"]]
    tv = rl["target_v"]
    cap = sc >= 3
    # held at a limit: on the side of the setpoint that keeps it there, unless capped
    assert np.all((v[lt == MAX] <= tv[lt == MAX]) | cap[lt == MAX])
    assert np.all((v[lt == MIN] >= tv[lt == MIN]) | cap[lt == MIN])
    # PV buses (at their setpoint) inside their reactive range (+ the band)
    _, q, _ = o.fill(x, np.ones(net.n, np.int32), np.zeros(net.n), jacobian=False)
    qg = q[rl["bus"]] + rl["load_q"]
    pv = lt == 0
    assert np.all(qg[pv] >= rl["min_q"][pv] - 1e-4) and np.all(qg[pv] <= rl["max_q"][pv] + 1e-4)
    np.testing.assert_allclose(v[pv], tv[pv], atol=1e-12)
    # the seeded buses went back to PV, at both limits
    assert (pv & seeded & (lt0 == MAX)).sum() >= 40 and (pv & seeded & (lt0 == MIN)).sum() >= 15


def test_switch_cap_blocks_the_switch_back(grid):
    # with a cap of zero no bus may leave its limit: seeded buses stay PQ on the far side of their setpoints
    net = grid("grid2000")
    o = po.OracleSystem(net)
    x0, tg, md, lt0, rl, seeded = _seeded_at_limits(o, net)
    x, mode, target, lt, res, sc = o.solve(x0, md, tg, lt0, prm=po.params(max_switch=0), switch_counts=True)
    assert po.STATUS[res.status] == "CONVERGED"
    # PV -> PQ switches still happen; nothing at a limit returns to PV
    assert np.array_equal(lt[seeded], lt0[seeded]) and not np.any((lt0 != 0) & (lt == 0))
    v = x[rl["voltage_row"]]
    tv = rl["target_v"]
    assert (seeded & (lt == MAX) & (v > tv)).sum() >= 40
    assert (seeded & (lt == MIN) & (v < tv)).sum() >= 15
