"""Pins of the oracle's Newton-Raphson: a two-bus closed form, the independent Y-bus mismatch at the solution,
the power balance, the reactive-limit invariants, islands against bridges, and the status rules the reference
states (newton_iteration.cu:276-397)."""
import numpy as np
import pytest

import pyoracle as po
from olfgpu import gridgen


def ybus_mismatch(net, x, mode, target, outages=(), islanded=None):
    """Independent mismatch: S = V conj(Y V) from a complex admittance matrix built from r, x, b, ratio."""
    n_bus = net.n_bus
    Y = np.zeros((n_bus, n_bus), complex)
    for e in range(net.n_branch):
        if e in set(outages):
            continue
        a, b = net.bus1[e], net.bus2[e]
        ys = 1 / complex(net.r[e], net.x[e])
        t = net.ratio[e]
        bc = 1j * net.b[e] / 2
        Y[a, a] += (ys + bc) * t * t
        Y[b, b] += ys + bc
        Y[a, b] -= ys * t
        Y[b, a] -= ys * t
    V = x[1::2] * np.exp(1j * x[0::2])
    S = V * np.conj(Y @ V)
    F = np.empty(net.n)
    F[0::2] = S.real - target[0::2]
    F[1::2] = S.imag - target[1::2]
    fixed = mode == 0
    F[fixed] = x[fixed] - target[fixed]
    return F


def test_two_bus_closed_form():
    # slack 1<0 feeding a PQ load through one line: |V2|^2 solves a quadratic (textbook)
    r, x = 0.02, 0.1
    P, Q = 0.8, 0.3            # load, pu
    net = gridgen.Network(2, np.array([0], np.int32), np.array([1], np.int32), np.array([r]), np.array([x]),
                          np.array([0.0]), np.array([1.0]), np.array([2, 0], np.int8), np.array([1.0, 1.0]),
                          np.zeros(2), np.zeros(2), np.array([0, P * gridgen.SB]), np.array([0, Q * gridgen.SB]),
                          np.zeros(2))
    osys = po.OracleSystem(net)
    x0, tgt, mode, _ = net.system()
    xs, _, _, _, res = osys.solve(x0, mode, tgt, prm=po.params(tol=1e-12, max_iter=20))
    assert po.STATUS[res.status] == "CONVERGED"
    # closed form: V2^4 + (2(rP + xQ) - 1) V2^2 + (r^2 + x^2)(P^2 + Q^2) = 0, high-voltage root
    a, bq, c = 1.0, 2 * (r * P + x * Q) - 1, (r * r + x * x) * (P * P + Q * Q)
    u = (-bq + np.sqrt(bq * bq - 4 * a * c)) / 2
    assert abs(xs[3] - np.sqrt(u)) < 1e-10
    # angle: V2 = V1 - z I with I = conj(S_load / V2) (the load current)
    V2 = xs[3] * np.exp(1j * xs[2])
    assert abs(V2 - (1.0 - complex(r, x) * np.conj(complex(P, Q) / V2))) < 1e-10


def test_flat_start_118_meets_tolerance_and_invariants(grid):
    net = grid("grid118")
    osys = po.OracleSystem(net)
    x0, tgt, mode, _ = net.system()
    xs, ms, ts, lt, res = osys.solve(x0, mode, tgt, prm=po.params(max_iter=15))
    assert po.STATUS[res.status] == "CONVERGED"
    assert res.norm < 1e-4
    F = ybus_mismatch(net, xs, ms, ts)
    assert np.abs(F).max() < 1e-4
    # power balance: total injection = series + shunt losses = sum over branches of p1 + p2 >= 0
    Fp, q, _ = osys.fill(xs, np.ones(net.n, np.int32), np.zeros(net.n), jacobian=False)
    losses = Fp[0::2].sum()
    assert losses > 0
    # reactive-limit invariants: still-PV buses inside their limits (+tol), switched buses at their limit
    rl = net.reactive_limits()
    qgen = q[rl["bus"]] + rl["load_q"]
    pv = ms[rl["voltage_row"]] == 0
    assert np.all(qgen[pv] >= rl["min_q"][pv] - 1e-4) and np.all(qgen[pv] <= rl["max_q"][pv] + 1e-4)
    sw = ~pv
    lim = np.where(lt[sw] == 1, rl["min_q"][sw], rl["max_q"][sw])
    assert np.allclose(qgen[sw], lim, atol=2e-4)


def test_islands_follow_bridges(grid):
    net = grid("grid2000")
    osys = po.OracleSystem(net)
    br = gridgen.bridges(net)
    rng = np.random.default_rng(0)
    for e in rng.choice(np.nonzero(br)[0], 10, replace=False):
        assert osys.islands([e]).sum() >= 1
    for e in rng.choice(np.nonzero(~br)[0], 10, replace=False):
        assert osys.islands([e]).sum() == 0


def test_statuses_forced_from_the_host(grid):
    net = grid("grid118")
    osys = po.OracleSystem(net)
    x0, tgt, mode, _ = net.system()
    bad = x0.copy()
    bad[5] = np.nan
    assert po.STATUS[osys.solve(bad, mode, tgt)[4].status] == "DIVERGED"
    # budget of one iteration: the first test happens at it == maxIter
    r1 = osys.solve(x0, mode, tgt, prm=po.params(max_iter=1, reactive_limits=False))[4]
    assert po.STATUS[r1.status] == "MAX_ITERATION_REACHED" and r1.iterations == 1
    # an already-solved state still takes one Newton step (>= 1 iteration)
    xs, ms, ts, lt, _ = osys.solve(x0, mode, tgt)
    r2 = osys.solve(xs, ms, ts, lt, prm=po.params(reactive_limits=False))[4]
    assert po.STATUS[r2.status] == "CONVERGED" and r2.iterations == 1


def test_n1_results_solve_the_post_contingency_grid(grid):
    net = grid("grid2000")
    osys = po.OracleSystem(net)
    x0, tgt, mode, _ = net.system()
    xb, mb, tb, ltb, res = osys.solve(x0, mode, tgt)
    assert po.STATUS[res.status] == "CONVERGED"
    ctg = gridgen.sample_contingencies(net, 16, 0.25, 9)
    out = osys.n1(xb, mb, tb, ltb, ctg)
    br = gridgen.bridges(net)
    for k, e in enumerate(ctg):
        assert out["status"][k] == 0
        isl = osys.islands([e])
        xs = out["x"][k]
        if br[e]:
            assert isl.any()
            assert np.allclose(xs[0::2][isl], 0.0, atol=1e-9) and np.allclose(xs[1::2][isl], 1.0, atol=1e-9)
        F = ybus_mismatch(net, xs, np.ones(net.n, np.int32), np.zeros(net.n), outages=[e])
        # main component P rows (non-slack) meet their targets
        main = ~isl
        main[0] = False
        assert np.abs(F[0::2][main] - tb[0::2][main]).max() < 1e-4


def test_n1_per_scenario_starts(grid):
    """oracle_n1's per-scenario start (x0): the base state repeated gives bitwise the base-start run; each
    contingency restarted from its own solved state converges after one Newton step to the same answer (within the
    tolerance's reach)."""
    net = grid("grid2000")
    osys = po.OracleSystem(net)
    x0, tgt, mode, _ = net.system()
    xb, mb, tb, ltb, res = osys.solve(x0, mode, tgt)
    ctg = gridgen.sample_contingencies(net, 12, 0.25, 19)
    a = osys.n1(xb, mb, tb, ltb, ctg)
    b = osys.n1(xb, mb, tb, ltb, ctg, x0=np.tile(xb, (len(ctg), 1)))
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["total_iterations"], b["total_iterations"])
    c = osys.n1(xb, mb, tb, ltb, ctg, x0=a["x"])
    ok = a["status"] == 0
    assert (c["status"][ok] == 0).all()
    assert np.abs(c["x"][ok] - a["x"][ok]).max() < 1e-3
    with pytest.raises(ValueError):
        osys.n1(xb, mb, tb, ltb, ctg, x0=np.tile(xb, (len(ctg) + 1, 1)))
