"""Pins of the oracle's distributed slack (oracle_nr_solve with an oracle_ds; the reference's per-bus water-fill,
native/cuda/outerloop/distributed_slack.cu:131-195,215-600, inside OLF's outer-loop sweep):

* lossless closed form: with r = 0 no power is lost, so the total from the anchor is minus the net injection I of the
  targets, whatever the state; each participant must end at delta_i = -f_i I and the slack at mismatch -(1 - sum f) I;
* saturation: the water-fill's answer is delta_i = min(c_i, f_i t) with one level t (the classic water-filling
  equilibrium), found here independently by bisection, to the loop's own residue (1e-5 pu);
* a lossy grid: the final state passes the loop's own gate, moves stay in their rooms and unsaturated participants
  share one delta / f;
* failure: rooms too small for the imbalance leave a residual, flagged, and FAIL rejects the case;
* N-1: a participant cut off by its outaged branches keeps its frozen target, the others share the lost power."""
import dataclasses

import numpy as np
import pytest

import pyoracle as po
from olfgpu import gridgen

SB = gridgen.SB


def _lossless(net):
    return dataclasses.replace(net, r=np.zeros_like(net.r))


def _solve(o, net, ds, tol_mw=1.0, reject=False, outages=(), **kw):
    x0, tg, md, _ = net.system()
    prm = po.params(**kw)
    return o.solve(x0, md, tg, prm=prm, ds=ds, ds_tol=tol_mw / SB, ds_reject=reject, outages=outages)


def test_lossless_closed_form():
    net = _lossless(gridgen.workload("grid118"))
    o = po.OracleSystem(net)
    ds = gridgen.distributed_slack(net)
    x, mode, target, lt, res = _solve(o, net, ds, tol_mw=0.01, reactive_limits=False, tol=1e-9)
    assert res.status == 0
    _, tg, _, _ = net.system()
    p_targets = tg[0::2].copy()
    p_targets[ds["slack_bus"]] = ds["slack_target_p"]
    imbalance = p_targets.sum()
    assert abs(imbalance) > 1e-3           # the 2 % surplus is there to distribute
    delta = target[2 * ds["bus"]] - ds["base_target_p"]
    np.testing.assert_allclose(delta, -ds["factor"] * imbalance, rtol=0, atol=1e-8)
    rho = 1 - ds["factor"].sum()
    m = o.bus_p(x, ds["slack_bus"]) - ds["slack_target_p"]
    assert abs(m + rho * imbalance) < 1e-8


def _waterfill(factor, lo, hi, total):
    """delta_i = clip(f_i t, lo_i, hi_i) with sum = total, by bisection on t"""
    a, b = -1e6, 1e6
    for _ in range(200):
        t = 0.5 * (a + b)
        s = np.clip(factor * t, lo, hi).sum()
        if s < total:
            a = t
        else:
            b = t
    return np.clip(factor * 0.5 * (a + b), lo, hi)


def test_saturated_water_fill_matches_bisection():
    net = _lossless(gridgen.workload("grid118"))
    o = po.OracleSystem(net)
    ds = gridgen.distributed_slack(net)
    # a load step the participants must take up, with some rooms cut so that they saturate
    net2 = dataclasses.replace(net, pload=net.pload * 1.0)
    net2.pload[5] += 300.0
    ds = gridgen.distributed_slack(net2, target=net.system()[1])
    rng = np.random.default_rng(4)
    cut = rng.random(len(ds["bus"])) < 0.4
    ds["max_delta"][cut] = ds["max_delta"][cut] * 0.02
    x0, tg, md, _ = net2.system()
    x, mode, target, lt, res = o.solve(x0, md, tg, prm=po.params(reactive_limits=False, tol=1e-9), ds=ds,
                                       ds_tol=0.01 / SB)
    assert res.status == 0 and res.slack_distribution_failed == 0
    p_targets = tg[0::2].copy()
    p_targets[ds["slack_bus"]] = ds["slack_target_p"]
    anchor_imbalance = p_targets.sum() - (tg[2 * ds["bus"]] - ds["base_target_p"]).sum()
    share = ds["factor"].sum() * -anchor_imbalance
    want = _waterfill(ds["factor"], ds["min_delta"], ds["max_delta"], share)
    delta = target[2 * ds["bus"]] - ds["base_target_p"]
    assert np.sum(delta >= ds["max_delta"] - 1e-15) > 3          # saturation happened
    # the water-fill stops once less than P_RESIDUE_EPS = 1e-5 pu is left to place
    assert abs(delta.sum() - share) <= 1e-5
    np.testing.assert_allclose(delta, want, rtol=0, atol=1e-5)
    np.testing.assert_array_equal(delta >= ds["max_delta"] - 1e-15, want >= ds["max_delta"] - 1e-12)


def test_lossy_grid_ends_inside_the_gate():
    net = gridgen.workload("grid2000")
    o = po.OracleSystem(net)
    ds = gridgen.distributed_slack(net)
    x, mode, target, lt, res = _solve(o, net, ds)
    assert res.status == 0 and res.outer_switches > 0
    delta = target[2 * ds["bus"]] - ds["base_target_p"]
    assert np.all(delta <= ds["max_delta"] + 1e-15) and np.all(delta >= ds["min_delta"] - 1e-15)
    free = (delta < ds["max_delta"]) & (delta > ds["min_delta"])
    ratio = delta[free] / ds["factor"][free]
    assert np.ptp(ratio) <= 1e-12 * max(1.0, np.abs(ratio).max())
    m = o.bus_p(x, ds["slack_bus"]) - ds["slack_target_p"]
    fs = ds["factor"].sum()
    rem = m - (1 - fs) * (m + delta.sum())
    assert abs(rem) <= fs * 1.0 / SB


def test_failure_flag_and_rejection():
    net = gridgen.workload("grid118")
    net.pload[7] += 400.0                  # more than the participants' rooms
    o = po.OracleSystem(net)
    ds = gridgen.distributed_slack(net, headroom=1.01)
    x, mode, target, lt, res = _solve(o, net, ds)
    assert res.status == 0 and res.slack_distribution_failed == 1
    delta = target[2 * ds["bus"]] - ds["base_target_p"]
    np.testing.assert_allclose(delta, ds["max_delta"], rtol=0, atol=1e-15)   # everyone at its limit
    x, mode, target, lt, res = _solve(o, net, ds, reject=True)
    assert po.STATUS[res.status] == "REJECTED"


def test_islanded_participants_keep_their_targets():
    net = gridgen.workload("grid2000")
    o = po.OracleSystem(net)
    ds = gridgen.distributed_slack(net)
    x0, tg, md, _ = net.system()
    xb, mb, tb, ltb, res = o.solve(x0, md, tg, ds=ds, ds_tol=1.0 / SB)
    assert res.status == 0
    # cut one participant off: every branch at a PV bus of degree <= 4
    deg = np.bincount(np.concatenate([net.bus1, net.bus2]), minlength=net.n_bus)
    bus = int(next(b for b in ds["bus"] if deg[b] <= 4))
    outs = np.nonzero((net.bus1 == bus) | (net.bus2 == bus))[0].astype(np.int32)
    ctg = np.full((1, 4), -1, np.int32)
    ctg[0, :len(outs)] = outs
    r = o.n1(xb, mb, tb, ltb, ctg, ds=ds, ds_tol=1.0 / SB, want_targets=True)
    assert r["status"][0] == 0
    assert o.islands(outs)[bus]
    i = int(np.nonzero(ds["bus"] == bus)[0][0])
    t_after = r["target"][0][2 * ds["bus"]]
    assert t_after[i] == 0.0                       # the frozen row keeps the island's installed angle target
    others = np.arange(len(ds["bus"])) != i
    moved = t_after[others] - tb[2 * ds["bus"]][others]
    assert np.abs(moved).max() > 1e-4              # the lost generation was taken up by the others
    free = (t_after[others] - ds["base_target_p"][others] < ds["max_delta"][others])
    ratio = (t_after[others] - ds["base_target_p"][others])[free] / ds["factor"][others][free]
    assert np.ptp(ratio) <= 1e-12 * max(1.0, np.abs(ratio).max())
