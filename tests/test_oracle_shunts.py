"""Pins of the oracle's bus-shunt family (p = g v^2, q = -b v^2 at the bus, reference
equations/generated/shunt.cuh:47-60; the workload has no shunts, Network.with_shunts adds them):

* the mismatch with shunts minus the one without is exactly (g v^2, -b v^2) per bus at any state;
* the Jacobian with shunts matches central finite differences of the mismatch, entry by entry;
* at the solved state the net injection equals the losses: branch losses plus shunt conductance g v^2."""
import numpy as np
import scipy.sparse as sp

import pyoracle as po
from olfgpu import gridgen


def test_shunt_terms_and_jacobian():
    base = gridgen.workload("grid118")
    net = base.with_shunts(0.2, seed=3)
    o0, o1 = po.OracleSystem(base), po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    rng = np.random.default_rng(5)
    x = x0.copy()
    x[0::2] += rng.uniform(-0.2, 0.2, net.n_bus)
    x[1::2] *= rng.uniform(0.9, 1.1, net.n_bus)
    mode = np.ones_like(md)                          # every row an equation: the raw injections
    F0, q0, _ = o0.fill(x, mode, np.zeros_like(tg))
    F1, q1, J1 = o1.fill(x, mode, np.zeros_like(tg))
    v = x[1::2]
    np.testing.assert_allclose(F1[0::2] - F0[0::2], net.shunt_g * v * v, rtol=0, atol=1e-12)
    np.testing.assert_allclose(F1[1::2] - F0[1::2], -net.shunt_b * v * v, rtol=0, atol=1e-12)
    np.testing.assert_allclose(q1 - q0, -net.shunt_b * v * v, rtol=0, atol=1e-12)
    n = o1.n
    Jd = sp.csr_matrix((J1, o1.colidx, o1.rowptr), shape=(n, n)).toarray()
    h = 1e-6
    for c in range(n):
        xp, xm = x.copy(), x.copy()
        xp[c] += h
        xm[c] -= h
        col = (o1.fill(xp, mode, np.zeros_like(tg), jacobian=False)[0]
               - o1.fill(xm, mode, np.zeros_like(tg), jacobian=False)[0]) / (2 * h)
        np.testing.assert_allclose(Jd[:, c], col, rtol=0, atol=2e-6)


def test_power_balance_with_shunts():
    net = gridgen.workload("grid2000").with_shunts(0.05)
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    x, mode, target, lt, res = o.solve(x0, md, tg)
    assert res.status == 0
    p_inj, _, _ = o.fill(x, np.ones_like(mode), np.zeros_like(target), jacobian=False)
    branch_losses = 0.0
    for e in range(net.n_branch):
        a, b = int(net.bus1[e]), int(net.bus2[e])
        out = po.closed_branch(o.p8[e], x[2 * a + 1], x[2 * a], x[2 * b + 1], x[2 * b])
        branch_losses += out[0] + out[14]
    shunt_p = (net.shunt_g * x[1::2] ** 2).sum()
    assert abs(p_inj[0::2].sum() - (branch_losses + shunt_p)) < 1e-9
    assert shunt_p > 0
