"""Pins of the oracle's fixed-iteration mode (tol <= 0 with no other criterion; the reference's device fixed-iteration
mode, native/cuda/solver/outer_loop_driver.cu:165-175, newton_iteration.cu:141-151,336-343): exactly max_iter Newton
steps, no convergence test, no linear-solve guard, no outer loop; MAX_ITERATION_REACHED, or DIVERGED when the final
mismatch is not finite.

* the state after k fixed steps is that of an independent loop of k dense Newton steps (numpy solves over the
  oracle's fill), for several k, including steps past convergence;
* with reactive limits on, no PV bus switches (no outer loop runs), while the converging run of the same case does;
* a non-finite start ends DIVERGED after the budget, a finite one MAX_ITERATION_REACHED, even when it has converged;
* the guard bound (lin_fail_ratio) that fails every normal solve at its 2nd iteration leaves the fixed mode alone."""
import numpy as np
import scipy.sparse as sp

import pyoracle as po
from olfgpu import gridgen


def dense_steps(o, x0, mode, target, k):
    x = x0.copy()
    for _ in range(k):
        F, _, J = o.fill(x, mode, target)
        A = sp.csr_matrix((J, o.colidx, o.rowptr), shape=(o.n, o.n)).toarray()
        x = x + np.linalg.solve(A, -F)
    return x


def test_fixed_steps_match_a_dense_loop():
    net = gridgen.workload("grid118")
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    for k in (1, 2, 3, 6):
        x, _, _, _, r = o.solve(x0, md, tg, prm=po.params(tol=0.0, max_iter=k))
        assert po.STATUS[r.status] == "MAX_ITERATION_REACHED" and r.iterations == k
        np.testing.assert_allclose(x, dense_steps(o, x0, md, tg, k), rtol=0, atol=1e-10)
        F, _, _ = o.fill(x, md, tg, jacobian=False)
        assert abs(r.norm - np.abs(F).max()) <= 1e-12 * max(1.0, r.norm)


def test_no_outer_loop_and_no_guard():
    net = gridgen.workload("grid2000")
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    _, mc, _, ltc, rc = o.solve(x0, md, tg)
    assert po.STATUS[rc.status] == "CONVERGED" and (ltc != 0).any()
    x, mf, _, ltf, rf = o.solve(x0, md, tg, prm=po.params(tol=0.0, max_iter=8, lin_fail_ratio=1e-300))
    assert po.STATUS[rf.status] == "MAX_ITERATION_REACHED" and rf.iterations == 8 and rf.solves <= 1
    assert (ltf == 0).all() and np.array_equal(mf, md)
    # the same bound fails a normal solve at its second iteration
    r = o.solve(x0, md, tg, prm=po.params(max_iter=8, lin_fail_ratio=1e-300))[4]
    assert po.STATUS[r.status] == "LINEAR_SOLVE_FAILED" and r.iterations == 1


def test_non_finite_start_ends_diverged_after_the_budget():
    net = gridgen.workload("grid118")
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    bad = x0.copy()
    bad[7] = np.nan
    r = o.solve(bad, md, tg, prm=po.params(tol=0.0, max_iter=3))[4]
    assert po.STATUS[r.status] == "DIVERGED" and r.iterations == 3
