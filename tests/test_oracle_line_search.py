"""Pins of the oracle's LINE_SEARCH state-vector scaling (OLF's LineSearchStateVectorScaling as the reference's
batched Newton runs it, native/cuda/solver/newton_iteration.cu:497-530): after every step but a solve's first, a step
that did not lower ||F||_2 (or made F non-finite) is walked back to x_prev + mu dx, mu = (4/3)^-b, b = 1..10, the
mismatch and Jacobian refilled at each point, until ||F||_2 is below the last tested one.

* where no step ever raises ||F||_2, LINE_SEARCH is plain Newton, bit for bit;
* from a start far enough that full steps overshoot, an independent Newton loop (dense numpy solves over the oracle's
  fill, the same rule) reaches the same iterate sequence: same iteration count, states within 1e-10, and it differs
  from the undamped run."""
import numpy as np
import scipy.sparse as sp

import pyoracle as po
from olfgpu import gridgen


def _newton_with_line_search(o, x0, mode, target, max_iter=15, tol=1e-4, fold=4.0 / 3.0, max_bt=10):
    n = o.n
    x = x0.copy()
    it = 0
    last = None
    xprev = dx = None
    while True:
        F, _, J = o.fill(x, mode, target)
        if it > 0:
            ssq = float(F @ F)
            fin = bool(np.all(np.isfinite(F)))
            if last is not None and (not fin or ssq >= last):
                for b in range(1, max_bt + 1):
                    x = xprev + (1.0 / fold ** b) * dx
                    F, _, J = o.fill(x, mode, target)
                    ssq = float(F @ F)
                    fin = bool(np.all(np.isfinite(F)))
                    if fin and ssq < last:
                        break
            last = ssq
            norm = np.abs(F).max()
            if norm < tol and it < max_iter:
                return x, it, "CONVERGED"
            if it == max_iter:
                return x, it, "MAX_ITERATION_REACHED"
        A = sp.csr_matrix((J, o.colidx, o.rowptr), shape=(n, n)).toarray()
        dx = np.linalg.solve(A, -F)
        xprev = x.copy()
        x = x + dx
        it += 1


def test_plain_newton_when_no_step_raises_the_norm():
    net = gridgen.workload("grid118")
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    xa, _, _, _, ra = o.solve(x0, md, tg, prm=po.params(reactive_limits=False, scaling_mode=0))
    xb, _, _, _, rb = o.solve(x0, md, tg, prm=po.params(reactive_limits=False, scaling_mode=1))
    assert ra.status == rb.status == 0 and ra.iterations == rb.iterations
    np.testing.assert_array_equal(xa, xb)


def test_far_start_matches_an_independent_backtracking_loop():
    net = gridgen.workload("grid118")
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    rng = np.random.default_rng(2)
    x0 = x0.copy()
    x0[0::2] += rng.uniform(-0.1, 0.1, net.n_bus)      # a start from which full steps overshoot
    x0[1::2] *= rng.uniform(0.9, 1.1, net.n_bus)
    x0[0] = 0.0
    fixed = md == 0
    x0[fixed] = tg[fixed]
    xs, its, status = _newton_with_line_search(o, x0, md, tg)
    xl, _, _, _, rl = o.solve(x0, md, tg, prm=po.params(reactive_limits=False, scaling_mode=1))
    xn, _, _, _, rn = o.solve(x0, md, tg, prm=po.params(reactive_limits=False, scaling_mode=0))
    assert status == "CONVERGED" and po.STATUS[rl.status] == status
    assert rl.iterations == its
    np.testing.assert_allclose(xl, xs, rtol=0, atol=1e-10)
    # undamped Newton does not converge from there within the budget
    assert po.STATUS[rn.status] == "MAX_ITERATION_REACHED"
