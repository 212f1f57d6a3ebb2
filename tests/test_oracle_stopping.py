"""Pin of the oracle's 2-norm stopping criterion (OLF's DefaultNewtonRaphsonStoppingCriteria as the reference's
batched Newton applies it, native/cuda/solver/newton_iteration.cu:241-300): with conv_eps_per_eq > 0 a solve stops
at the first iterate (after at least one step) with ||F||_2 < sqrt(conv_eps_per_eq^2 n_active). An independent
Newton loop (dense numpy solves over the oracle's fill) with that rule must stop at the same iteration and state,
for thresholds placed between consecutive iterates' norms, so that a squared norm, a wrong equation count or the
infinity norm would stop elsewhere."""
import numpy as np
import scipy.sparse as sp

import pyoracle as po
from olfgpu import gridgen


def _norms_and_states(o, x0, mode, target, iters):
    n = o.n
    x = x0.copy()
    out = []
    for _ in range(iters + 1):
        F, _, J = o.fill(x, mode, target)
        out.append((np.sqrt(F @ F), np.abs(F).max(), x.copy()))
        A = sp.csr_matrix((J, o.colidx, o.rowptr), shape=(n, n)).toarray()
        x = x + np.linalg.solve(A, -F)
    return out


def test_two_norm_criterion_stops_where_an_independent_loop_does():
    net = gridgen.workload("grid118")
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    hist = _norms_and_states(o, x0, md, tg, 5)
    n = o.n
    for k in (2, 3, 4):
        # a threshold between the 2-norms of iterates k-1 and k: the loop must stop at k
        thr = np.sqrt(hist[k][0] * hist[k - 1][0])
        eps = thr / np.sqrt(n)
        assert hist[k][0] < thr <= hist[k - 1][0]
        x, _, _, _, r = o.solve(x0, md, tg, prm=po.params(reactive_limits=False, conv_eps_per_eq=eps, tol=0.0))
        assert po.STATUS[r.status] == "CONVERGED" and r.iterations == k, (k, r.iterations)
        np.testing.assert_allclose(x, hist[k][2], rtol=0, atol=1e-10)
        # the same threshold over a smaller equation count is tighter: at least as many iterations
        _, _, _, _, r2 = o.solve(x0, md, tg, prm=po.params(reactive_limits=False, conv_eps_per_eq=eps, tol=0.0,
                                                           n_active=n // 4))
        assert r2.iterations >= k


def _abs_f(o, x, mode, target):
    return np.abs(o.fill(x, mode, target, jacobian=False)[0])


def test_per_row_criterion_stops_where_an_independent_loop_does():
    """OLF's PerEquationTypeStoppingCriteria as per-row thresholds (the reference's setBatchRowCriteria,
    native/cuda/solver/convergence.cu:28-60): the solve stops at the first iterate whose every row is below its
    threshold, the identity rows' threshold for FIXED_VARIABLE rows and the equation rows' for the others. Row
    thresholds of three times each equation row's |F| at iterate k (plus a floor) must stop it exactly at k; a
    huge identity-row threshold must not leak into the equation rows (it would stop at iterate 1); a per-row
    threshold read at the wrong row would miss the iterate."""
    net = gridgen.workload("grid118")
    o = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    hist = _norms_and_states(o, x0, md, tg, 5)
    eq = np.asarray(md) != 0
    for k in (2, 3):
        fk = _abs_f(o, hist[k][2], md, tg)
        thr_eq = 3.0 * fk + 1e-14
        thr_fixed = np.full(o.n, 1e9)
        # iterate k - 1 fails some equation row, iterate k passes all
        assert (_abs_f(o, hist[k - 1][2], md, tg)[eq] >= thr_eq[eq]).any() and (fk[eq] < thr_eq[eq]).all()
        x, _, _, _, r = o.solve(x0, md, tg, prm=po.params(reactive_limits=False, tol=0.0,
                                                          row_thresholds=(thr_fixed, thr_eq)))
        assert po.STATUS[r.status] == "CONVERGED" and r.iterations == k, (k, r.iterations)
        np.testing.assert_allclose(x, hist[k][2], rtol=0, atol=1e-10)
        # thresholds shifted by one row: no longer this iterate's
        _, _, _, _, r2 = o.solve(x0, md, tg, prm=po.params(reactive_limits=False, tol=0.0, max_iter=k,
                                                           row_thresholds=(thr_fixed, np.roll(thr_eq, 1))))
        assert not (po.STATUS[r2.status] == "CONVERGED" and r2.iterations == k)
