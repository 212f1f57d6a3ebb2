"""Pins of the oracle's MAX_VOLTAGE_CHANGE state-vector scaling (OLF's MaxVoltageChangeStateVectorScaling as the
reference applies it per scenario, native/cuda/solver/state_vector_scaling.cu:57-97): every Newton step dx is scaled
as a whole by f = min(1, maxDv / max|dV|, maxDphi / max|dphi|), the maxima over the voltage-magnitude and the
voltage-angle variables, and x += f dx.

* with limits no step reaches, the scaling is plain Newton, bit for bit;
* from a perturbed start with tight limits, an independent Newton loop (dense numpy solves over the oracle's fill,
  the rule above) gives the same iteration count and states within 1e-10; the scaled run takes more iterations than
  the unscaled one, so the limits act;
* a loop that clamps each component instead of scaling the step, or swaps the two limits, gives another iterate
  sequence (checked here, so the independent loop is not blind to those mistakes)."""
import numpy as np
import scipy.sparse as sp

import pyoracle as po
from olfgpu import gridgen


def _newton_with_mvc(o, x0, mode, target, max_dv, max_dphi, max_iter=15, tol=1e-4, variant="scale"):
    n = o.n
    x = x0.copy()
    it = 0
    while True:
        F, _, J = o.fill(x, mode, target)
        if it > 0:
            norm = np.abs(F).max()
            if norm < tol and it < max_iter:
                return x, it, "CONVERGED"
            if it == max_iter:
                return x, it, "MAX_ITERATION_REACHED"
        A = sp.csr_matrix((J, o.colidx, o.rowptr), shape=(n, n)).toarray()
        dx = np.linalg.solve(A, -F)
        dv = np.abs(dx[1::2]).max()
        dphi = np.abs(dx[0::2]).max()
        if variant == "scale":
            f = min(1.0, max_dv / dv if dv > max_dv else 1.0, max_dphi / dphi if dphi > max_dphi else 1.0)
            x = x + f * dx
        elif variant == "swapped":
            f = min(1.0, max_dphi / dv if dv > max_dphi else 1.0, max_dv / dphi if dphi > max_dv else 1.0)
            x = x + f * dx
        else:   # per-component clamp
            lim = np.empty(n)
            lim[0::2] = max_dphi
            lim[1::2] = max_dv
            x = x + np.clip(dx, -lim, lim)
        it += 1


def _perturbed_start(net):
    x0, tg, md, _ = net.system()
    rng = np.random.default_rng(5)
    x0 = x0.copy()
    x0[0::2] += rng.uniform(-0.05, 0.05, net.n_bus)
    x0[1::2] *= rng.uniform(0.95, 1.05, net.n_bus)
    fixed = md == 0
    x0[fixed] = tg[fixed]
    return x0, tg, md


def test_plain_newton_when_no_step_reaches_the_limits():
    net = gridgen.workload("grid118")
    o = po.OracleSystem(net)
    x0, tg, md = _perturbed_start(net)
    xa, _, _, _, ra = o.solve(x0, md, tg, prm=po.params(reactive_limits=False, scaling_mode=0))
    xb, _, _, _, rb = o.solve(x0, md, tg, prm=po.params(reactive_limits=False, scaling_mode=2, max_dv=10.0,
                                                         max_dphi=10.0))
    assert ra.status == rb.status == 0 and ra.iterations == rb.iterations
    np.testing.assert_array_equal(xa, xb)


def test_tight_limits_match_an_independent_scaled_loop():
    net = gridgen.workload("grid118")
    o = po.OracleSystem(net)
    x0, tg, md = _perturbed_start(net)
    dv, dphi = 0.03, 0.1
    xs, its, status = _newton_with_mvc(o, x0, md, tg, dv, dphi)
    xm, _, _, _, rm = o.solve(x0, md, tg, prm=po.params(reactive_limits=False, scaling_mode=2, max_dv=dv,
                                                         max_dphi=dphi))
    xn, _, _, _, rn = o.solve(x0, md, tg, prm=po.params(reactive_limits=False, scaling_mode=0))
    assert status == "CONVERGED" and po.STATUS[rm.status] == status
    assert rm.iterations == its
    np.testing.assert_allclose(xm, xs, rtol=0, atol=1e-10)
    assert rm.iterations > rn.iterations
    # the mistakes this pin is meant to catch change the iterate sequence
    for variant in ("swapped", "clamp"):
        xw, itw, stw = _newton_with_mvc(o, x0, md, tg, dv, dphi, variant=variant)
        assert itw != its or np.abs(xw - xs).max() > 1e-8, variant
