"""Host logic of the solver parameters (CPU): the batch order (longest_first and its tie key) and the
reactive-limit band that the stopping criterion sets."""
import numpy as np

from olfgpu import gridgen
from olfgpu.solver import (NewtonRaphsonParameters, PerEquationTypeStoppingCriteria, longest_first,
                           outage_flow_magnitude)


def test_longest_first_orders_by_cost_then_tie_then_position():
    cost = np.array([2.0, 1.0, 2.0, 2.0, 1.0])
    assert longest_first(cost).tolist() == [0, 2, 3, 1, 4]
    tie = np.array([0.1, 5.0, 0.3, 0.2, 5.0])
    assert longest_first(cost, tie).tolist() == [2, 3, 0, 1, 4]


def test_outage_flow_magnitude_is_the_side1_flow(grid):
    net = grid("grid118")
    x = net.system()[0]
    x[0::2] = np.linspace(0.0, -0.2, net.n_bus)
    s1 = np.abs(gridgen.branch_flows(net, x)[0])
    br = np.array([3, 7, 11])
    np.testing.assert_allclose(outage_flow_magnitude(net, x, br), s1[br])
    multi = np.array([[3, 7, -1, -1], [11, -1, -1, -1]])
    np.testing.assert_allclose(outage_flow_magnitude(net, x, multi), [max(s1[3], s1[7]), s1[11]])


def test_reactive_limit_band_follows_the_stopping_criterion():
    """OLF's createReactiveLimitsOuterLoop rule (GpuSaRunScope.java:339-350): the per-equation-type criterion's
    maxReactivePowerMismatch over the base (not divided twice, GpuSolveSettings.java:35-41), the uniform
    criterion's convEpsPerEq, else the explicit maxReactivePowerMismatch."""
    assert NewtonRaphsonParameters().reactive_limit_band() == 1e-4
    assert NewtonRaphsonParameters(maxReactivePowerMismatch=3e-4).reactive_limit_band() == 3e-4
    assert NewtonRaphsonParameters(convEpsPerEq=2e-5, maxReactivePowerMismatch=3e-4).reactive_limit_band() == 2e-5
    crit = PerEquationTypeStoppingCriteria(maxReactivePowerMismatch=5.0, base_mva=100.0)
    nrp = NewtonRaphsonParameters(convEpsPerEq=2e-5, maxReactivePowerMismatch=3e-4, stoppingCriteria=crit)
    assert nrp.reactive_limit_band() == 0.05
    # the same band as the criterion's own Q-row equation threshold
    assert crit.row_thresholds(4)[1][1::2].tolist() == [0.05, 0.05]
