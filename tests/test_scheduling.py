"""Host logic of the batch order (CPU): longest_first and its tie key."""
import numpy as np

from olfgpu import gridgen
from olfgpu.solver import longest_first, outage_flow_magnitude


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
