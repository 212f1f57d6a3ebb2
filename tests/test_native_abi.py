"""CPU-side checks of the C ABI: the library loads and exports every entry point of include/olfhip.h, the host plan
(pattern, order, islands) agrees with the oracle, contract violations come back as error codes, and the parameter
surface keeps the reference's names and defaults. No compute call needs a GPU here."""
import os
import re

import numpy as np
import pytest

import pyoracle as po
from olfgpu import _native as N
from olfgpu import gridgen
from olfgpu.solver import GpuLoadFlowParameters, NewtonRaphsonParameters

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(REPO, "include", "olfhip.h")).read()
    declared = sorted(set(re.findall(r"\b(olfhip_[a-z_0-9]+)\s*\(", header)))
    assert declared, "no declarations parsed"
    L = N.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(N.EXPORTS) == declared


def test_device_count_never_fails():
    assert N.device_count() >= 0


def test_pattern_matches_the_oracle_pattern(grid):
    net = grid("grid2000")
    st = N.Structure.from_network(net)
    rp, ci = st.pattern()
    osys = po.OracleSystem(net)
    assert np.array_equal(rp, osys.rowptr)
    assert np.array_equal(ci, osys.colidx)


def test_plan_statistics_and_order(grid):
    net = grid("grid9241")
    st = N.Structure.from_network(net)
    s = st.statistics()
    assert s["buses"] == 9241
    assert s["pattern_nnz"] == 4 * 9241 + 8 * 16000 - 8 * 0  # no parallel branches in the generator
    assert s["lu_nnz"] >= s["pattern_nnz"]
    assert 10 < s["levels"] < 1000
    order = st.order()
    assert sorted(order.tolist()) == list(range(9241))


def test_islands_match_the_oracle(grid):
    net = grid("grid2000")
    st = N.Structure.from_network(net)
    osys = po.OracleSystem(net)
    ctg = gridgen.sample_contingencies(net, 40, 0.5, 4)
    for e in ctg:
        assert np.array_equal(st.islanded_buses([e]), osys.islands([e]))
    two = [int(ctg[0]), int(ctg[1])]
    assert np.array_equal(st.islanded_buses(two), osys.islands(two))


def test_unsupported_and_invalid_inputs_return_codes(grid):
    net = grid("grid118")
    idx = net.branch_indices().copy()
    idx[0, 9] = 7  # a ratio-control column: not modelled
    with pytest.raises(N.OlfHipError) as e:
        N.Structure(net.n, net.pi_model(), idx)
    assert e.value.code == N.E_UNSUPPORTED
    idx = net.branch_indices().copy()
    idx[3, 1] = idx[3, 0]  # Q row not the bus's V row
    with pytest.raises(N.OlfHipError) as e:
        N.Structure(net.n, net.pi_model(), idx)
    assert e.value.code == N.E_UNSUPPORTED
    with pytest.raises(N.OlfHipError) as e:
        N.Structure(net.n + 1, net.pi_model(), net.branch_indices())
    assert e.value.code == N.E_ARGUMENT
    assert "even" in N.lib().olfhip_last_error().decode()


def test_robust_mode_contract(grid):
    """The robust mode solves one scenario per tile and is a Newton-solve mode only; both are refused before any
    device work."""
    net = grid("grid118")
    st = N.Structure.from_network(net)
    with pytest.raises(N.OlfHipError) as e:
        N.BatchContext(st, 4, N.LU_ROBUST, net.reactive_limits(), lanes_per_tile=2)
    assert e.value.code == N.E_ARGUMENT and "one scenario per tile" in str(e.value)
    with pytest.raises(N.OlfHipError) as e:
        st.lu_factorize_and_solve(np.zeros((1, st.statistics()["pattern_nnz"])), np.zeros((1, net.n)), mode=N.LU_ROBUST)
    assert e.value.code == N.E_ARGUMENT and "Newton-solve mode" in str(e.value)
    with pytest.raises(N.OlfHipError) as e:
        N.BatchContext(st, 4, 7, net.reactive_limits())
    assert e.value.code == N.E_ARGUMENT


def test_parameters_keep_the_reference_surface():
    p = GpuLoadFlowParameters()
    assert p.NAME == "gpu-load-flow-parameters"
    assert p.MODULE == "open-loadflow-gpu-default-parameters"
    assert (p.linearSolver, p.reorderingAlgorithm, p.batchedLuOrdering, p.batchedLuSplitLongRows,
            p.matchingAlgorithm, p.pivotType) == ("BATCHED_LU", "AUTO", "CUDSS", True, "AUTO", "BY_ORDERING")
    assert np.isnan(p.pivotThreshold)
    assert (p.iterativeRefinementBaseSteps, p.iterativeRefinementMaxSteps, p.iterativeRefinementTolerance) == (1, 4, 1e-12)
    q = GpuLoadFlowParameters.load({"linearSolver": "BATCHED_LU_FP32_REFINED", "batchedLuSplitLongRows": "false"})
    assert q.lu_mode == N.LU_FP32_REFINED and q.batchedLuSplitLongRows is False
    with pytest.raises(ValueError):
        GpuLoadFlowParameters(linearSolver="CUDSS")
    with pytest.raises(ValueError):
        GpuLoadFlowParameters.load({"noSuchParameter": 1})
    assert NewtonRaphsonParameters().scaling_code() == N.SCALING_NONE
    assert NewtonRaphsonParameters(stateVectorScalingMode="LINE_SEARCH").scaling_code() == N.SCALING_LINE_SEARCH
    with pytest.raises(ValueError):
        NewtonRaphsonParameters(stateVectorScalingMode="NOT_A_MODE").scaling_code()


def test_outage_neighbourhoods_match_the_oracle(grid):
    """The warm-start network's outage neighbourhoods (a host routine of the library) against the oracle's Dijkstra,
    itself pinned by brute force."""
    import mlp_oracle as mo
    from olfgpu import warmstart as W

    net = grid("grid118")
    for k in (2, 17, 64):
        got = W.neighbourhoods(net.n_bus, net.bus1, net.bus2, net.x, k)
        assert np.array_equal(got, mo.neighbourhoods(net.n_bus, net.bus1, net.bus2, net.x, k)), k
    big = grid("grid2000")
    sel = np.arange(0, big.n_branch, 37)
    got = W.neighbourhoods(big.n_bus, big.bus1, big.bus2, big.x, 96)[sel]
    ref = mo.neighbourhoods(big.n_bus, big.bus1, big.bus2, big.x, 96)[sel]
    assert np.array_equal(got, ref)
    with pytest.raises(N.OlfHipError) as e:
        W.neighbourhoods(net.n_bus, net.bus1, net.bus2, net.x, net.n_bus + 1)
    assert e.value.code == N.E_ARGUMENT


def test_limit_table_contract_violations_return_codes(grid):
    """The limit table's arguments are checked before any device use."""
    net = grid("grid118")
    st = N.Structure.from_network(net)
    x0 = net.system()[0]
    good = gridgen.limit_table(net, x0)
    for key, bad in (("group_side", 3), ("group_quantity", 7), ("branch", net.n_branch)):
        t = {k: np.array(v, copy=True) for k, v in good.items()}
        t[key][0] = bad
        with pytest.raises(N.OlfHipError) as e:
            N.LimitTable(st, t)
        assert e.value.code == N.E_ARGUMENT, key
    t = {k: np.array(v, copy=True) for k, v in good.items()}
    t["bus_high"] = np.zeros(net.n_bus + 1)
    t["bus_low"] = np.zeros(net.n_bus + 1)
    with pytest.raises(N.OlfHipError) as e:
        N.LimitTable(st, t)
    assert e.value.code == N.E_ARGUMENT


def test_voltage_angle_limit_contract(grid):
    """Voltage-angle limits: a bus out of range fails with a code before any device use; arrays of unequal lengths
    are refused by the Python owner."""
    net = grid("grid118")
    st = N.Structure.from_network(net)
    good = gridgen.limit_table(net, net.system()[0], angle_limits=4)
    for key, bad in (("angle_from", net.n_bus), ("angle_to", -1)):
        t = {k: np.array(v, copy=True) for k, v in good.items()}
        t[key][2] = bad
        with pytest.raises(N.OlfHipError) as e:
            N.LimitTable(st, t)
        assert e.value.code == N.E_ARGUMENT, key
    t = dict(good, angle_low=good["angle_low"][:3])
    with pytest.raises(ValueError):
        N.LimitTable(st, t)
