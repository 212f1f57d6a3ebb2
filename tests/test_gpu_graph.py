"""The solve graphs a context keeps (olfhip_context_graph_statistics): consecutive solves with other inputs (other
scenario counts, with and without per-scenario row-mode overrides, other reactive-limit seeds) replay the context's
instantiated graphs, updated in place where only kernel arguments changed; every graph solve equals, bitwise, the
same solve launched eagerly (OLFHIP_NO_GRAPH=1) between the replays."""
import os

import numpy as np
import pytest

import pyoracle as po
from olfgpu import _native as N
from olfgpu import gridgen
from olfgpu.solver import GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters

pytestmark = pytest.mark.gpu


def test_cached_graph_replays_equal_eager_solves(grid):
    if N.device_count() == 0:
        pytest.fail("no GPU visible: the gpu-marked tests need an MI355X")
    net = grid("grid2000")
    osys = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    xb, mb, tb, ltb, res = osys.solve(x0, md, tg)
    st = N.Structure.from_network(net)
    sa = GpuBatchedSecurityAnalysisSolver(st, net.reactive_limits(),
                                          GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"))
    runs = 0
    try:
        for count, bridges, seed in ((48, 0.25, 1), (48, 0.25, 2), (16, 0.0, 3), (40, 0.1, 4), (40, 0.1, 5)):
            print(f"[graph] grid2000 {count} contingencies, bridges {bridges}, seed {seed}", flush=True)
            ctg = gridgen.sample_contingencies(net, count, bridges, seed)
            os.environ.pop("OLFHIP_NO_GRAPH", None)
            g = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
            runs += 1
            os.environ["OLFHIP_NO_GRAPH"] = "1"
            e = sa.solve_batch_n1(xb, tb, mb, ltb, ctg)
            os.environ.pop("OLFHIP_NO_GRAPH", None)
            assert np.array_equal(g.status, e.status) and np.array_equal(g.total_iterations, e.total_iterations)
            np.testing.assert_array_equal(g.x, e.x)
    finally:
        os.environ.pop("OLFHIP_NO_GRAPH", None)
    inst, upd = sa._ctx.graph_statistics()
    print(f"[graph] {runs} graph solves: {inst} instantiations, {upd} in-place updates", flush=True)
    assert upd >= 4 and inst < 2 * runs
