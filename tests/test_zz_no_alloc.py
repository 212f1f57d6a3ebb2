"""GPU test of the allocation-free screen (olfhip_screen_violations_no_alloc) through the C ABI: on two solves in a
row, the second with fewer scenarios, its counts equal those of the screen that allocates (olfhip_screen_violations)
and each scenario's candidate records are the same set (emission order is not fixed); a size beyond its scratch
array is refused with OLFHIP_E_ARGUMENT."""
import numpy as np
import pytest

from olfgpu import _native as N
from olfgpu import gridgen
from olfgpu.solver import GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters, GpuNewtonRaphson

pytestmark = pytest.mark.gpu


def test_no_alloc_screen_matches_the_screen(grid):
    if N.device_count() == 0:
        pytest.fail("no GPU visible: the gpu-marked tests need an MI355X")
    net = grid("grid2000")
    st = N.Structure.from_network(net)
    rl = net.reactive_limits()
    x0, tg, md, _ = net.system()
    base = GpuNewtonRaphson(st, rl).run(x0, tg, md)
    assert base.status == "CONVERGED"
    lt = N.LimitTable(st, gridgen.limit_table(net, base.x))
    sa = GpuBatchedSecurityAnalysisSolver(st, rl, GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"))
    capacity = 512
    for count, seed in ((48, 6), (16, 7)):
        print(f"[no-alloc] grid2000 {count} contingencies, capacity {capacity}", flush=True)
        ctg = gridgen.sample_contingencies(net, count, 0.25, seed)
        r = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg,
                              states_to_host=False)
        ctx = sa._ctx
        c1, r1 = lt.screen(ctx, count, capacity=capacity)
        c2, r2 = lt.screen_no_alloc(ctx, count, capacity=capacity)
        assert np.array_equal(c1, c2)
        assert (c1 <= capacity).all() and c1.sum() > 0
        assert (c1[r.status != 0] == 0).all()
        for k in range(count):
            assert sorted(r1[k, :c1[k]].tolist()) == sorted(r2[k, :c2[k]].tolist()), k
    # the scratch array's limit, and the context's scenario count
    too_many = N.NO_ALLOC_SCRATCH_BYTES // (4 * count) + 1
    with pytest.raises(N.OlfHipError) as e:
        lt.screen_no_alloc(ctx, count, capacity=too_many)
    assert e.value.code == N.E_ARGUMENT and "scratch" in str(e.value)
    with pytest.raises(N.OlfHipError) as e:
        lt.screen_no_alloc(ctx, count + 1, capacity=capacity)
    assert e.value.code == N.E_ARGUMENT
