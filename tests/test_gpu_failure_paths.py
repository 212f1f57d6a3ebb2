"""GPU tests of the paths that end a scenario without convergence, and of the batch mechanics around them, through
the C ABI against the oracle on the same inputs:

* the linear-solve guard (reference native/cuda/solver/newton_iteration.cu:612-652): with its bound forced from the
  host below any residual (linear_solve_fail_ratio = 1e-300), every scenario that needs a second Newton step of its
  first solve ends LINEAR_SOLVE_FAILED after one iteration, the others CONVERGED, as in the oracle;
* a bad last solve at the budget (:369-389): with max_iter = 1 and the same bound every scenario ends REJECTED in a
  batch (MAX_ITERATION_REACHED under the default bound), as in the oracle; solve_batch_n1 hands exactly those
  scenarios to the robust pass, whose answers (rejected again under the same bound) are not kept;
* the fixed-iteration mode (tol <= 0): every scenario runs exactly max_iter steps, states equal to the oracle's;
* a batch solved in several chunks gives each contingency bitwise the answer of one chunk;
* the autotuner returns a valid tile shape whose answers agree with the default's."""
import numpy as np
import pytest

import pyoracle as po
from olfgpu import _native as N
from olfgpu import gridgen
from olfgpu.solver import GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters, NewtonRaphsonParameters

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g2000(grid):
    if N.device_count() == 0:
        pytest.fail("no GPU visible: the gpu-marked tests need an MI355X")
    net = grid("grid2000")
    osys = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    xb, mb, tb, ltb, res = osys.solve(x0, md, tg)
    assert po.STATUS[res.status] == "CONVERGED"
    return net, osys, N.Structure.from_network(net), (xb, mb, tb, ltb)


def solver(g, lu="BATCHED_LU", robust_retry=False, **nr):
    net, osys, st, base = g
    return GpuBatchedSecurityAnalysisSolver(st, net.reactive_limits(), GpuLoadFlowParameters(linearSolver=lu),
                                            NewtonRaphsonParameters(**nr), robust_retry=robust_retry)


def run(g, sa, ctg):
    xb, mb, tb, ltb = g[3]
    return sa.solve_batch_n1(xb, tb, mb, ltb, ctg)


@pytest.mark.parametrize("lu", ["BATCHED_LU", "BATCHED_LU_FP32_REFINED"])
def test_linear_solve_guard_matches_oracle(g2000, lu):
    net, osys, st, (xb, mb, tb, ltb) = g2000
    print(f"[guard] grid2000 forced guard bound {lu}", flush=True)
    ctg = gridgen.sample_contingencies(net, 64, 0.1, 3)
    ref = osys.n1(xb, mb, tb, ltb, ctg, prm=po.params(batch_of_one=False, lin_fail_ratio=1e-300))
    r = run(g2000, solver(g2000, lu, linearSolveFailRatio=1e-300), ctg)
    assert np.array_equal(r.status, ref["status"])
    assert np.array_equal(r.iterations, ref["iterations"])
    failed = r.status == N.LINEAR_SOLVE_FAILED
    assert failed.sum() >= 40 and (r.status == N.CONVERGED).sum() >= 3
    assert (r.iterations[failed] == 1).all()


def test_bad_solve_at_the_budget_is_rejected_and_retried(g2000):
    net, osys, st, (xb, mb, tb, ltb) = g2000
    print("[budget] grid2000 max_iter 1, forced guard bound", flush=True)
    ctg = gridgen.sample_contingencies(net, 48, 0.1, 4)
    ref = osys.n1(xb, mb, tb, ltb, ctg, prm=po.params(batch_of_one=False, max_iter=1, lin_fail_ratio=1e-300))
    assert (ref["status"] == N.REJECTED).all()
    plain = osys.n1(xb, mb, tb, ltb, ctg, prm=po.params(batch_of_one=False, max_iter=1))
    assert (plain["status"] == N.MAX_ITERATION_REACHED).all()
    r = run(g2000, solver(g2000, maxIterations=1, linearSolveFailRatio=1e-300), ctg)
    assert np.array_equal(r.status, ref["status"]) and (r.iterations == 1).all()
    r = run(g2000, solver(g2000, maxIterations=1), ctg)
    assert np.array_equal(r.status, plain["status"])
    # the robust pass takes exactly the rejected scenarios; rejected again, they keep the first pass's answers
    r = run(g2000, solver(g2000, robust_retry=True, maxIterations=1, linearSolveFailRatio=1e-300), ctg)
    assert r.robust_attempted == len(ctg) and r.robust_recovered == 0 and r.robust_errors == []
    assert np.array_equal(r.status, ref["status"])


@pytest.mark.parametrize("lu", ["BATCHED_LU", "BATCHED_LU_FP32_REFINED"])
def test_fixed_iteration_mode_matches_oracle(g2000, lu):
    net, osys, st, (xb, mb, tb, ltb) = g2000
    print(f"[fixed] grid2000 tol 0, 3 steps {lu}", flush=True)
    ctg = gridgen.sample_contingencies(net, 40, 0.1, 5)
    ref = osys.n1(xb, mb, tb, ltb, ctg, prm=po.params(batch_of_one=False, tol=0.0, max_iter=3))
    assert (ref["status"] == N.MAX_ITERATION_REACHED).all() and (ref["iterations"] == 3).all()
    r = run(g2000, solver(g2000, lu, maxMismatch=0.0, maxIterations=3), ctg)
    assert np.array_equal(r.status, ref["status"]) and (r.iterations == 3).all()
    assert np.abs(r.x - ref["x"]).max() <= 1e-6
    np.testing.assert_allclose(r.norm, ref["norm"], rtol=1e-3, atol=1e-9)


def test_chunks_give_one_chunks_answers(g2000):
    net, osys, st, base = g2000
    print("[chunks] grid2000 48 contingencies in chunks of 16", flush=True)
    ctg = gridgen.sample_contingencies(net, 48, 0.2, 6)
    one = run(g2000, solver(g2000, "BATCHED_LU_FP32_REFINED"), ctg)
    sa = solver(g2000, "BATCHED_LU_FP32_REFINED")
    sa.max_chunk = 16
    parts = run(g2000, sa, ctg)
    assert one.chunks == 1 and parts.chunks == 3
    assert np.array_equal(one.status, parts.status) and np.array_equal(one.total_iterations, parts.total_iterations)
    np.testing.assert_array_equal(one.x, parts.x)


def test_autotune_picks_a_valid_shape(g2000):
    net, osys, st, (xb, mb, tb, ltb) = g2000
    print("[autotune] grid2000 32 contingencies", flush=True)
    ctg = gridgen.sample_contingencies(net, 32, 0.0, 7)
    inputs, keep = N.BatchContext.make_inputs(len(ctg), xb, tb, mb, outaged_branches=ctg,
                                              reactive_limits_seed_state=ltb)
    best = N.autotune(st, inputs, N.LU_FP64, net.reactive_limits())
    assert best["lanes_per_tile"] in (1, 2, 4, 8, 16, 32) and best["threads_per_tile"] in (256, 512, 1024)
    assert best["ms"] > 0
    ref = run(g2000, solver(g2000), ctg)
    sa = solver(g2000)
    sa.lanes_per_tile, sa.threads_per_tile = best["lanes_per_tile"], best["threads_per_tile"]
    r = run(g2000, sa, ctg)
    assert np.array_equal(r.status, ref.status)
    assert np.abs(r.x - ref.x).max() < 1e-9


def test_sparse_start_seeds_equal_full_starts(g2000):
    """olfhip_set_x0_seeds (NewtonSettingsNative.setBatchX0Seeds): start values of chosen (scenario, row) pairs over
    the tiled base start give bitwise the solve of the same starts passed in full (x0_per_scenario), and the oracle's
    solve from those starts to 1e-6; clearing the seeds gives back the base-start solve."""
    net, osys, st, (xb, mb, tb, ltb) = g2000
    print("[seeds] grid2000 24 contingencies, sparse start seeds", flush=True)
    ctg = gridgen.sample_contingencies(net, 24, 0.0, 8)
    n = net.n
    rng = np.random.default_rng(8)
    idx, val = [], []
    x0p = np.tile(xb, (len(ctg), 1))
    for s_, e in enumerate(ctg):
        for b in (net.bus1[e], net.bus2[e]):
            for r, d in ((2 * b, rng.uniform(-0.01, 0.01)), (2 * b + 1, rng.uniform(-0.005, 0.005))):
                idx.append(s_ * n + r)
                val.append(xb[r] + d)
                x0p[s_, r] = xb[r] + d
    ctx = N.BatchContext(st, len(ctg), N.LU_FP64, net.reactive_limits())
    inp, keep = N.BatchContext.make_inputs(len(ctg), xb, tb, mb, outaged_branches=ctg, reactive_limits_seed_state=ltb)
    base = ctx.solve(inp)
    ctx.set_x0_seeds(np.array(idx), np.array(val))
    seeded = ctx.solve(inp)
    ctx.set_x0_seeds(None)
    again = ctx.solve(inp)
    full, keep2 = N.BatchContext.make_inputs(len(ctg), xb, tb, mb, outaged_branches=ctg, reactive_limits_seed_state=ltb,
                                             x0_per_scenario=x0p)
    ref = ctx.solve(full)
    np.testing.assert_array_equal(seeded["x"], ref["x"])
    assert np.array_equal(seeded["total_iterations"], ref["total_iterations"])
    np.testing.assert_array_equal(again["x"], base["x"])
    assert np.abs(seeded["x"] - base["x"]).max() > 0
    orc = osys.n1(xb, mb, tb, ltb, ctg, x0=x0p)
    assert np.array_equal(orc["status"], seeded["status"])
    assert np.abs(orc["x"] - seeded["x"]).max() <= 1e-6
    # a seed of a scenario beyond the solve's count is refused
    with pytest.raises(N.OlfHipError):
        ctx.set_x0_seeds(np.array([len(ctg) * n]), np.array([1.0]))   # beyond the context's capacity
    ctx.set_x0_seeds(np.array([(len(ctg) - 1) * n]), np.array([0.0]))
    small, keep3 = N.BatchContext.make_inputs(len(ctg) - 1, xb, tb, mb, outaged_branches=ctg[:-1],
                                              reactive_limits_seed_state=ltb)
    with pytest.raises(N.OlfHipError):
        ctx.solve(small)
