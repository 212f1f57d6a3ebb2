"""GPU tests of the learned warm start: the matrix-core GEMM against numpy (lane layout and rounding of
v_mfma_f32_16x16x16bf16_1k), the network's gradients against the CPU oracle, and what the trained network gains on
N-1 solves while every answer stays the base-start answer (same statuses, states within 1e-6)."""
import numpy as np
import pytest

import mlp_oracle as mo
from olfgpu import _native as N
from olfgpu import gridgen
from olfgpu import warmstart as W
from olfgpu.solver import GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters, GpuNewtonRaphson

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(16, 16, 16), (37, 53, 40), (128, 64, 512), (300, 70, 1000), (64, 640, 33)])
def test_matrix_core_gemm_matches_numpy(shape):
    M, Nn, K = shape
    print(f"[gemm] {shape}", flush=True)
    rng = np.random.default_rng(M * 7 + K)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, Nn)).astype(np.float32)
    Cg = W.gemm_test(A, B)
    ref = mo.bf16(A) @ mo.bf16(B)
    err = np.abs(Cg - ref).max() / np.abs(ref).max()
    assert err < 1e-5, err


def test_network_gradients_match_cpu_oracle():
    print("[mlp] gradients", flush=True)
    rng = np.random.default_rng(4)
    n, E, F, H, B = 40, 12, W.N_FEATURES, 32, 24
    net = W.WarmStartMLP(n, E, F, H, seed=5)
    x_mean = rng.normal(0, 0.1, n).astype(np.float32)
    x_std = rng.uniform(0.5, 2, n).astype(np.float32)
    net.set_normalization(x_mean, x_std, np.zeros(n, np.float32), np.ones(n, np.float32))
    p = {k: net.parameters(k).astype(np.float64) for k in mo.PARAMS}
    # non-zero biases so every path is exercised
    for k in ("b1", "b2", "b3"):
        p[k] = rng.normal(0, 0.1, p[k].shape)
        net.set_parameters(k, p[k])
    xb = rng.normal(size=(B, n))
    br = rng.integers(0, E, B).astype(np.int32)
    ft = rng.normal(size=(B, F)).astype(np.float32)
    tg = rng.normal(size=(B, n)).astype(np.float32)
    X = mo.inputs(xb, x_mean.astype(np.float64), x_std.astype(np.float64), ft)
    y_ref = mo.forward(p, X, br)["y"]
    y = net.forward(xb, br, ft)
    assert np.abs(y - y_ref).max() < 1e-3 * np.abs(y_ref).max()
    loss_ref, g_ref = mo.loss_and_grads(p, X, br, tg)
    for k in mo.PARAMS:
        g, loss = net.gradients(xb, br, ft, tg, k)
        assert abs(loss - loss_ref) < 1e-4 * loss_ref
        rel = np.abs(g - g_ref[k]).max() / (np.abs(g_ref[k]).max() + 1e-30)
        assert rel < 2e-2, (k, rel)


def test_trained_warm_start_cuts_newton_iterations_and_keeps_answers(grid):
    """Trained on post-contingency states the solver itself computes (load scalings U(0.9, 1.1)), the network's
    starts need fewer Newton iterations than the base-case start on held-out contingencies, and every answer is the
    base-start answer: same statuses, |V| and angle within 1e-6."""
    net = grid("grid2000")
    print("[warm] grid2000 training set", flush=True)
    st = N.Structure.from_network(net)
    data = W.make_training_set(net, st, n_samples=2400, levels=12, seed=11)
    assert len(data["branch"]) > 2000
    print(f"[warm] training on {len(data['branch'])} samples", flush=True)
    mlp, hist = W.train(data, net.n_branch, hidden=256, epochs=12, batch=128, lr=2e-3, seed=1)
    assert hist[-1] < 0.5 * hist[0], hist
    rl = net.reactive_limits()
    x0, tg, md, _ = net.system()
    base = GpuNewtonRaphson(st, rl).run(x0, tg, md)
    ctg = gridgen.sample_contingencies(net, 256, 0.05, 99)
    feats = W.endpoint_features(net, base.x, ctg, 1.0, gridgen.bridges(net))
    sa = GpuBatchedSecurityAnalysisSolver(st, rl, GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"))
    cold = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg)
    warm = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg, warm_start=mlp,
                             warm_start_features=feats)
    print(f"[warm] iterations cold {cold.total_iterations.sum()} warm {warm.total_iterations.sum()} "
          f"fallbacks {warm.warm_start_fallbacks}", flush=True)
    assert np.array_equal(warm.status, cold.status)
    ok = cold.status == 0
    assert np.abs(warm.x[ok, 1::2] - cold.x[ok, 1::2]).max() <= 1e-6
    assert np.abs(warm.x[ok, 0::2] - cold.x[ok, 0::2]).max() <= 1e-6
    assert warm.total_iterations.sum() < cold.total_iterations.sum()
