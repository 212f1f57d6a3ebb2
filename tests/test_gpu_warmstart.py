"""GPU tests of the learned warm start: the matrix-core GEMM against numpy (lane layout and rounding of
v_mfma_f32_16x16x16bf16_1k), the network's gradients against the CPU oracle, and what the trained network gains on
N-1 solves while the Newton mismatch stays the sole criterion of every answer (same statuses as the base-case start)."""
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


def _device_to_host(ptr, shape):
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    out = np.zeros(shape, np.float64)
    rc = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(out.nbytes),
                       ctypes.c_int(2))  # hipMemcpyDeviceToHost
    assert rc == 0, rc
    return out


def test_network_passes_match_cpu_oracle():
    """Layer-1 input from the neighbourhoods, forward, every parameter gradient and the warm-start scatter against
    the CPU oracle (which rounds the matrix operands to BF16 where the device does)."""
    print("[mlp] forward, gradients and warm start vs oracle", flush=True)
    rng = np.random.default_rng(4)
    net = gridgen.generate(40, 70, seed=9)
    K, F, H, B, R = 6, W.N_FEATURES, 32, 24, 3
    mlp = W.WarmStartMLP(net, n_local=K, hidden=H, seed=5)
    hoods = mo.neighbourhoods(net.n_bus, net.bus1, net.bus2, net.x, K)
    assert np.array_equal(mlp.hoods, hoods)
    n = 2 * K
    x_mean = rng.normal(0, 0.1, n).astype(np.float32)
    x_std = rng.uniform(0.5, 2, n).astype(np.float32)
    y_mean = rng.normal(0, 0.1, n).astype(np.float32)
    y_std = rng.uniform(0.5, 2, n).astype(np.float32)
    mlp.set_normalization(x_mean, x_std, y_mean, y_std)
    x_rows = np.zeros((R, net.n))
    x_rows[:, 0::2] = rng.normal(0, 0.2, (R, net.n_bus))
    x_rows[:, 1::2] = 1 + rng.normal(0, 0.03, (R, net.n_bus))
    mlp.set_base_states(x_rows)
    p = {k: mlp.parameters(k).astype(np.float64) for k in mo.PARAMS}
    # non-zero biases so every path is exercised
    for k in ("b1", "b2", "b3"):
        p[k] = rng.normal(0, 0.1, p[k].shape)
        mlp.set_parameters(k, p[k])
    rows = rng.integers(0, R, B).astype(np.int32)
    br = rng.integers(0, net.n_branch, B).astype(np.int32)
    br[1] = br[0]                                   # a repeated branch: its embedding gradient accumulates
    ft = rng.normal(size=(B, F)).astype(np.float32)
    tg = rng.normal(size=(B, n)).astype(np.float32)
    X = mo.inputs(x_rows, rows, hoods, br, x_mean.astype(np.float64), x_std.astype(np.float64), ft)
    y_ref = mo.forward(p, X, br)["y"]
    y = mlp.forward(rows, br, ft)
    assert np.abs(y - y_ref).max() < 1e-3 * np.abs(y_ref).max()
    loss_ref, g_ref = mo.loss_and_grads(p, X, br, tg)
    for k in mo.PARAMS:
        g, loss = mlp.gradients(rows, br, ft, tg, k)
        assert abs(loss - loss_ref) < 1e-4 * loss_ref
        rel = np.abs(g - g_ref[k]).max() / (np.abs(g_ref[k]).max() + 1e-30)
        assert rel < 2e-2, (k, rel)
    scale = rng.uniform(0.5, 2, B).astype(np.float32)
    ptr = mlp.predict_raw(br, ft, scale, base_row=rows)
    x0 = _device_to_host(ptr, (B, net.n))
    ref = mo.warm_start(x_rows, rows, hoods, br, mlp.forward(rows, br, ft), y_mean, y_std, scale)
    assert np.abs(x0 - ref).max() < 1e-6


def test_trained_warm_start_cuts_newton_iterations_and_keeps_answers(grid):
    """Trained on post-contingency states the solver itself computes (every branch covered, load scalings
    U(0.9, 1.1)), the network's starts need fewer Newton iterations than the base-case start on a contingency set at
    the base load (a load level not trained on). Every answer is judged by the Newton mismatch: statuses are those of
    the base-case start, and a warm-started state differs from the base-start one only within the stopping
    tolerance's reach: both satisfy ||F|| < 1e-4, so they may differ by up to ||J^-1|| times that (the test bounds the
    difference by 1e-3 and prints it)."""
    net = grid("grid2000")
    print("[warm] grid2000 training set", flush=True)
    st = N.Structure.from_network(net)
    mlp = W.WarmStartMLP(net, hidden=512, seed=1)
    data = W.make_training_set(net, st, mlp.hoods, n_samples=6000, levels=12, seed=11)
    assert len(data["branch"]) > 5000
    print(f"[warm] training on {len(data['branch'])} samples", flush=True)
    hist = W.train(mlp, data, epochs=100, batch=128, lr=2e-3, seed=1)
    print(f"[warm] loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    assert hist[-1] < 0.1 * hist[0], hist
    rl = net.reactive_limits()
    x0, tg, md, _ = net.system()
    base = GpuNewtonRaphson(st, rl).run(x0, tg, md)
    ctg = gridgen.sample_contingencies(net, 400, 0.05, 99)
    feats = W.endpoint_features(net, base.x, ctg, 1.0, gridgen.bridges(net))
    sa = GpuBatchedSecurityAnalysisSolver(st, rl, GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"))
    cold = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg)
    warm = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg, warm_start=mlp,
                             warm_start_features=feats)
    ok = cold.status == 0
    dx = np.abs(warm.x[ok] - cold.x[ok]).max()
    print(f"[warm] iterations cold {cold.total_iterations.sum()} {np.bincount(cold.total_iterations)} warm "
          f"{warm.total_iterations.sum()} {np.bincount(warm.total_iterations)} fallbacks "
          f"{warm.warm_start_fallbacks}; max |x_warm - x_cold| {dx:.2e}", flush=True)
    assert np.array_equal(warm.status, cold.status)
    assert dx <= 1e-3
    assert warm.total_iterations.sum() <= 0.9 * cold.total_iterations.sum()
