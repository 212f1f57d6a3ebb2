"""Pin of the warm-start network's CPU oracle: analytic FP64 gradients of every parameter array against central
finite differences of the loss, and the BF16 rounding helper against exact cases."""
import numpy as np

import mlp_oracle as mo


def small_problem(seed=0, B=6, N=9, F=3, E=5, H=8):
    rng = np.random.default_rng(seed)
    p = {"Win": rng.normal(0, 0.4, (H, N + F)), "Emb": rng.normal(0, 0.3, (E, H)), "b1": rng.normal(0, 0.1, H),
         "W2": rng.normal(0, 0.4, (H, H)), "b2": rng.normal(0, 0.1, H), "W3": rng.normal(0, 0.4, (N, H)),
         "b3": rng.normal(0, 0.1, N)}
    X = rng.normal(size=(B, N + F))
    br = rng.integers(0, E, B)
    br[0] = br[1]            # a repeated branch: its embedding gradient accumulates
    tg = rng.normal(size=(B, N))
    return p, X, br, tg


def test_fp64_gradients_match_finite_differences():
    p, X, br, tg = small_problem()
    _, g = mo.loss_and_grads(p, X, br, tg, emulate_bf16=False)
    h = 1e-6
    rng = np.random.default_rng(1)
    for name in mo.PARAMS:
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in p[name].shape)
            q = {k: v.copy() for k, v in p.items()}
            q[name][idx] += h
            lp, _ = mo.loss_and_grads(q, X, br, tg, emulate_bf16=False)
            q[name][idx] -= 2 * h
            lm, _ = mo.loss_and_grads(q, X, br, tg, emulate_bf16=False)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[name][idx]) < 1e-7 + 1e-5 * abs(fd), (name, idx, fd, g[name][idx])


def test_bf16_rounding():
    assert mo.bf16(1.0) == 1.0
    assert mo.bf16(1.0 + 2.0 ** -8) == 1.0                      # tie to even (down)
    assert mo.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6      # tie to even (up)
    assert mo.bf16(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7          # representable
    x = np.random.default_rng(2).normal(size=1000)
    assert np.all(np.abs(mo.bf16(x) - x) <= np.abs(x) * 2.0 ** -8 + 1e-30)
