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


def _brute_neighbourhoods(n_bus, bus1, bus2, weight, k):
    """All-pairs shortest paths (Floyd-Warshall), then per branch the endpoints and the k - 2 nearest other buses by
    (distance to the nearer endpoint, bus id); -1 padded for buses out of reach."""
    D = np.full((n_bus, n_bus), np.inf)
    np.fill_diagonal(D, 0.0)
    for a, b, w in zip(bus1, bus2, weight):
        D[a, b] = min(D[a, b], abs(w))
        D[b, a] = min(D[b, a], abs(w))
    for m in range(n_bus):
        D = np.minimum(D, D[:, m:m + 1] + D[m:m + 1, :])
    out = np.full((len(bus1), k), -1, np.int32)
    for e, (a, b) in enumerate(zip(bus1, bus2)):
        ends = [a] if a == b else [a, b]
        d = np.minimum(D[a], D[b])
        rest = sorted((d[v], v) for v in range(n_bus) if v not in ends and np.isfinite(d[v]))
        row = ends + [v for _, v in rest]
        out[e, :min(k, len(row))] = row[:k]
    return out


def test_neighbourhoods_match_brute_force():
    rng = np.random.default_rng(5)
    for trial in range(6):
        n = int(rng.integers(5, 25))
        m = int(rng.integers(n, 3 * n))
        bus1 = rng.integers(0, n, m)
        bus2 = rng.integers(0, n, m)
        keep = bus1 != bus2
        bus1, bus2 = bus1[keep], bus2[keep]
        # integer weights: exact path sums, so distance ties are real ties broken by bus id in both
        w = rng.integers(1, 6, len(bus1)).astype(float) * np.where(rng.random(len(bus1)) < 0.2, -1, 1)
        for k in (2, 5, n):
            got = mo.neighbourhoods(n, bus1, bus2, w, k)
            assert np.array_equal(got, _brute_neighbourhoods(n, bus1, bus2, w, k)), (trial, k)


def test_local_values_and_warm_start_scatter():
    # 4 buses, neighbourhood of branch 0 = [2, 0, 3, -1]; states are (phi, V) per bus
    hoods = np.array([[2, 0, 3, -1]], np.int32)
    x = np.array([[0.1, 1.01, -0.2, 0.99, 0.3, 1.02, 0.05, 0.97]])
    loc = mo.local_values(x, None, hoods, [0])
    np.testing.assert_allclose(loc[0], [0.0, 0.1 - 0.3, 0.05 - 0.3, 0.0, 0.02, 0.01, -0.03, 0.0], atol=1e-15)
    y = np.arange(8, dtype=float).reshape(1, 8)
    ws = mo.warm_start(x, None, hoods, [0], y, np.zeros(8), np.full(8, 0.5), [2.0])
    d = ws[0] - x[0]
    # slot c adds y[c] * 0.5 * 2 = y[c]: angle rows get slots 0..2, voltage rows slots 4..6; bus 1 untouched
    np.testing.assert_allclose(d, [1.0, 5.0, 0.0, 0.0, 0.0, 4.0, 2.0, 6.0], atol=1e-15)
