"""CPU oracle of the warm-start network (numpy, FP64) — TEST INFRASTRUCTURE ONLY.

The same computations as the device code (olf-gpu_amd/csrc/mlp.hip): the outage neighbourhoods (multi-source
Dijkstra, restated with heapq), the layer-1 input built from a base state's neighbourhood, the 3-layer MLP with GELU
and an embedding row added in the first layer, mean squared error on normalised targets, and the warm start it
produces. The neighbourhoods are pinned by brute force (all-pairs shortest paths on small graphs) in
tests/test_mlp_oracle.py, the gradients by finite differences. With ``emulate_bf16`` the
matrix operands are rounded to BF16 (round to nearest even, from FP32) where the device rounds them, so the device
gradients can be compared tightly; without it the passes are plain FP64, which the finite-difference test pins.
The reference has no learned warm start, so nothing here restates reference code.
"""
from __future__ import annotations

import heapq

import numpy as np
from scipy.special import erf

PARAMS = ("Win", "Emb", "b1", "W2", "b2", "W3", "b3")


def bf16(a):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def _r(a, on):
    return bf16(a) if on else np.asarray(a, dtype=np.float64)


def neighbourhoods(n_bus, bus1, bus2, weight, k):
    """[E][k]: endpoint 1, endpoint 2, then the buses settled by a Dijkstra over |weight| started from both
    endpoints at distance 0, in order of (distance, bus id); -1 padded."""
    adj = [[] for _ in range(n_bus)]
    for a, b, w in zip(bus1, bus2, weight):
        adj[int(a)].append((int(b), abs(float(w))))
        adj[int(b)].append((int(a), abs(float(w))))
    out = np.full((len(bus1), k), -1, np.int32)
    for e, (a, b) in enumerate(zip(bus1, bus2)):
        a, b = int(a), int(b)
        ends = [a] if a == b else [a, b]
        row = list(ends)
        dist = {u: 0.0 for u in ends}
        heap = [(0.0, u) for u in ends]
        heapq.heapify(heap)
        done = set()
        while heap and len(done) < k:
            d, u = heapq.heappop(heap)
            if u in done or d > dist[u]:
                continue
            done.add(u)
            if u not in ends:
                row.append(u)
            for v, w in adj[u]:
                if d + w < dist.get(v, np.inf):
                    dist[v] = d + w
                    heapq.heappush(heap, (d + w, v))
        out[e, :len(row)] = row[:k]
    return out


def local_values(x_rows, base_row, hoods, branch):
    """[B][2k] neighbourhood values of the base states: angle minus endpoint 1's angle, voltage minus 1; 0 on
    padding slots."""
    B = len(branch)
    k = hoods.shape[1]
    out = np.zeros((B, 2 * k))
    for i in range(B):
        x = x_rows[0 if base_row is None else base_row[i]]
        h = hoods[branch[i]]
        for c in range(k):
            if h[c] >= 0:
                out[i, c] = x[2 * h[c]] - x[2 * h[0]]
                out[i, k + c] = x[2 * h[c] + 1] - 1.0
    return out


def inputs(x_rows, base_row, hoods, branch, x_mean, x_std, feat, emulate_bf16=True):
    loc = local_values(x_rows, base_row, hoods, branch)
    xn = ((loc - x_mean) / x_std).astype(np.float32).astype(np.float64)
    X = np.concatenate([xn, np.asarray(feat, np.float64).reshape(len(xn), -1)], axis=1)
    return _r(X, emulate_bf16)


def warm_start(x_rows, base_row, hoods, branch, y, y_mean, y_std, scale):
    """[B][n] starts: the base state plus (y * y_std + y_mean) * scale on the neighbourhood's rows."""
    k = hoods.shape[1]
    out = np.array([x_rows[0 if base_row is None else base_row[i]] for i in range(len(branch))], np.float64)
    for i in range(len(branch)):
        h = hoods[branch[i]]
        s = 1.0 if scale is None else float(scale[i])
        for c in range(2 * k):
            bus = h[c % k]
            if bus >= 0:
                out[i, 2 * bus + (c >= k)] += (float(y[i, c]) * float(y_std[c]) + float(y_mean[c])) * s
    return out


def forward(p, X, branch, emulate_bf16=True):
    e = emulate_bf16
    c1 = X @ _r(p["Win"], e).T
    z1 = c1 + p["b1"] + p["Emb"][branch]
    h1 = _r(gelu(z1), e)
    z2 = h1 @ _r(p["W2"], e).T + p["b2"]
    h2 = _r(gelu(z2), e)
    y = h2 @ _r(p["W3"], e).T + p["b3"]
    return dict(z1=z1, h1=h1, z2=z2, h2=h2, y=y)


def loss_and_grads(p, X, branch, target, emulate_bf16=True):
    e = emulate_bf16
    a = forward(p, X, branch, e)
    B, N = a["y"].shape
    d = a["y"] - target
    loss = float((d * d).mean())
    dy = 2.0 * d / (B * N)
    dyr = _r(dy, e)
    g = {"W3": dyr.T @ a["h2"], "b3": dy.sum(0)}
    dz2 = (dyr @ _r(p["W3"], e)) * gelu_grad(a["z2"])
    dz2r = _r(dz2, e)
    g["W2"] = dz2r.T @ a["h1"]
    g["b2"] = dz2.sum(0)
    dz1 = (dz2r @ _r(p["W2"], e)) * gelu_grad(a["z1"])
    g["Win"] = _r(dz1, e).T @ X
    g["b1"] = dz1.sum(0)
    gE = np.zeros_like(p["Emb"])
    np.add.at(gE, branch, dz1)
    g["Emb"] = gE
    return loss, g
