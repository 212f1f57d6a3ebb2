"""Learned warm start for the batched N-1: the network (over the C ABI), its features, its training set made by the
Newton solver itself, and its training loop.

The network maps a scenario's base-case state (all bus V/phi), its outaged branch (embedding) and features of the
branch's endpoints to the post-contingency state change. Training samples are post-contingency states the batched
solver computes under random load scalings; a warm-started scenario that does not converge is solved again from the
base-case start, so the converged Newton mismatch stays the sole judge of every answer.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _native as N
from .gridgen import SB

PARAMS = ("Win", "Emb", "b1", "W2", "b2", "W3", "b3")
N_FEATURES = 16


def _lib():
    L = N.lib()
    if not getattr(L, "_mlp_bound", False):
        vp, i, d, f = C.c_void_p, C.c_int, C.c_double, C.c_float
        L.olfhip_mlp_create.argtypes = [i, i, i, i, C.c_ulonglong, C.POINTER(vp)]
        L.olfhip_mlp_destroy.argtypes = [vp]
        L.olfhip_mlp_destroy.restype = None
        L.olfhip_mlp_set_normalization.argtypes = [vp, vp, vp, vp, vp]
        L.olfhip_mlp_train_step.argtypes = [vp, i, vp, vp, vp, vp, f, C.POINTER(d)]
        L.olfhip_mlp_forward.argtypes = [vp, i, vp, vp, vp, vp]
        L.olfhip_mlp_gradients.argtypes = [vp, i, vp, vp, vp, vp, i, vp, C.POINTER(d)]
        L.olfhip_mlp_parameters.argtypes = [vp, i, vp, vp]
        L.olfhip_mlp_predict_warm_start.argtypes = [vp, i, vp, i, vp, vp, C.POINTER(C.c_void_p)]
        L.olfhip_mlp_gemm_test.argtypes = [i, i, i, vp, vp, vp]
        L._mlp_bound = True
    return L


def gemm_test(A, B):
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    Cm = np.zeros((A.shape[0], B.shape[1]), np.float32)
    N.check(_lib().olfhip_mlp_gemm_test(A.shape[0], B.shape[1], A.shape[1], N.ptr(A), N.ptr(B), N.ptr(Cm)))
    return Cm


class WarmStartMLP:
    """Owner of an olfhip_mlp."""

    def __init__(self, n_state, n_branch, n_features=N_FEATURES, hidden=512, seed=0):
        self.n_state, self.n_branch, self.n_features, self.hidden = n_state, n_branch, n_features, hidden
        h = C.c_void_p()
        N.check(_lib().olfhip_mlp_create(n_state, n_branch, n_features, hidden, seed, C.byref(h)))
        self.h = h
        self.shapes = {"Win": (hidden, n_state + n_features), "Emb": (n_branch, hidden), "b1": (hidden,),
                       "W2": (hidden, hidden), "b2": (hidden,), "W3": (n_state, hidden), "b3": (n_state,)}

    def close(self):
        if getattr(self, "h", None):
            _lib().olfhip_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_normalization(self, x_mean, x_std, y_mean, y_std):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (x_mean, x_std, y_mean, y_std)]
        N.check(_lib().olfhip_mlp_set_normalization(self.h, *[N.ptr(a) for a in arrs]))
        self.norm = arrs

    def _batch(self, x_base, branch, feat):
        x = np.ascontiguousarray(x_base, np.float64)
        br = np.ascontiguousarray(branch, np.int32)
        ft = np.ascontiguousarray(feat, np.float32).reshape(len(br), self.n_features)
        return x, br, ft

    def train_step(self, x_base, branch, feat, target, lr):
        x, br, ft = self._batch(x_base, branch, feat)
        tg = np.ascontiguousarray(target, np.float32)
        loss = C.c_double()
        N.check(_lib().olfhip_mlp_train_step(self.h, len(br), N.ptr(x), N.ptr(br), N.ptr(ft), N.ptr(tg), lr,
                                             C.byref(loss)))
        return loss.value

    def forward(self, x_base, branch, feat):
        x, br, ft = self._batch(x_base, branch, feat)
        y = np.zeros((len(br), self.n_state), np.float32)
        N.check(_lib().olfhip_mlp_forward(self.h, len(br), N.ptr(x), N.ptr(br), N.ptr(ft), N.ptr(y)))
        return y

    def gradients(self, x_base, branch, feat, target, which):
        x, br, ft = self._batch(x_base, branch, feat)
        tg = np.ascontiguousarray(target, np.float32)
        g = np.zeros(self.shapes[which], np.float32)
        loss = C.c_double()
        N.check(_lib().olfhip_mlp_gradients(self.h, len(br), N.ptr(x), N.ptr(br), N.ptr(ft), N.ptr(tg),
                                            PARAMS.index(which), N.ptr(g), C.byref(loss)))
        return g, loss.value

    def parameters(self, which):
        v = np.zeros(self.shapes[which], np.float32)
        N.check(_lib().olfhip_mlp_parameters(self.h, PARAMS.index(which), N.ptr(v), None))
        return v

    def set_parameters(self, which, values):
        v = np.ascontiguousarray(values, np.float32).reshape(self.shapes[which])
        N.check(_lib().olfhip_mlp_parameters(self.h, PARAMS.index(which), None, N.ptr(v)))

    def predict_warm_start(self, x_base, branch, feat):
        """Device pointer (int) to [K][n_state] start states for olfhip_solve_inputs.x0_per_scenario_device."""
        x = np.ascontiguousarray(x_base, np.float64)
        per = 1 if x.ndim == 2 else 0
        br = np.ascontiguousarray(branch, np.int32)
        ft = np.ascontiguousarray(feat, np.float32).reshape(len(br), self.n_features)
        out = C.c_void_p()
        N.check(_lib().olfhip_mlp_predict_warm_start(self.h, len(br), N.ptr(x), per, N.ptr(br), N.ptr(ft),
                                                     C.byref(out)))
        return out.value


def endpoint_features(net, x_base, branches, load_scale, bridge_mask=None):
    """16 features of each outaged branch at the base state: endpoint V and phi, their difference, the base-case
    flows p1 q1 p2 q2 (textbook pi-branch), the branch's r, x, b and ratio, whether it is a bridge, the load scale
    and a constant."""
    br = np.asarray(branches, dtype=np.int64)
    a, b = net.bus1[br], net.bus2[br]
    x = np.asarray(x_base, np.float64)
    th, v = x[0::2], x[1::2]
    Va = v[a] * np.exp(1j * th[a]) * net.ratio[br]
    Vb = v[b] * np.exp(1j * th[b])
    ys = 1.0 / (net.r[br] + 1j * net.x[br])
    bc = 1j * net.b[br] / 2
    S1 = Va * np.conj((Va - Vb) * ys + Va * bc)
    S2 = Vb * np.conj((Vb - Va) * ys + Vb * bc)
    f = np.zeros((len(br), N_FEATURES), np.float32)
    f[:, 0] = v[a] - 1.0
    f[:, 1] = v[b] - 1.0
    f[:, 2] = th[a]
    f[:, 3] = th[b]
    f[:, 4] = th[a] - th[b]
    f[:, 5] = S1.real
    f[:, 6] = S1.imag
    f[:, 7] = S2.real
    f[:, 8] = S2.imag
    f[:, 9] = net.r[br] * 10
    f[:, 10] = net.x[br] * 10
    f[:, 11] = net.b[br] * 10
    f[:, 12] = net.ratio[br] - 1.0
    f[:, 13] = 0.0 if bridge_mask is None else bridge_mask[br].astype(np.float32)
    f[:, 14] = load_scale - 1.0
    f[:, 15] = 1.0
    return f


def make_training_set(net, structure, n_samples=20000, levels=40, seed=3, scale_range=(0.9, 1.1),
                      bridge_fraction=0.05, verbose=False):
    """Post-contingency states solved by the batched Newton solver itself, under uniform random load scalings.

    Returns a dict with x_base [levels, n], level index, branch, features and the state change target [S, n]
    (islanded rows set to zero change: they are frozen whatever the start)."""
    from . import gridgen
    from .solver import GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters, GpuNewtonRaphson

    rng = np.random.default_rng(seed)
    scales = rng.uniform(scale_range[0], scale_range[1], levels)
    per = int(np.ceil(n_samples / levels))
    bridges = gridgen.bridges(net)
    xb_all, lvl, brs, feats, targets = [], [], [], [], []
    t0 = time.perf_counter()
    for k, s in enumerate(scales):
        ns = net.scaled(s)
        rl = ns.reactive_limits()
        x0, tg, md, _ = ns.system()
        base = GpuNewtonRaphson(structure, rl).run(x0, tg, md)
        if base.status != "CONVERGED":
            continue
        ctg = gridgen.sample_contingencies(net, per, bridge_fraction, int(rng.integers(1 << 31)))
        sa = GpuBatchedSecurityAnalysisSolver(structure, rl, GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"))
        res = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg)
        ok = res.status == 0
        d = res.x[ok] - base.x[None, :]
        for j, e in enumerate(ctg[ok]):
            if bridges[e]:
                isl = structure.islanded_buses([e])
                d[j, 0::2][isl] = 0.0
                d[j, 1::2][isl] = 0.0
        xb_all.append(base.x)
        li = len(xb_all) - 1
        lvl.append(np.full(ok.sum(), li, np.int32))
        brs.append(ctg[ok])
        feats.append(endpoint_features(net, base.x, ctg[ok], s, bridges))
        targets.append(d.astype(np.float32))
        if verbose:
            print(f"  level {k}: scale {s:.3f}, {ok.sum()}/{len(ctg)} converged, {time.perf_counter() - t0:.1f}s",
                  flush=True)
    return dict(x_base=np.array(xb_all), level=np.concatenate(lvl), branch=np.concatenate(brs).astype(np.int32),
                features=np.concatenate(feats), target=np.concatenate(targets), scales=scales,
                seconds=time.perf_counter() - t0)


def train(data, n_branch, hidden=512, epochs=20, batch=256, lr=1e-3, seed=0, verbose=False):
    """Fit a WarmStartMLP to a training set; returns (network, loss history)."""
    xb, lvl, br, ft, tg = data["x_base"], data["level"], data["branch"], data["features"], data["target"]
    n = tg.shape[1]
    x_mean = xb.mean(0)
    x_std = np.maximum(xb.std(0), 1e-3)
    # one scale per variable type (angle rows, voltage rows), no offset: a row's change is large only near the
    # outage, and those large changes are what decide whether a start converges in one Newton step — a per-row scale
    # would blow up rows that hardly move
    y_mean = np.zeros(n)
    y_std = np.empty(n)
    y_std[0::2] = max(float(np.sqrt((tg[:, 0::2].astype(np.float64) ** 2).mean())), 1e-6)
    y_std[1::2] = max(float(np.sqrt((tg[:, 1::2].astype(np.float64) ** 2).mean())), 1e-6)
    net = WarmStartMLP(n, n_branch, ft.shape[1], hidden, seed)
    net.set_normalization(x_mean, x_std, y_mean, y_std)
    tgn = ((tg - y_mean) / y_std).astype(np.float32)
    rng = np.random.default_rng(seed)
    hist = []
    S = len(br)
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(S)
        lr_ep = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))
        tot = 0.0
        nb = 0
        for q in range(0, S - batch + 1, batch):
            idx = order[q:q + batch]
            tot += net.train_step(xb[lvl[idx]], br[idx], ft[idx], tgn[idx], lr_ep)
            nb += 1
        hist.append(tot / max(nb, 1))
        if verbose:
            print(f"  epoch {ep}: loss {hist[-1]:.4f} ({time.perf_counter() - t0:.1f}s)", flush=True)
    return net, hist
