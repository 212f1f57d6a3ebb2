"""Learned warm start for the batched N-1: the network (over the C ABI), its features, its training set made by the
Newton solver itself, and its training loop.

The network predicts the post-contingency change on the outage's neighbourhood (the n_local buses electrically
nearest the outaged branch) from the neighbourhood's base state, the branch (an embedding row) and endpoint features.
What it learns is the response shape per unit of the branch's base flow: targets are divided by the flow scale
``|S1| + 0.01`` (S1 the base-case complex flow at side 1, pu) and the prediction is multiplied by it, so one
training level serves every load level. Angle changes are taken relative to the mean change of the neighbourhood's
outer quarter, so the start has no step where the neighbourhood ends: a Newton step's remaining error depends on the
start's branch-wise differences, and a step there would cost more than the prediction gains.

Training samples are post-contingency states the batched solver computes under random load scalings; every branch of
the grid is covered before any is repeated. A warm-started scenario that does not converge is solved again from the
base-case start, so the converged Newton mismatch alone decides every answer.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _native as N

PARAMS = ("Win", "Emb", "b1", "W2", "b2", "W3", "b3")
N_FEATURES = 16
N_LOCAL = 96


def _lib():
    L = N.lib()
    if not getattr(L, "_mlp_bound", False):
        vp, i, d, f = C.c_void_p, C.c_int, C.c_double, C.c_float
        L.olfhip_neighbourhoods.argtypes = [i, i, vp, vp, vp, i, vp]
        L.olfhip_mlp_create.argtypes = [i, i, vp, vp, vp, i, i, i, C.c_ulonglong, C.POINTER(vp)]
        L.olfhip_mlp_destroy.argtypes = [vp]
        L.olfhip_mlp_destroy.restype = None
        L.olfhip_mlp_set_normalization.argtypes = [vp, vp, vp, vp, vp]
        L.olfhip_mlp_set_base_states.argtypes = [vp, i, vp]
        L.olfhip_mlp_train_step.argtypes = [vp, i, vp, vp, vp, vp, f, C.POINTER(d)]
        L.olfhip_mlp_forward.argtypes = [vp, i, vp, vp, vp, vp]
        L.olfhip_mlp_gradients.argtypes = [vp, i, vp, vp, vp, vp, i, vp, C.POINTER(d)]
        L.olfhip_mlp_parameters.argtypes = [vp, i, vp, vp]
        L.olfhip_mlp_predict_warm_start.argtypes = [vp, i, vp, vp, vp, vp, C.POINTER(C.c_void_p)]
        L.olfhip_mlp_gemm_test.argtypes = [i, i, i, vp, vp, vp]
        L._mlp_bound = True
    return L


def _ptr_or_none(a):
    return None if a is None else N.ptr(a)


def gemm_test(A, B):
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    Cm = np.zeros((A.shape[0], B.shape[1]), np.float32)
    N.check(_lib().olfhip_mlp_gemm_test(A.shape[0], B.shape[1], A.shape[1], N.ptr(A), N.ptr(B), N.ptr(Cm)))
    return Cm


def neighbourhoods(n_bus, bus1, bus2, weight, n_local):
    """[E][n_local] bus ids of every branch's neighbourhood (host computation in the library)."""
    b1 = np.ascontiguousarray(bus1, np.int32)
    b2 = np.ascontiguousarray(bus2, np.int32)
    w = np.ascontiguousarray(weight, np.float64)
    out = np.zeros((len(b1), n_local), np.int32)
    N.check(_lib().olfhip_neighbourhoods(n_bus, len(b1), N.ptr(b1), N.ptr(b2), N.ptr(w), n_local, N.ptr(out)))
    return out


def flow_scale(features):
    """Per-scenario output scale |S1| + 0.01 from endpoint features 5 and 6 (base-case P1, Q1 in pu)."""
    f = np.asarray(features, np.float64)
    return (np.sqrt(f[:, 5] ** 2 + f[:, 6] ** 2) + 0.01).astype(np.float32)


def local_state(x_rows, base_row, hoods, branch):
    """[B][2K] local values of the base states: angles relative to endpoint 1, voltages minus 1 (0 on padding)."""
    h = hoods[np.asarray(branch)]
    x = np.asarray(x_rows, np.float64)[np.asarray(base_row)]
    valid = h >= 0
    hc = np.where(valid, h, 0)
    th = np.take_along_axis(x, 2 * hc, 1) - np.take_along_axis(x, 2 * hc[:, :1], 1)
    v = np.take_along_axis(x, 2 * hc + 1, 1) - 1.0
    return np.concatenate([np.where(valid, th, 0.0), np.where(valid, v, 0.0)], 1)


def local_change(dx, hoods, branch):
    """[B][2K] neighbourhood change of full state changes dx [B][n]: angle changes minus the mean change of the
    neighbourhood's outer quarter (slots [3k/4, k) of its k valid slots), voltage changes; 0 on padding."""
    h = hoods[np.asarray(branch)]
    valid = h >= 0
    hc = np.where(valid, h, 0)
    dth = np.where(valid, np.take_along_axis(dx, 2 * hc, 1), 0.0)
    dv = np.where(valid, np.take_along_axis(dx, 2 * hc + 1, 1), 0.0)
    k = valid.sum(1)
    slot = np.arange(h.shape[1])[None, :]
    ring = valid & (slot >= (3 * k // 4)[:, None])
    ref = (dth * ring).sum(1) / np.maximum(ring.sum(1), 1)
    dth = np.where(valid, dth - ref[:, None], 0.0)
    return np.concatenate([dth, dv], 1)


class WarmStartMLP:
    """Owner of an olfhip_mlp for one grid (gridgen.Network)."""

    def __init__(self, net, n_local=N_LOCAL, n_features=N_FEATURES, hidden=512, seed=0):
        self.n_bus, self.n_branch = net.n_bus, net.n_branch
        self.n_local = min(n_local, net.n_bus)
        self.n_features, self.hidden = n_features, hidden
        self.n_rows = 2 * self.n_local
        b1 = np.ascontiguousarray(net.bus1, np.int32)
        b2 = np.ascontiguousarray(net.bus2, np.int32)
        w = np.ascontiguousarray(net.x, np.float64)
        h = C.c_void_p()
        N.check(_lib().olfhip_mlp_create(net.n_bus, net.n_branch, N.ptr(b1), N.ptr(b2), N.ptr(w), self.n_local,
                                         n_features, hidden, seed, C.byref(h)))
        self.h = h
        self.hoods = neighbourhoods(net.n_bus, b1, b2, w, self.n_local)
        D = self.n_rows + n_features
        self.shapes = {"Win": (hidden, D), "Emb": (net.n_branch, hidden), "b1": (hidden,), "W2": (hidden, hidden),
                       "b2": (hidden,), "W3": (self.n_rows, hidden), "b3": (self.n_rows,)}
        self._base = None

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

    def set_base_states(self, x_rows):
        x = np.ascontiguousarray(np.atleast_2d(x_rows), np.float64)
        N.check(_lib().olfhip_mlp_set_base_states(self.h, x.shape[0], N.ptr(x)))
        self._base = x.copy()

    def _batch(self, base_row, branch, feat):
        br = np.ascontiguousarray(branch, np.int32)
        rows = None if base_row is None else np.ascontiguousarray(base_row, np.int32)
        ft = np.ascontiguousarray(feat, np.float32).reshape(len(br), self.n_features)
        return rows, br, ft

    def train_step(self, base_row, branch, feat, target, lr):
        rows, br, ft = self._batch(base_row, branch, feat)
        tg = np.ascontiguousarray(target, np.float32)
        loss = C.c_double()
        N.check(_lib().olfhip_mlp_train_step(self.h, len(br), _ptr_or_none(rows), N.ptr(br), N.ptr(ft), N.ptr(tg),
                                             lr, C.byref(loss)))
        return loss.value

    def forward(self, base_row, branch, feat):
        rows, br, ft = self._batch(base_row, branch, feat)
        y = np.zeros((len(br), self.n_rows), np.float32)
        N.check(_lib().olfhip_mlp_forward(self.h, len(br), _ptr_or_none(rows), N.ptr(br), N.ptr(ft), N.ptr(y)))
        return y

    def gradients(self, base_row, branch, feat, target, which):
        rows, br, ft = self._batch(base_row, branch, feat)
        tg = np.ascontiguousarray(target, np.float32)
        g = np.zeros(self.shapes[which], np.float32)
        loss = C.c_double()
        N.check(_lib().olfhip_mlp_gradients(self.h, len(br), _ptr_or_none(rows), N.ptr(br), N.ptr(ft), N.ptr(tg),
                                            PARAMS.index(which), N.ptr(g), C.byref(loss)))
        return g, loss.value

    def parameters(self, which):
        v = np.zeros(self.shapes[which], np.float32)
        N.check(_lib().olfhip_mlp_parameters(self.h, PARAMS.index(which), N.ptr(v), None))
        return v

    def set_parameters(self, which, values):
        v = np.ascontiguousarray(values, np.float32).reshape(self.shapes[which])
        N.check(_lib().olfhip_mlp_parameters(self.h, PARAMS.index(which), None, N.ptr(v)))

    def predict_raw(self, branch, feat, scale=None, base_row=None):
        """Device pointer (int) to [B][n] start states from the base states already set."""
        rows, br, ft = self._batch(base_row, branch, feat)
        sc = None if scale is None else np.ascontiguousarray(scale, np.float32)
        out = C.c_void_p()
        N.check(_lib().olfhip_mlp_predict_warm_start(self.h, len(br), _ptr_or_none(rows), N.ptr(br), N.ptr(ft),
                                                     _ptr_or_none(sc), C.byref(out)))
        return out.value

    def predict_warm_start(self, x_base, branch, feat):
        """Device pointer (int) to [B][n] start states for olfhip_solve_inputs.x0_per_scenario_device: x_base plus
        the predicted neighbourhood change, scaled by each outage's base flow."""
        x = np.asarray(x_base, np.float64).reshape(1, -1)
        if self._base is None or self._base.shape != x.shape or not np.array_equal(self._base, x):
            self.set_base_states(x)
        return self.predict_raw(branch, feat, flow_scale(feat))

    def warm_starts_host(self, x_base, branch, feat) -> np.ndarray:
        """The start states predict_warm_start hands the solver, copied to the host ([B][n])."""
        ptr = self.predict_warm_start(x_base, branch, feat)
        return N.device_to_host(ptr, (len(np.atleast_1d(branch)), 2 * self.n_bus))


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


def make_training_set(net, structure, hoods, n_samples=20000, levels=20, seed=3, scale_range=(0.9, 1.1),
                      verbose=False):
    """Post-contingency states solved by the batched Newton solver itself, at `levels` load scalings drawn from
    U(scale_range); the outages are consecutive slices of one random permutation of all branches, so every branch is
    covered before any repeats.

    Returns a dict with the base states x_base [levels, n], per sample its level, branch, endpoint features,
    neighbourhood target (local_change / flow_scale, unnormalised) [S, 2K] and the Newton iterations its solve took
    from the base start (all solves of its outer loop). Islanded buses count as unchanged (the solver freezes them
    whatever the start); outages whose solve does not converge are left out."""
    from . import gridgen
    from .solver import GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters, GpuNewtonRaphson

    rng = np.random.default_rng(seed)
    scales = rng.uniform(scale_range[0], scale_range[1], levels)
    per = int(np.ceil(n_samples / levels))
    order = rng.permutation(net.n_branch)
    bridges = gridgen.bridges(net)
    xb_all, lvl, brs, feats, targets, iters, lvl_scale = [], [], [], [], [], [], []
    t0 = time.perf_counter()
    pos = 0
    for k, s in enumerate(scales):
        idx = (pos + np.arange(per)) % net.n_branch
        pos = (pos + per) % net.n_branch
        ctg = order[idx].astype(np.int32)
        ns = net.scaled(s)
        rl = ns.reactive_limits()
        x0, tg, md, _ = ns.system()
        base = GpuNewtonRaphson(structure, rl).run(x0, tg, md)
        if base.status != "CONVERGED":
            continue
        sa = GpuBatchedSecurityAnalysisSolver(structure, rl,
                                              GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"))
        res = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg)
        ok = res.status == 0
        c = ctg[ok]
        d = res.x[ok] - base.x[None, :]
        for j, e in enumerate(c):
            if bridges[e]:
                isl = structure.islanded_buses([e])
                d[j, 0::2][isl] = 0.0
                d[j, 1::2][isl] = 0.0
        f = endpoint_features(net, base.x, c, s, bridges)
        xb_all.append(base.x)
        lvl_scale.append(s)
        lvl.append(np.full(len(c), len(xb_all) - 1, np.int32))
        brs.append(c)
        feats.append(f)
        targets.append((local_change(d, hoods, c) / flow_scale(f)[:, None]).astype(np.float32))
        iters.append(res.total_iterations[ok].astype(np.int32))
        if verbose:
            print(f"  level {k}: scale {s:.3f}, {ok.sum()}/{len(ctg)} converged, {time.perf_counter() - t0:.1f}s",
                  flush=True)
    return dict(x_base=np.array(xb_all), level=np.concatenate(lvl), branch=np.concatenate(brs).astype(np.int32),
                features=np.concatenate(feats), target=np.concatenate(targets), iterations=np.concatenate(iters),
                scales=scales, level_scale=np.array(lvl_scale), seconds=time.perf_counter() - t0)


def iteration_costs(data, n_branch):
    """Expected Newton iterations of each branch's outage: the mean over the training set's solves of it (at their
    load scalings), the mean of all solves for a branch the set did not cover."""
    total = np.bincount(data["branch"], weights=data["iterations"], minlength=n_branch)
    count = np.bincount(data["branch"], minlength=n_branch)
    cost = np.full(n_branch, float(data["iterations"].mean()))
    seen = count > 0
    cost[seen] = total[seen] / count[seen]
    return cost


def warm_iteration_costs(net, structure, mlp, data):
    """Expected Newton iterations of each branch's outage from the trained network's warm start: the training set's
    outages solved again at their load scalings, each from its prediction (the iterations of the warm-started attempt,
    all solves of its outer loop), averaged per branch; a branch the set did not cover gets the mean. Offline, once
    per network and model, as the training is: a schedule cost for the solves that start from the same network."""
    from . import gridgen
    from .solver import GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters, GpuNewtonRaphson

    bridges = gridgen.bridges(net)
    total = np.zeros(net.n_branch)
    count = np.zeros(net.n_branch)
    for k, s in enumerate(data["level_scale"]):
        ctg = data["branch"][data["level"] == k]
        if len(ctg) == 0:
            continue
        ns = net.scaled(float(s))
        rl = ns.reactive_limits()
        x0, tg, md, _ = ns.system()
        base = GpuNewtonRaphson(structure, rl).run(x0, tg, md)
        if base.status != "CONVERGED":
            continue
        sa = GpuBatchedSecurityAnalysisSolver(structure, rl,
                                              GpuLoadFlowParameters(linearSolver="BATCHED_LU_FP32_REFINED"))
        feats = endpoint_features(net, base.x, ctg, float(s), bridges)
        res = sa.solve_batch_n1(base.x, base.target, base.row_mode, base.reactive_limit_types, ctg, warm_start=mlp,
                                warm_start_features=feats, confirm_band=0)
        its = res.total_iterations.astype(np.float64)
        np.add.at(total, ctg, its)
        np.add.at(count, ctg, 1.0)
    cost = np.full(net.n_branch, float(total.sum() / max(count.sum(), 1.0)))
    seen = count > 0
    cost[seen] = total[seen] / count[seen]
    return cost


def train(mlp, data, epochs=60, batch=128, lr=2e-3, seed=0, verbose=False):
    """Fit `mlp` to a training set (cosine learning-rate decay); returns the loss history per epoch."""
    xb, lvl, br, ft, tg = data["x_base"], data["level"], data["branch"], data["features"], data["target"]
    K = mlp.n_local
    xl = local_state(xb, lvl, mlp.hoods, br)
    x_mean = xl.mean(0)
    x_std = np.maximum(xl.std(0), 1e-3)
    # one output scale per variable type, no offset
    y_mean = np.zeros(2 * K)
    y_std = np.empty(2 * K)
    y_std[:K] = max(float(np.sqrt((tg[:, :K].astype(np.float64) ** 2).mean())), 1e-6)
    y_std[K:] = max(float(np.sqrt((tg[:, K:].astype(np.float64) ** 2).mean())), 1e-6)
    mlp.set_normalization(x_mean, x_std, y_mean, y_std)
    mlp.set_base_states(xb)
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
            tot += mlp.train_step(lvl[idx], br[idx], ft[idx], tgn[idx], lr_ep)
            nb += 1
        hist.append(tot / max(nb, 1))
        if verbose:
            print(f"  epoch {ep}: loss {hist[-1]:.4f} ({time.perf_counter() - t0:.1f}s)", flush=True)
    mlp._base = None
    return hist
