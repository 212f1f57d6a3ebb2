"""Seeded synthetic meshed transmission grids: the workloads the N-1 path is measured on.

The reference measures on real networks (case9241pegase, rte6515); the workloads here are synthetic grids built
from a seed, with every distribution fixed by the workload definition:

* buses at uniform random 2-D positions in the unit square;
* a random spanning tree: buses are visited in a random order and each one links to the nearest bus already
  placed (a geometric tree, so that the grid is planar-like, as a transmission grid is);
* chords to the 3 nearest neighbours of each bus (buses visited in a second random order, neighbours by
  increasing distance) until the branch count is reached;
* per branch x ~ U(0.01, 0.2) pu, r = x / U(3, 10), total charging b ~ U(0, 0.05) pu split b/2 per side;
  exactly 10 % of the branches are transformers with ratio U(0.95, 1.05) on side 1;
* bus 0 is the slack; exactly 20 % of the other buses are PV with V setpoint U(0.98, 1.05) and symmetric
  reactive limits +-U(50, 300) MVar (the slack also draws a setpoint and a capacity);
* every bus has a load P ~ lognormal(mu=3, sigma=0.8) MW and Q = P * U(0.2, 0.4);
* total generation is 1.02 x total load, split over the slack and PV generators in proportion to their
  capacity draw (the same U(50, 300) draw as the reactive limit).

Per-unit base 400 MVA. The workload definition gives impedances in pu and loads in MW but no base; at OLF's
100 MVA (``PerUnit.SB``) every generated grid is past its loadability (the flat-start 118-bus case reaches its
15-iteration budget at ||F|| = 23 pu), and at 1000 MVA line charging drives most PV buses to their reactive
limits; at 400 MVA all three base cases converge.

The packing into the solver's arrays follows the reference's extractor layout
(``GpuAcFamilyPacker.java:100-147``, ``StructureNative.java:39-46``): closed branches as 8 doubles
``(y, g1, g2, b1, b2, ksi, r1, a1)`` and 12 ints ``(P1 Q1 P2 Q2 ph1 v1 ph2 v2 a1col r1col altQ1 altQ2)`` over
stable rows (bus i owns row/col ``2i`` = phi / P and ``2i+1`` = V / Q or V target, ``docs/gpu-vs-olf.md:44-58``).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SB = 400.0  # MVA per-unit base: at 100 MVA every generated grid is past its loadability

# bus types
PQ, PV, SLACK = 0, 1, 2

# row modes (reference GpuRowMode, docs/gpu-vs-olf.md:60-76)
FIXED_VARIABLE, EQUATION, BUS_TARGET_Q = 0, 1, 4
# variable types (reference GpuAcVariableType, OLF AcVariableType ordinals)
BUS_V, BUS_PHI = 0, 1


@dataclass
class Network:
    """A synthetic grid in physical units (MW, MVar, pu impedances)."""

    n_bus: int
    bus1: np.ndarray          # int32 [E]
    bus2: np.ndarray          # int32 [E]
    r: np.ndarray             # [E] pu
    x: np.ndarray             # [E] pu
    b: np.ndarray             # [E] pu, total charging
    ratio: np.ndarray         # [E] side-1 ratio (1 for lines)
    bus_type: np.ndarray      # int8 [N]
    vset: np.ndarray          # [N] pu (PV and slack)
    qmin: np.ndarray          # [N] MVar (PV)
    qmax: np.ndarray          # [N] MVar (PV)
    pload: np.ndarray         # [N] MW
    qload: np.ndarray         # [N] MVar
    pgen: np.ndarray          # [N] MW
    pos: np.ndarray = field(repr=False, default=None)
    shunt_g: np.ndarray | None = field(repr=False, default=None)   # [N] pu bus shunt conductance (None: no shunts)
    shunt_b: np.ndarray | None = field(repr=False, default=None)   # [N] pu bus shunt susceptance

    @property
    def n_branch(self) -> int:
        return int(self.bus1.shape[0])

    @property
    def n(self) -> int:
        return 2 * self.n_bus

    def scaled(self, factor: float) -> "Network":
        """The same grid with every load and the generation scaled by `factor` (generation stays 1.02 x load)."""
        import dataclasses
        return dataclasses.replace(self, pload=self.pload * factor, qload=self.qload * factor, pgen=self.pgen * factor)

    # ---- packing (the solver's inputs) ----
    def pi_model(self) -> np.ndarray:
        """Closed-branch PiModel pack, 8 doubles per branch: y, g1, g2, b1, b2, ksi, r1, a1."""
        z = np.hypot(self.r, self.x)
        p = np.zeros((self.n_branch, 8))
        p[:, 0] = 1.0 / z
        p[:, 3] = self.b / 2
        p[:, 4] = self.b / 2
        p[:, 5] = np.arctan2(self.r, self.x)
        p[:, 6] = self.ratio
        return p

    def branch_indices(self) -> np.ndarray:
        """Closed-branch index pack, 12 ints per branch, on stable rows."""
        a = 2 * self.bus1.astype(np.int32)
        c = 2 * self.bus2.astype(np.int32)
        ix = np.full((self.n_branch, 12), -1, dtype=np.int32)
        ix[:, 0] = a
        ix[:, 1] = a + 1
        ix[:, 2] = c
        ix[:, 3] = c + 1
        ix[:, 4] = a
        ix[:, 5] = a + 1
        ix[:, 6] = c
        ix[:, 7] = c + 1
        return ix

    def system(self):
        """Flat-start state, targets, row modes and variable types over the 2N stable rows."""
        n = self.n
        x0 = np.zeros(n)
        x0[1::2] = 1.0
        target = np.zeros(n)
        mode = np.full(n, EQUATION, dtype=np.int32)
        var_type = np.empty(n, dtype=np.int32)
        var_type[0::2] = BUS_PHI
        var_type[1::2] = BUS_V
        target[0::2] = (self.pgen - self.pload) / SB
        target[1::2] = -self.qload / SB
        voltage = self.bus_type != PQ
        target[1::2][voltage] = self.vset[voltage]
        mode[1::2][voltage] = FIXED_VARIABLE
        slack = np.nonzero(self.bus_type == SLACK)[0]
        target[2 * slack] = 0.0
        mode[2 * slack] = FIXED_VARIABLE
        return x0, target, mode, var_type

    def shunts(self):
        """Bus-shunt pack for the structure: per shunt (g, b) in pu and its (P row, Q row, V column) stable rows;
        (None, None) without shunts."""
        if self.shunt_g is None:
            return None, None
        buses = np.nonzero((self.shunt_g != 0) | (self.shunt_b != 0))[0]
        params = np.stack([self.shunt_g[buses], self.shunt_b[buses]], axis=1)
        idx = np.stack([2 * buses, 2 * buses + 1, 2 * buses + 1], axis=1).astype(np.int32)
        return params, idx

    def with_shunts(self, fraction: float = 0.05, seed: int = 17) -> "Network":
        """The same grid with bus shunts on a fraction of the non-slack buses (not part of the workload definition;
        for the shunt family's parity): b ~ U(-0.15, 0.3) pu (reactors and capacitors), g ~ U(0, 0.005) pu."""
        import dataclasses
        rng = np.random.default_rng(seed)
        m = int(round(fraction * (self.n_bus - 1)))
        buses = 1 + rng.permutation(self.n_bus - 1)[:m]
        g = np.zeros(self.n_bus)
        b = np.zeros(self.n_bus)
        g[buses] = rng.uniform(0.0, 0.005, m)
        b[buses] = rng.uniform(-0.15, 0.3, m)
        return dataclasses.replace(self, shunt_g=g, shunt_b=b)

    def with_series_resonance(self, branches, x: float = 0.05) -> "Network":
        """The same grid with, beside each listed branch (bus a to bus b), a path through a new load-free PQ bus i:
        a lossless branch a-i of reactance x and one i-b of reactance -x, no charging (not part of the workload
        definition; a test network for pivoting). The two reactances cancel, so the path ties a and b together
        and bus i's own 2x2 Jacobian block is singular at the flat start and at the solution, while the whole
        Jacobian is not: an elimination that takes bus i before a and b without pivoting meets a zero pivot block."""
        import dataclasses
        branches = np.asarray(branches, dtype=np.int64).reshape(-1)
        m = branches.size
        new = self.n_bus + np.arange(m)

        def grow(a, fill):
            return None if a is None else np.concatenate([a, np.full(m, fill, dtype=a.dtype)])
        bus1 = np.concatenate([self.bus1, self.bus1[branches], new]).astype(self.bus1.dtype)
        bus2 = np.concatenate([self.bus2, new, self.bus2[branches]]).astype(self.bus2.dtype)
        e = np.zeros(2 * m)
        pos = None if self.pos is None else np.concatenate(
            [self.pos, 0.5 * (self.pos[self.bus1[branches]] + self.pos[self.bus2[branches]])])
        return dataclasses.replace(
            self, n_bus=self.n_bus + m, bus1=bus1, bus2=bus2, r=np.concatenate([self.r, e]),
            x=np.concatenate([self.x, np.full(m, x), np.full(m, -x)]), b=np.concatenate([self.b, e]),
            ratio=np.concatenate([self.ratio, np.ones(2 * m)]), bus_type=grow(self.bus_type, PQ),
            vset=grow(self.vset, 1.0), qmin=grow(self.qmin, 0.0), qmax=grow(self.qmax, 0.0),
            pload=grow(self.pload, 0.0), qload=grow(self.qload, 0.0), pgen=grow(self.pgen, 0.0), pos=pos,
            shunt_g=grow(self.shunt_g, 0.0), shunt_b=grow(self.shunt_b, 0.0))

    def reactive_limits(self):
        """Reactive-limit entries of the PV buses (pu): voltage row, minQ, maxQ, loadQ, targetV, bus."""
        pv = np.nonzero(self.bus_type == PV)[0].astype(np.int32)
        return dict(
            voltage_row=(2 * pv + 1).astype(np.int32),
            min_q=self.qmin[pv] / SB,
            max_q=self.qmax[pv] / SB,
            load_q=self.qload[pv] / SB,
            target_v=self.vset[pv].copy(),
            bus=pv,
        )


def generate(n_bus: int, n_branch: int, seed: int, load_scale: float = 1.0) -> Network:
    """Build the seeded grid. Deterministic for a given (n_bus, n_branch, seed, load_scale)."""
    if n_branch < n_bus - 1:
        raise ValueError("a connected grid needs at least n_bus - 1 branches")
    rng = np.random.default_rng(seed)
    pos = rng.random((n_bus, 2))
    order = rng.permutation(n_bus)
    chord_order = rng.permutation(n_bus)

    # geometric random spanning tree: each bus in `order` links to the nearest bus placed before it
    best_d = np.full(n_bus, np.inf)
    best_j = np.full(n_bus, -1, dtype=np.int64)
    placed = np.zeros(n_bus, dtype=bool)
    first = order[0]
    placed[first] = True
    d = ((pos - pos[first]) ** 2).sum(1)
    upd = d < best_d
    best_d[upd] = d[upd]
    best_j[upd] = first
    edges_a, edges_b = [], []
    seen = set()
    for t in range(1, n_bus):
        i = int(order[t])
        j = int(best_j[i])
        edges_a.append(j)
        edges_b.append(i)
        seen.add((min(i, j), max(i, j)))
        placed[i] = True
        d = ((pos - pos[i]) ** 2).sum(1)
        upd = (d < best_d) & ~placed
        best_d[upd] = d[upd]
        best_j[upd] = i

    # chords to the k nearest neighbours until the branch count is reached
    k = 3
    while len(edges_a) < n_branch:
        nn = _knn(pos, k)
        for i in chord_order:
            i = int(i)
            for j in nn[i]:
                j = int(j)
                key = (min(i, j), max(i, j))
                if key in seen:
                    continue
                seen.add(key)
                edges_a.append(i)
                edges_b.append(j)
                if len(edges_a) >= n_branch:
                    break
            if len(edges_a) >= n_branch:
                break
        k += 1
        if k > n_bus - 1:
            raise ValueError("too many branches for a simple graph")

    e = n_branch
    bus1 = np.asarray(edges_a, dtype=np.int32)
    bus2 = np.asarray(edges_b, dtype=np.int32)
    x = rng.uniform(0.01, 0.2, e)
    r = x / rng.uniform(3.0, 10.0, e)
    b = rng.uniform(0.0, 0.05, e)
    ratio = np.ones(e)
    n_tr = int(round(0.1 * e))
    tr = rng.permutation(e)[:n_tr]
    ratio[tr] = rng.uniform(0.95, 1.05, n_tr)

    n_pv = int(round(0.2 * n_bus))
    pv = 1 + rng.permutation(n_bus - 1)[:n_pv]
    gens = np.concatenate([[0], pv])
    bus_type = np.zeros(n_bus, dtype=np.int8)
    bus_type[pv] = PV
    bus_type[0] = SLACK
    vset = np.ones(n_bus)
    vset[gens] = rng.uniform(0.98, 1.05, n_pv + 1)
    cap = rng.uniform(50.0, 300.0, n_pv + 1)
    qmin = np.zeros(n_bus)
    qmax = np.zeros(n_bus)
    qmin[pv] = -cap[1:]
    qmax[pv] = cap[1:]

    pload = rng.lognormal(3.0, 0.8, n_bus) * load_scale
    qload = pload * rng.uniform(0.2, 0.4, n_bus)
    pgen = np.zeros(n_bus)
    pgen[gens] = 1.02 * pload.sum() * cap / cap.sum()
    return Network(n_bus, bus1, bus2, r, x, b, ratio, bus_type, vset, qmin, qmax, pload, qload, pgen, pos)


def _knn(pos: np.ndarray, k: int) -> np.ndarray:
    """The k nearest other buses of every bus, by increasing distance (ties: lower index)."""
    try:
        from scipy.spatial import cKDTree
        _, nn = cKDTree(pos).query(pos, k=k + 1)
        return nn[:, 1:]
    except ImportError:  # pragma: no cover - scipy ships with the image
        out = np.empty((pos.shape[0], k), dtype=np.int64)
        for i in range(pos.shape[0]):
            d = ((pos - pos[i]) ** 2).sum(1)
            d[i] = np.inf
            out[i] = np.argsort(d, kind="stable")[:k]
        return out


def bridges(net: Network) -> np.ndarray:
    """Boolean mask of bridge branches (their outage splits the grid), by an iterative Tarjan low-link walk."""
    n, e = net.n_bus, net.n_branch
    adj_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(adj_ptr, net.bus1 + 1, 1)
    np.add.at(adj_ptr, net.bus2 + 1, 1)
    adj_ptr = np.cumsum(adj_ptr)
    adj_to = np.empty(2 * e, dtype=np.int64)
    adj_br = np.empty(2 * e, dtype=np.int64)
    fill = adj_ptr[:-1].copy()
    for br in range(e):
        a, c = int(net.bus1[br]), int(net.bus2[br])
        adj_to[fill[a]] = c
        adj_br[fill[a]] = br
        fill[a] += 1
        adj_to[fill[c]] = a
        adj_br[fill[c]] = br
        fill[c] += 1
    disc = np.full(n, -1, dtype=np.int64)
    low = np.zeros(n, dtype=np.int64)
    is_bridge = np.zeros(e, dtype=bool)
    timer = 0
    for root in range(n):
        if disc[root] >= 0:
            continue
        disc[root] = low[root] = timer
        timer += 1
        stack = [(root, -1, int(adj_ptr[root]))]
        while stack:
            u, via, p = stack[-1]
            if p < adj_ptr[u + 1]:
                stack[-1] = (u, via, p + 1)
                w, br = int(adj_to[p]), int(adj_br[p])
                if br == via:
                    continue
                if disc[w] < 0:
                    disc[w] = low[w] = timer
                    timer += 1
                    stack.append((w, br, int(adj_ptr[w])))
                else:
                    low[u] = min(low[u], disc[w])
            else:
                stack.pop()
                if stack:
                    parent = stack[-1][0]
                    low[parent] = min(low[parent], low[u])
                    if low[u] > disc[parent]:
                        is_bridge[via] = True
    return is_bridge


def sample_contingencies(net: Network, count: int, bridge_fraction: float, seed: int) -> np.ndarray:
    """Single-branch outages: round(count*bridge_fraction) bridges, the rest non-bridges, uniformly without
    replacement, in one shuffled order."""
    rng = np.random.default_rng(seed)
    br = bridges(net)
    nb = int(round(count * bridge_fraction))
    bridge_ids = np.nonzero(br)[0]
    plain_ids = np.nonzero(~br)[0]
    if nb > bridge_ids.size or count - nb > plain_ids.size:
        raise ValueError("not enough branches of the requested kind")
    pick = np.concatenate([rng.permutation(bridge_ids)[:nb], rng.permutation(plain_ids)[:count - nb]])
    return rng.permutation(pick).astype(np.int32)


# The workloads, by name: (buses, branches, grid seed, load scale)
WORKLOADS = {
    "grid118": (118, 186, 118, 1.0),
    "grid2000": (2000, 3100, 2000, 1.0),
    "grid9241": (9241, 16000, 9241, 1.1),
}


def workload(name: str, load_scale: float | None = None) -> Network:
    nb, ne, seed, scale = WORKLOADS[name]
    return generate(nb, ne, seed, scale if load_scale is None else load_scale)


def branch_flows(net: Network, x: np.ndarray):
    """Textbook pi-branch flows at a state (stable rows): S1, S2 (pu, complex) and |V1|, |V2| per branch, with the
    ideal transformer of ratio r on side 1."""
    x = np.asarray(x, np.float64)
    th, v = x[0::2], x[1::2]
    a, b = net.bus1, net.bus2
    Va = v[a] * np.exp(1j * th[a]) * net.ratio
    Vb = v[b] * np.exp(1j * th[b])
    ys = 1.0 / (net.r + 1j * net.x)
    bc = 1j * net.b / 2
    S1 = Va * np.conj((Va - Vb) * ys + Va * bc)
    S2 = Vb * np.conj((Vb - Va) * ys + Vb * bc)
    return S1, S2, v[a], v[b]


def limit_table(net: Network, x_base: np.ndarray, seed: int = 7, angle_limits: int = 0) -> dict:
    """Synthetic operational limits for violation detection (the workload definition carries none), rated from the
    base case so that outages overload some neighbours:

    * 10 % of the branches unmonitored; every other branch has a current group on each side with thresholds
      [1.5, 1.25, 1.0] x PATL, most severe first, PATL = max(1.1 - 1.5 x the larger base-case end current, 0.02) pu;
    * 5 % of the monitored branches also an active-power group on side 1 ([1.3, 1.0] x a rating of the same kind on
      |P1|), every monitored transformer an apparent-power group on side 1 ([1.0] x a rating on |S1|);
    * every bus a low limit U(0.90, 0.95) and a high limit U(1.05, 1.10) pu, 10 % of each NaN (no limit);
    * angle_limits voltage-angle limits across as many distinct branches (bus1 -> bus2), the base-case difference
      +- U(0.0002, 0.003) rad (N-1 moves them by that much), 10 % of each bound NaN (drawn apart, so the other
      entries do not depend on it).
    Returns the table as arrays in the C ABI's layout (olfhip_limit_table_desc)."""
    rng = np.random.default_rng(seed)
    S1, S2, v1, v2 = branch_flows(net, x_base)
    i_base = np.maximum(np.abs(S1) / v1, np.abs(S2) / v2)
    monitored = rng.random(net.n_branch) >= 0.1
    factor = rng.uniform(1.1, 1.5, net.n_branch)
    with_p = rng.random(net.n_branch) < 0.05
    is_tr = net.ratio != 1.0
    group_start = [0]
    group_side, group_qty, limit_start, threshold = [], [], [0], []

    def group(side, qty, rating, mults):
        group_side.append(side)
        group_qty.append(qty)
        threshold.extend(m * rating for m in mults)
        limit_start.append(len(threshold))

    for e in range(net.n_branch):
        if monitored[e]:
            patl = max(factor[e] * i_base[e], 0.02)
            group(1, 0, patl, (1.5, 1.25, 1.0))
            group(2, 0, patl, (1.5, 1.25, 1.0))
            if with_p[e]:
                group(1, 1, max(factor[e] * abs(S1[e].real), 0.02), (1.3, 1.0))
            if is_tr[e]:
                group(1, 2, max(factor[e] * abs(S1[e]), 0.02), (1.0,))
        group_start.append(len(group_side))
    low = rng.uniform(0.90, 0.95, net.n_bus)
    high = rng.uniform(1.05, 1.10, net.n_bus)
    low[rng.random(net.n_bus) < 0.1] = np.nan
    high[rng.random(net.n_bus) < 0.1] = np.nan
    table = dict(branch=np.arange(net.n_branch, dtype=np.int32), group_start=np.array(group_start, np.int32),
                 group_side=np.array(group_side, np.int32), group_quantity=np.array(group_qty, np.int32),
                 limit_start=np.array(limit_start, np.int32), threshold=np.array(threshold, np.float64),
                 bus_high=high, bus_low=low)
    if angle_limits:
        ra = np.random.default_rng(seed + 1)
        br = ra.choice(net.n_branch, angle_limits, replace=False)
        fr, to = net.bus1[br].astype(np.int32), net.bus2[br].astype(np.int32)
        d = x_base[2 * to] - x_base[2 * fr]
        ahigh = d + ra.uniform(0.0002, 0.003, angle_limits)
        alow = d - ra.uniform(0.0002, 0.003, angle_limits)
        ahigh[ra.random(angle_limits) < 0.1] = np.nan
        alow[ra.random(angle_limits) < 0.1] = np.nan
        table.update(angle_from=fr, angle_to=to, angle_high=ahigh, angle_low=alow)
    return table


def distributed_slack(net: Network, target: np.ndarray | None = None, headroom: float = 1.25) -> dict:
    """Distributed slack over the PV generators, OLF's default balance type PROPORTIONAL_TO_GENERATION_P_MAX: each
    generator's factor is its Pmax over the Pmax of every generator, the slack's included, so the slack keeps its own
    share (1 - the participants' sum). The workload definition gives no active limits: Pmax = headroom x the
    generator's dispatch, Pmin = 0. The anchor is the participants' P-row targets in `target` (default: the
    dispatch of `net.system()`), the slack's own P target its dispatch minus its load; pu on SB."""
    if target is None:
        target = net.system()[1]
    gens = np.nonzero(net.bus_type != PQ)[0]
    pv = np.nonzero(net.bus_type == PV)[0].astype(np.int32)
    pmax = headroom * net.pgen
    base = np.asarray(target, dtype=np.float64)[2 * pv].copy()
    dispatch = base * SB + net.pload[pv]
    slack = int(np.nonzero(net.bus_type == SLACK)[0][0])
    return dict(bus=pv, factor=pmax[pv] / pmax[gens].sum(), base_target_p=base,
                max_delta=(pmax[pv] - dispatch) / SB, min_delta=(0.0 - dispatch) / SB, slack_bus=slack,
                slack_target_p=(net.pgen[slack] - net.pload[slack]) / SB, base_mva=SB)
