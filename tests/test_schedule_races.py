"""CPU race check of the LU level schedule (olfhip_structure_schedule), the property the kernels' one barrier per
level relies on (the reference checks the same of its transposed plan, native/host/lu_transposed_check.cpp:10-13):

* every factor slot is written by exactly one item (no two items, of one level or of any, write one slot);
* every contribution an item of level l reads (L~(k,j), U(j,k), L~(i,j), U(j,i), D(j)^-1) is written by an item of a
  level below l, or by no item at all: then it is a pair of the Jacobian's own pattern, which the fill wrote before
  level 0 (an item with an empty contribution list is left out of the schedule); the level a record states is its
  node j's;
* an item's contributions ascend in the contributing node j, and j is a node below the item's own;
* the backward solve, one phase per level from the root, reads x(i) of S(k) only from levels above node k's.

A planted error (a contribution moved to its reader's level, a slot written twice) fails these checks; the test
plants one of each on a copy of the records."""
import numpy as np
import pytest

from olfgpu import _native as N


def race_report(sc, n_bus, adjacent=None):
    """Problems found in a schedule (empty: none). adjacent(k, i) tells whether nodes k and i are coupled in the
    Jacobian's pattern: a pair slot no item writes holds the Jacobian's value from the fill and must be such a pair."""
    items, contrib = sc["item"], sc["contrib"]
    lip, lnp = sc["level_item_ptr"], sc["level_node_ptr"]
    n_levels = len(lip) - 1
    item_level = np.repeat(np.arange(n_levels), np.diff(lip))
    node_level = np.empty(n_bus, np.int64)
    node_level[sc["level_nodes"]] = np.repeat(np.arange(n_levels), np.diff(lnp))
    diag = items[:, 3] >= 0
    # slots each item writes: its diagonal slot, or the pair L~(i,k) and U(k,i) right after it
    writes = np.concatenate([items[diag, 0], items[~diag, 0], items[~diag, 0] + 1])
    writer = np.concatenate([np.nonzero(diag)[0], np.nonzero(~diag)[0], np.nonzero(~diag)[0]])
    problems = []
    uniq, counts = np.unique(writes, return_counts=True)
    if (counts > 1).any():
        problems.append(f"slots written twice: {uniq[counts > 1][:5].tolist()}")
    # slots of node k: its diagonal at its base, then L~(i,k) and U(k,i) for the e-th member i of S(k); a slot no
    # item writes keeps the Jacobian value the fill wrote (ready before level 0, level -1)
    diag_slot = np.empty(n_bus, np.int64)
    diag_slot[items[diag, 3]] = items[diag, 0]
    members = sc["s_members"]
    bwd = sc["bwd"]
    n_slots = int(max(writes.max(), (diag_slot + 1 + 2 * (bwd[np.argsort(bwd[:, 3]), 2] -
                                                          bwd[np.argsort(bwd[:, 3]), 1])).max())) + 1
    slot_level = np.full(n_slots, -1)
    slot_level[writes] = item_level[writer]
    written = np.zeros(n_slots, bool)
    written[writes] = True
    if adjacent is not None:
        for k_rec in bwd:
            k = k_rec[3]
            for e, i in enumerate(members[k_rec[1]:k_rec[2]]):
                sl = diag_slot[k] + 1 + 2 * e
                if not (written[sl] and written[sl + 1]) and not adjacent(k, i):
                    problems.append(f"fill-in pair ({k}, {i}) read before any item writes it")
    for q in range(len(items)):
        c0, c1 = items[q, 1], items[q, 2]
        if c1 <= c0:
            continue
        rec = contrib[c0:c1]
        lvl = item_level[q]
        if items[q, 3] >= 0:
            reads = np.concatenate([rec[:, 0], rec[:, 0] + 1, rec[:, 2]])
        else:
            reads = np.concatenate([rec[:, 0], rec[:, 0] + 1, rec[:, 1], rec[:, 1] + 1, rec[:, 2]])
        if (slot_level[reads] >= lvl).any():
            problems.append(f"item {q} (level {lvl}) reads a slot of its own level or above")
        j = rec[:, 1] if items[q, 3] >= 0 else None
        if j is not None:
            if (np.diff(j) <= 0).any():
                problems.append(f"item {q}: contributions not ascending in j")
            if (node_level[j] != rec[:, 3]).any() or (diag_slot[j] != rec[:, 2]).any():
                problems.append(f"item {q}: a record states another level or diagonal slot than node j's")
        if (slot_level[rec[:, 2]] != rec[:, 3]).any():
            problems.append(f"item {q}: a record's level is not its writer's")
    # backward: node k (level l) reads x(i), i in S(k), finalised by the phases of levels above l
    for k_rec in bwd:
        k = k_rec[3]
        s = members[k_rec[1]:k_rec[2]]
        if s.size and (node_level[s] <= node_level[k]).any():
            problems.append(f"backward of node {k} reads a node of its level or below")
    return problems


def coupling(net, structure):
    """adjacent(k, i) over node numbers: the two buses share a branch."""
    perm = structure.order()
    pairs = set(zip(net.bus1.tolist(), net.bus2.tolist()))
    return lambda k, i: (int(perm[k]), int(perm[i])) in pairs or (int(perm[i]), int(perm[k])) in pairs


@pytest.mark.parametrize("name", ["grid118", "grid2000", "grid9241"])
def test_level_schedule_has_no_races(grid, name):
    net = grid(name)
    st = N.Structure.from_network(net)
    assert race_report(st.schedule(), net.n_bus, coupling(net, st)) == []


def test_planted_races_are_found(grid):
    net = grid("grid2000")
    sc = N.Structure.from_network(net).schedule()
    lip = sc["level_item_ptr"]
    items = sc["item"]
    # a contribution whose writer sits at its reader's level: exchange the writer of a slot an item of level 3
    # reads with that item (levels are ranges of the item list)
    writer_of = {}
    for w in range(len(items)):
        writer_of[int(items[w, 0])] = w
        if items[w, 3] < 0:
            writer_of[int(items[w, 0]) + 1] = w
    q, w = next((q, writer_of[int(c[0])]) for q in range(lip[3], lip[4])
                for c in sc["contrib"][items[q, 1]:items[q, 2]] if int(c[0]) in writer_of)
    assert w < lip[3]
    bad = {k: v.copy() for k, v in sc.items()}
    bad["item"][[w, q]] = bad["item"][[q, w]]
    assert race_report(bad, net.n_bus) != []
    assert race_report(sc, net.n_bus) == []
    # a slot written twice
    bad = {k: v.copy() for k, v in sc.items()}
    bad["item"][1, 0] = bad["item"][0, 0]
    assert any("twice" in p for p in race_report(bad, net.n_bus))
