"""CPU checks of the LU level schedule the kernels run (olfhip_structure_schedule): bucket structure, and an emulation
of the kernels' work-unit mapping for every tile width (each item's threads contiguous and aligned inside one wave,
every (item, lane, thread) unit exactly once)."""
import numpy as np
import pytest

from olfgpu import _native as N


def split_for(length):
    s = 1
    while s < 64 and 2 * s < length:
        s *= 2
    return s


def _units(buckets, T, diag_only=False):
    """The kernels' unitOf over a run of buckets: per unit (record, lane, thread, split)."""
    out = []
    for first, end, split, diag in buckets:
        if diag_only and diag != 1:
            continue
        S = min(split, 64 // T)
        for r in range(first, end):
            for s in range(S):
                for t in range(T):
                    out.append((r, t, s, S))
    return out


@pytest.mark.parametrize("name", ["grid118", "grid2000", "grid9241"])
def test_buckets_cover_levels_with_descending_splits(grid, name):
    net = grid(name)
    st = N.Structure.from_network(net)
    sc = st.schedule()
    items, buckets, lbp = sc["item"], sc["bucket"], sc["level_bucket_ptr"]
    lip, lnp = sc["level_item_ptr"], sc["level_node_ptr"]
    nlev = len(lip) - 1
    for l in range(nlev):
        bks = buckets[lbp[l]:lbp[l + 1]]
        # contiguous cover of the level's items; splits non-increasing, diagonal before pair on equal splits, one
        # bucket per (split, kind)
        assert bks[0][0] == lip[l] and bks[-1][1] == lip[l + 1]
        assert all(bks[q][1] == bks[q + 1][0] for q in range(len(bks) - 1))
        key = [(-b[2], -b[3]) for b in bks]
        assert all(key[q] < key[q + 1] for q in range(len(key) - 1))
        for first, end, split, diag in bks:
            its = items[first:end]
            assert np.all((its[:, 3] >= 0) == bool(diag))
            assert np.all([split_for(c1 - c0) == split for c0, c1 in its[:, 1:3]])
        # the level's diagonal items are its nodes
        assert sorted(items[lip[l]:lip[l + 1]][:, 3][items[lip[l]:lip[l + 1]][:, 3] >= 0].tolist()) == \
            sorted(sc["level_nodes"][lnp[l]:lnp[l + 1]].tolist())
    assert np.array_equal(items[sc["diag_item_of"]][:, 3], np.arange(net.n_bus))
    # backward buckets: every node of the level once, splits of |S(k)| descending
    bwd, bb, lbb = sc["bwd"], sc["bwd_bucket"], sc["level_bwd_bucket_ptr"]
    for l in range(nlev):
        bks = bb[lbb[l]:lbb[l + 1]]
        assert bks[0][0] == lnp[l] and bks[-1][1] == lnp[l + 1]
        assert np.all(np.diff(bks[:, 2]) < 0)
        for first, end, split, _ in bks:
            assert np.all([split_for(b - a) == split for a, b in bwd[first:end][:, 1:3]])
    assert sorted(bwd[:, 3].tolist()) == list(range(net.n_bus))


@pytest.mark.parametrize("T", [1, 2, 4, 8, 16, 32])
def test_work_units_are_wave_aligned(grid, T):
    net = grid("grid2000")
    sc = N.Structure.from_network(net).schedule()
    runs = ((sc["bucket"], sc["level_bucket_ptr"], False), (sc["bucket"], sc["level_bucket_ptr"], True),
            (sc["bwd_bucket"], sc["level_bwd_bucket_ptr"], False))
    for recs, lp, diag_only in runs:
        for l in range(0, len(lp) - 1, 3):
            units = _units(recs[lp[l]:lp[l + 1]], T, diag_only)
            seen = set()
            for u, (r, t, s, S) in enumerate(units):
                # a record's T*S units start at a multiple of T*S and stay in one 64-wide wave
                assert u % T == t
                base = u - (s * T + t)
                assert base % (T * S) == 0 and base // 64 == (base + T * S - 1) // 64
                seen.add((r, t, s))
            assert len(seen) == len(units)
