"""Host logic of OLF's LimitViolationManager over the scan's feed (CPU): one violation per key, the last one fed, and
the increased-violations comparison with the base case (olfgpu.violations.manage).

Known answers, each from the rule as OLF states it (GpuBaseViolations.java:100-170, the reference's restatement of
LimitViolationManager.violationWeakenedOrEquivalent; docs/increased-violations-thresholds.md:20-48 for the voltage
legs): a dropped epsilon, a max in place of the min, a key that includes the type, the first record kept in place of
the last, or a base lookup by type would each fail one case."""
import numpy as np
import pytest

from olfgpu import _native as N
from olfgpu.violations import IncreasedViolationsParameters, manage

NB = 10   # table branches


def rec(*rows):
    """rows of (scenario, kind, element, side, quantity, limit, value, threshold)"""
    return np.array([tuple(r) for r in rows], N.VIOLATION_DTYPE)


def cur(s, e, side, value, limit=0, threshold=1.0, q=0):
    return (s, 0, e, side, q, limit, value, threshold)


def high(s, bus, value, threshold=1.05):
    return (s, 1, bus, 0, 0, -1, value, threshold)


def low(s, bus, value, threshold=0.95):
    return (s, 2, bus, 0, 0, -1, value, threshold)


def kept(post, base=None, **kw):
    return [tuple(r)[1:5] + (float(r["value"]),) for r in manage(rec(*post), NB, rec(*base) if base else None, **kw)]


def test_no_base_keeps_the_last_violation_per_key_at_its_first_place():
    post = [cur(0, 3, 1, 1.3), cur(0, 3, 2, 1.2), cur(0, 3, 1, 0.9, limit=5, threshold=0.8, q=1),
            high(0, 4, 1.07), low(0, 4, 0.80), high(0, 5, 1.08), cur(1, 3, 1, 1.4)]
    got = kept(post)
    # side 1's active power overwrites its current (one key per side and limits group), in the current's place;
    # bus 4's low overwrites its high; scenario 1 has a map of its own
    assert got == [(0, 3, 1, 1, 0.9), (0, 3, 2, 0, 1.2), (2, 4, 0, 0, 0.80), (1, 5, 0, 0, 1.08), (0, 3, 1, 0, 1.4)]


def test_flow_within_the_proportional_threshold_is_dropped():
    base = [cur(0, 2, 1, 1.2)]
    at = 1.2 * 1.1
    assert kept([cur(3, 2, 1, at)], base) == []                       # |v2| = |v1| (1 + 0.1): equivalent
    assert kept([cur(3, 2, 1, -at)], base) == []                      # magnitudes compared
    assert len(kept([cur(3, 2, 1, at * (1 + 1e-9))], base)) == 1       # beyond: reported
    assert len(kept([cur(3, 2, 1, at)], base,
                    params=IncreasedViolationsParameters(flowProportionalThreshold=0.05))) == 1
    assert len(kept([cur(3, 2, 2, 1.2)], base)) == 1                  # the other side: another key
    # a value within eps of the band counts as equal (OLF's EQUALITY_EPSILON)
    assert kept([cur(3, 2, 1, 1.2 * (1.1 + 0.5e-12))], base) == []


def test_limit_order_decides_before_the_value():
    base = [cur(0, 2, 1, 1.30, limit=1, threshold=1.25)]
    # a smaller limit exceeded: weaker, dropped whatever the value
    assert kept([cur(1, 2, 1, 5.0, limit=2, threshold=1.0)], base) == []
    # a larger limit exceeded: increased, reported however close the value
    assert len(kept([cur(1, 2, 1, 1.30, limit=0, threshold=1.5)], base)) == 1


def test_base_key_keeps_its_last_violation_and_types_must_match():
    # base: side 1 current then apparent power; the map keeps the apparent power
    base = [cur(0, 6, 1, 1.2), cur(0, 6, 1, 2.0, limit=7, threshold=1.9, q=2)]
    assert len(kept([cur(2, 6, 1, 1.2)], base)) == 1                           # current: no current in the map
    assert kept([cur(2, 6, 1, 2.0, limit=7, threshold=1.9, q=2)], base) == []  # apparent power: equivalent


def test_voltage_thresholds_default_to_a_zero_band():
    base = [high(0, 1, 1.07), low(0, 2, 0.90)]
    one_ulp_up = np.nextafter(1.07, 2.0)
    assert kept([high(1, 1, 1.07)], base) == []
    # a ulp is within |v1| eps: equivalent; 1e-9 pu is beyond the zero band
    assert kept([high(1, 1, one_ulp_up)], base) == []
    assert len(kept([high(1, 1, 1.07 + 1e-9)], base)) == 1
    assert kept([low(1, 2, 0.90 + 1e-9)], base) == []                   # a low voltage that rose: weaker
    assert len(kept([low(1, 2, 0.90 - 1e-9)], base)) == 1


@pytest.mark.parametrize("kv", [400.0, np.full(8, 400.0)])
def test_voltage_legs_combine_with_min_in_kv(kv):
    base = [high(0, 1, 1.05 * 1.0)]                                     # 420 kV
    v1 = 420.0
    only_abs = IncreasedViolationsParameters(highVoltageAbsoluteThreshold=0.5)
    # the absolute leg alone: min(0.5, v1 * 0) = 0, no band
    assert len(kept([high(1, 1, (v1 + 0.1) / 400.0)], base, params=only_abs, bus_nominal_kv=kv)) == 1
    both = IncreasedViolationsParameters(highVoltageAbsoluteThreshold=0.5, highVoltageProportionalThreshold=0.001)
    # min(0.5 kV, 0.42 kV) = 0.42 kV
    assert kept([high(1, 1, (v1 + 0.41) / 400.0)], base, params=both, bus_nominal_kv=kv) == []
    assert len(kept([high(1, 1, (v1 + 0.43) / 400.0)], base, params=both, bus_nominal_kv=kv)) == 1
    lowp = IncreasedViolationsParameters(lowVoltageAbsoluteThreshold=0.3, lowVoltageProportionalThreshold=0.01)
    lbase = [low(0, 3, 0.9)]                                            # 360 kV: min(0.3, 3.6) = 0.3 kV
    assert kept([low(1, 3, (360.0 - 0.29) / 400.0)], lbase, params=lowp, bus_nominal_kv=kv) == []
    assert len(kept([low(1, 3, (360.0 - 0.31) / 400.0)], lbase, params=lowp, bus_nominal_kv=kv)) == 1


def test_a_filtered_record_leaves_the_earlier_one_under_its_key():
    # side 1: current (new) then active power (equivalent to the base's): the current stays
    base = [cur(0, 4, 1, 0.9, limit=3, threshold=0.8, q=1)]
    post = [cur(5, 4, 1, 1.3), cur(5, 4, 1, 0.9, limit=3, threshold=0.8, q=1)]
    assert kept(post, base) == [(0, 4, 1, 0, 1.3)]


def test_base_from_several_states_is_refused():
    with pytest.raises(ValueError):
        manage(rec(cur(1, 1, 1, 1.0)), NB, rec(cur(0, 1, 1, 1.0), cur(1, 2, 1, 1.0)))


def angle(s, a, kind, value, threshold):
    return (s, kind, a, 0, 0, -1, value, threshold)


def test_voltage_angle_violations_have_a_key_per_limit_and_are_kept():
    # limit 2's high and low share its key (the last fed kept); limit 2 and bus 2 are different keys; a base
    # violation of the same angle limit filters nothing
    post = [high(0, 2, 1.07), angle(0, 2, 3, 0.4, 0.35), angle(0, 2, 4, 0.4, 0.5), angle(0, 3, 3, 0.6, 0.5)]
    assert kept(post) == [(1, 2, 0, 0, 1.07), (4, 2, 0, 0, 0.4), (3, 3, 0, 0, 0.6)]
    base = [angle(0, 3, 3, 0.6, 0.5), high(0, 2, 1.07)]
    # OLF's comparison of angle violations with the base case is not modelled: refused unless asked to keep them
    with pytest.raises(ValueError):
        kept(post, base)
    with pytest.raises(ValueError):
        kept([high(0, 2, 1.08)], base)
    assert kept(post, base, keep_voltage_angle_violations=True) == [(4, 2, 0, 0, 0.4), (3, 3, 0, 0, 0.6)]
    # a per-bus nominal voltage table indexes buses only, never an angle limit's index
    assert len(kept([angle(1, 9, 3, 0.6, 0.5)], base, bus_nominal_kv=np.full(3, 400.0),
                    keep_voltage_angle_violations=True)) == 1


def test_branch_element_beyond_the_table_is_refused():
    with pytest.raises(ValueError):
        manage(rec(cur(0, NB, 1, 1.0)), NB)


def test_near_base_boundary_flags_values_at_the_increased_violation_boundary():
    """A record whose comparison with the base case's record is decided within the relative band of the boundary
    (flows: |v1| (1 + flowProportionalThreshold); high voltage: v1 + min(absolute, v1 proportional) in kV) is
    flagged; one of another limit, a value off the boundary, or a key absent from the base is not."""
    from olfgpu.violations import near_base_boundary
    p = IncreasedViolationsParameters(flowProportionalThreshold=0.001, highVoltageAbsoluteThreshold=0.5,
                                      highVoltageProportionalThreshold=0.001)
    base = rec(cur(0, 2, 1, 1.2), high(0, 4, 1.06))
    edge = 1.2 * 1.001
    v_edge = (1.06 * 400.0 + min(0.5, 1.06 * 400.0 * 0.001)) / 400.0
    post = rec(cur(7, 2, 1, edge * (1 + 5e-5)), cur(7, 2, 1, -edge * (1 - 5e-5)), cur(7, 2, 1, edge * 1.01),
               cur(7, 2, 1, edge, limit=3, threshold=0.9), cur(7, 3, 1, edge), high(7, 4, v_edge * (1 + 2e-5)),
               high(7, 4, 1.08), low(7, 4, v_edge))
    got = near_base_boundary(post, NB, base, p, 400.0, band=1e-4)
    assert got.tolist() == [True, True, False, False, False, True, False, False]
    # a band of zero flags nothing off the boundary itself
    assert not near_base_boundary(post, NB, base, p, 400.0, band=0.0)[:3].any()


def test_near_base_boundary_equals_a_record_by_record_restatement():
    """On random feeds (keys shared between branch sides and buses of the same element number, repeated base keys, a
    kilovolt table), near_base_boundary equals a plain loop over the records that looks up the base case's last
    record of the same key (a branch side, or a bus for high and low alike), requires the same limit type and
    threshold, and tests |value - boundary| <= band |boundary|."""
    from olfgpu.violations import near_base_boundary
    rng = np.random.default_rng(5)
    for trial in range(40):
        n, m = int(rng.integers(1, 300)), int(rng.integers(1, 12))
        v = np.zeros(n, N.VIOLATION_DTYPE)
        v["scenario"] = np.sort(rng.integers(0, 6, n))
        v["kind"] = rng.integers(0, 3, n)
        v["element"] = rng.integers(0, 8, n)
        v["side"] = np.where(v["kind"] == 0, rng.integers(1, 3, n), 0)
        v["quantity"] = np.where(v["kind"] == 0, rng.integers(0, 3, n), 0)
        v["threshold"] = rng.integers(1, 3, n) * 0.5
        v["value"] = rng.uniform(0.45, 1.2, n) * np.where(rng.random(n) < 0.2, -1.0, 1.0)
        b = v[rng.integers(0, n, m)].copy()
        b["scenario"] = 0
        kv = rng.uniform(100.0, 400.0, 8)
        p = IncreasedViolationsParameters(0.05, 0.01, 2.0, 0.02, 1.0)
        # some records placed on their base record's boundary
        for i in rng.integers(0, n, n // 3):
            j = int(rng.integers(0, m))
            v[i] = b[j]
            v1 = b["value"][j] * (1.0 if b["kind"][j] == 0 else kv[b["element"][j]])
            bound = (abs(v1) * 1.05 if b["kind"][j] == 0 else
                     v1 + min(1.0, v1 * 0.02) if b["kind"][j] == 1 else v1 - min(2.0, v1 * 0.01))
            scale = 1.0 if b["kind"][j] == 0 else kv[b["element"][j]]
            v["value"][i] = bound * (1.0 + rng.uniform(-2e-4, 2e-4)) / scale
        band = 1e-4
        want = []
        for r in v:
            key = (0, r["element"], r["side"]) if r["kind"] == 0 else (1, r["element"], 0)
            ref = None
            for q in b:
                if ((0, q["element"], q["side"]) if q["kind"] == 0 else (1, q["element"], 0)) == key:
                    ref = q   # the last one fed
            if ref is None or r["threshold"] != ref["threshold"]:
                want.append(False)
                continue
            if r["kind"] == 0:
                if ref["kind"] != 0 or ref["quantity"] != r["quantity"]:
                    want.append(False)
                    continue
                bound, val = abs(ref["value"]) * 1.05, abs(r["value"])
            else:
                if ref["kind"] != r["kind"]:
                    want.append(False)
                    continue
                v1, val = ref["value"] * kv[ref["element"]], r["value"] * kv[r["element"]]
                bound = v1 + min(1.0, v1 * 0.02) if r["kind"] == 1 else v1 - min(2.0, v1 * 0.01)
            want.append(abs(val - bound) <= band * abs(bound))
        got = near_base_boundary(v, NB, b, p, kv, band=band)
        assert got.tolist() == want, trial
