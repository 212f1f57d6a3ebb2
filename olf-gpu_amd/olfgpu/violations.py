"""OLF's LimitViolationManager over the violations the scan feeds it, and its comparison with the base case.

The device screen plus the exact host decisions (LimitTable.detect) give every violation the scan finds, in the
order the scan feeds them to OLF's manager (GpuLimitScanner.java:93-135): per scenario, branches in table order (a
group's first exceeded limit), then buses (high before low). OLF's manager then does two things with that feed,
which the reference leaves to it and restates for its device screen (GpuBaseViolations.java:27-41,100-170):

* it keeps one violation per key, the last one fed: a branch's key is (branch, side, operational limits group), so
  the current, active-power and apparent-power groups of one side share a key (the synthetic tables carry one limits
  group per side); a bus's key is the bus, high and low sharing it; a voltage-angle limit's key is the limit;
* after a contingency, it drops a violation that is weaker than or equivalent to the base case's violation under the
  same key (violationWeakenedOrEquivalent): same type and a smaller limit, or the same limit and a value within the
  increased-violations thresholds (powsybl-core's SecurityAnalysisParameters.IncreasedViolationsParameters):

      flows:        |v2| <= |v1| (1 + flowProportionalThreshold + eps)
      high voltage:  v2  <=  v1 + min(highVoltageAbsoluteThreshold, v1 highVoltageProportionalThreshold) + |v1| eps
      low voltage:   v2  >=  v1 - min(lowVoltageAbsoluteThreshold,  v1 lowVoltageProportionalThreshold)  - |v1| eps

  with v1 the base value, v2 the post-contingency one, voltages in kV (the table's per-unit value times the bus's
  nominal voltage) and eps = 1e-12, OLF's EQUALITY_EPSILON (GpuBaseViolations.java:129-140). The voltage
  thresholds default to zero, so with the defaults a pre-existing voltage violation comes back as soon as its value
  moves past the base value by more than that relative margin; the two legs are combined with min, so an absolute
  threshold alone leaves a zero band (docs/increased-violations-thresholds.md:20-48). Voltage-angle violations are
  kept (parity unpinned: the reference leaves their comparison to OLF's manager, GpuBaseViolations.java:52-54).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native as N

EQUALITY_EPSILON = 1e-12

KIND_BRANCH, KIND_HIGH_VOLTAGE, KIND_LOW_VOLTAGE, KIND_HIGH_ANGLE, KIND_LOW_ANGLE = 0, 1, 2, 3, 4


@dataclass
class IncreasedViolationsParameters:
    """powsybl-core's SecurityAnalysisParameters.IncreasedViolationsParameters, with its defaults. Voltage thresholds
    absolute in kV, proportional as fractions; flowProportionalThreshold a fraction of the base value."""

    flowProportionalThreshold: float = 0.1
    lowVoltageProportionalThreshold: float = 0.0
    lowVoltageAbsoluteThreshold: float = 0.0
    highVoltageProportionalThreshold: float = 0.0
    highVoltageAbsoluteThreshold: float = 0.0


_CLASS = np.int64(1) << 32    # key classes: branch sides, buses, voltage-angle limits


def _keys(v: np.ndarray) -> np.ndarray:
    """OLF's manager key of each record, as an integer: 2 element + side - 1 for a branch side, then the bus, then
    the voltage-angle limit (its id: high and low share it), each class in a range of its own."""
    e = v["element"].astype(np.int64)
    return np.where(v["kind"] == KIND_BRANCH, 2 * e + v["side"] - 1,
                    np.where(v["kind"] <= KIND_LOW_VOLTAGE, _CLASS + e, 2 * _CLASS + e))


def _type(v: np.ndarray) -> np.ndarray:
    """OLF's LimitViolationType of each record as an integer: 0-2 the branch quantities, 3 high, 4 low voltage,
    5 high, 6 low voltage angle."""
    return np.where(v["kind"] == KIND_BRANCH, v["quantity"], 2 + v["kind"])


def _last_per_key(keys: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """For an array of keys in feed order: (the index of each key's last occurrence, of its first occurrence), both
    listed in the order of the first occurrences (a dict's insertion order)."""
    if keys.size == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    uniq, first = np.unique(keys, return_index=True)
    _, last_rev = np.unique(keys[::-1], return_index=True)
    last = keys.size - 1 - last_rev
    order = np.argsort(first, kind="stable")
    return last[order], first[order]


def weakened_or_equivalent(ref: np.ndarray, rep: np.ndarray, params: IncreasedViolationsParameters,
                           ref_kv: np.ndarray, rep_kv: np.ndarray) -> np.ndarray:
    """violationWeakenedOrEquivalent(reference, reported) element-wise over matched record arrays; ref_kv / rep_kv:
    the nominal voltage that turns each record's value into kV (1 for a branch). GpuBaseViolations.java:143-160."""
    same_type = _type(ref) == _type(rep)
    smaller = rep["threshold"] < ref["threshold"]
    same = rep["threshold"] == ref["threshold"]
    v1 = ref["value"] * ref_kv
    v2 = rep["value"] * rep_kv
    eps = EQUALITY_EPSILON
    flow = np.abs(v2) <= np.abs(v1) * (1.0 + params.flowProportionalThreshold + eps)
    high = v2 <= v1 + np.minimum(params.highVoltageAbsoluteThreshold, v1 * params.highVoltageProportionalThreshold) \
        + np.abs(v1) * eps
    low = v2 >= v1 - np.minimum(params.lowVoltageAbsoluteThreshold, v1 * params.lowVoltageProportionalThreshold) \
        - np.abs(v1) * eps
    within = np.where(rep["kind"] == KIND_BRANCH, flow, np.where(rep["kind"] == KIND_HIGH_VOLTAGE, high, low))
    # voltage-angle violations are kept: the reference's restatement leaves their comparison to OLF's own manager
    # (GpuBaseViolations.java:52-54), which is not restated here
    return same_type & (smaller | (same & within)) & (rep["kind"] <= KIND_LOW_VOLTAGE)


def near_base_boundary(violations: np.ndarray, n_table_branches: int, base: np.ndarray,
                       params: IncreasedViolationsParameters | None = None, bus_nominal_kv=None,
                       band: float = 1e-4) -> np.ndarray:
    """Per record of `violations`: whether the manager's comparison with the base case's record of the same key
    (same limit type and threshold) is decided by a value within relative `band` of its boundary (the base value
    raised by the increased-violation threshold), so that another iterate inside the solver's tolerance can decide
    it the other way."""
    v = np.asarray(violations, N.VIOLATION_DTYPE)
    out = np.zeros(v.size, bool)
    b = np.asarray(base, N.VIOLATION_DTYPE)
    if v.size == 0 or b.size == 0:
        return out
    params = params or IncreasedViolationsParameters()
    kv = None if bus_nominal_kv is None else np.asarray(bus_nominal_kv, np.float64)

    def to_kv(r):
        if kv is None:
            return np.ones(r.size)
        bus = (r["kind"] == KIND_HIGH_VOLTAGE) | (r["kind"] == KIND_LOW_VOLTAGE)
        bus_kv = kv if kv.ndim == 0 else kv[np.where(bus, r["element"], 0)]
        return np.where(bus, bus_kv, 1.0)

    # the base case's last record of each key, found for every record through a lookup table per key class (branch
    # sides, buses): a gather per record instead of a search
    blast, _ = _last_per_key(_keys(b))
    brec = b[blast]
    kind, el = v["kind"], v["element"]
    size = int(max(el.max(), brec["element"].max())) + 1
    lut_branch = np.full(2 * size, -1, np.int64)
    lut_bus = np.full(size, -1, np.int64)
    bb = brec["kind"] == KIND_BRANCH
    bv = (brec["kind"] == KIND_HIGH_VOLTAGE) | (brec["kind"] == KIND_LOW_VOLTAGE)
    lut_branch[2 * brec["element"][bb] + brec["side"][bb] - 1] = np.nonzero(bb)[0]
    lut_bus[brec["element"][bv]] = np.nonzero(bv)[0]
    pos = np.where(kind == KIND_BRANCH, lut_branch[np.clip(2 * el + v["side"] - 1, 0, 2 * size - 1)],
                   np.where(kind <= KIND_LOW_VOLTAGE, lut_bus[el], -1))
    hit = np.nonzero(pos >= 0)[0]
    if hit.size == 0:
        return out
    p = pos[hit]
    kind = kind[hit]
    bkind = brec["kind"][p]
    is_bus = (kind == KIND_HIGH_VOLTAGE) | (kind == KIND_LOW_VOLTAGE)
    if kv is None:
        rep_kv = ref_kv = 1.0
    elif kv.ndim == 0:
        rep_kv = np.where(is_bus, kv, 1.0)
        ref_kv = np.where((bkind == KIND_HIGH_VOLTAGE) | (bkind == KIND_LOW_VOLTAGE), kv, 1.0)
    else:
        rep_kv = np.where(is_bus, kv[np.where(is_bus, el[hit], 0)], 1.0)
        bbus = (bkind == KIND_HIGH_VOLTAGE) | (bkind == KIND_LOW_VOLTAGE)
        ref_kv = np.where(bbus, kv[np.where(bbus, brec["element"][p], 0)], 1.0)
    v1 = brec["value"][p] * ref_kv
    v2 = v["value"][hit] * rep_kv
    bound = np.where(kind == KIND_BRANCH, np.abs(v1) * (1.0 + params.flowProportionalThreshold),
                     np.where(kind == KIND_HIGH_VOLTAGE,
                              v1 + np.minimum(params.highVoltageAbsoluteThreshold, v1 * params.highVoltageProportionalThreshold),
                              v1 - np.minimum(params.lowVoltageAbsoluteThreshold, v1 * params.lowVoltageProportionalThreshold)))
    val = np.where(kind == KIND_BRANCH, np.abs(v2), v2)
    # same limit type (LimitViolationType: branch quantity, or the voltage kind) and the same threshold
    same_type = np.where(kind == KIND_BRANCH, (bkind == KIND_BRANCH) & (brec["quantity"][p] == v["quantity"][hit]),
                         bkind == kind)
    decided = same_type & (v["threshold"][hit] == brec["threshold"][p])
    out[hit] = decided & (np.abs(val - bound) <= band * np.abs(bound))
    return out


def manage(violations: np.ndarray, n_table_branches: int, base: np.ndarray | None = None,
           params: IncreasedViolationsParameters | None = None, bus_nominal_kv=None,
           keep_voltage_angle_violations: bool = False) -> np.ndarray:
    """OLF's LimitViolationManager fed `violations` (N.VIOLATION_DTYPE in the scan's feed order, any number of
    scenarios): per scenario, the records it keeps, in its map's order (a key's first place, its last kept record).

    base: the base case's violations (one scenario, the scan's records of the base state), or None for the base case
    itself (no filtering); params: the increased-violations thresholds (default: powsybl-core's); bus_nominal_kv:
    per-bus nominal voltage in kV (scalar or [n_buses]; None: 1, the voltage thresholds then in per unit).

    OLF's comparison of a voltage-angle violation with the base case's is not restated here (the reference leaves
    it to OLF's own manager), so filtering against a base case refuses voltage-angle records (ValueError) unless
    keep_voltage_angle_violations is set: they are then all kept, which can list more of them than OLF."""
    v = np.asarray(violations, N.VIOLATION_DTYPE)
    if base is not None and not keep_voltage_angle_violations:
        b = np.asarray(base, N.VIOLATION_DTYPE)
        if (v["kind"] >= KIND_HIGH_ANGLE).any() or (b["kind"] >= KIND_HIGH_ANGLE).any():
            raise ValueError("filtering voltage-angle violations against the base case is not modelled; pass "
                             "keep_voltage_angle_violations=True to keep them all")
    params = params or IncreasedViolationsParameters()
    kv = None if bus_nominal_kv is None else np.asarray(bus_nominal_kv, np.float64)

    def to_kv(r):
        if kv is None:
            return np.ones(r.size)
        bus = (r["kind"] == KIND_HIGH_VOLTAGE) | (r["kind"] == KIND_LOW_VOLTAGE)
        bus_kv = kv if kv.ndim == 0 else kv[np.where(bus, r["element"], 0)]
        return np.where(bus, bus_kv, 1.0)

    for r in (v,) if base is None else (v, np.asarray(base, N.VIOLATION_DTYPE)):
        br = r["kind"] == KIND_BRANCH
        if br.any() and (r["element"][br] >= n_table_branches).any():
            raise ValueError("a branch violation's element is beyond the table's branches")
    keys = _keys(v)
    keep = np.ones(v.size, bool)
    if base is not None and v.size:
        b = np.asarray(base, N.VIOLATION_DTYPE)
        if b.size and (b["scenario"] != b["scenario"][0]).any():
            raise ValueError("base violations must come from one state")
        bkeys = _keys(b)
        blast, _ = _last_per_key(bkeys)
        bk = bkeys[blast]
        order = np.argsort(bk)
        bk, brec = bk[order], b[blast][order]
        pos = np.minimum(np.searchsorted(bk, keys), max(bk.size - 1, 0))
        hit = bk[pos] == keys if bk.size else np.zeros(v.size, bool)
        if hit.any():
            idx = np.nonzero(hit)[0]
            ref = brec[pos[idx]]
            keep[idx] = ~weakened_or_equivalent(ref, v[idx], params, to_kv(ref), to_kv(v[idx]))
    kept = np.nonzero(keep)[0]
    if kept.size == 0:
        return v[:0].copy()
    # one map per scenario: the key's first kept place, its last kept record
    skey = v["scenario"][kept].astype(np.int64) * (4 * _CLASS) + keys[kept]
    last, first = _last_per_key(skey)
    place = np.lexsort((kept[first], v["scenario"][kept[first]]))
    return v[kept[last[place]]].copy()
