"""GPU tests of the bank refill through the C ABI (olfhip_begin_batch_install, olfhip_solve_inputs.pause_at_ended,
olfhip_read_handed_back_slots; the reference's BatchSolveNative.beginBatchInstall / pauseAtEnded /
readBatchHandedBackSlots, driven as GpuBatchChunkSolver.java:545-610,866-915 drives them):

* a queue of contingencies run through a small bank, each solve pausing once a few scenarios have ended and the next
  contingencies installed into the slots it handed back while the others go on solving, gives every contingency
  bitwise the answer of one solve of the whole set (states, statuses, iteration counts), in FP64 and FP32-refined;
* the rows of a solve are the scenarios it handed back, in install order, with their slots;
* a slot chosen twice, beyond the bank, or still holding a scenario not handed back is refused."""
import numpy as np
import pytest

import pyoracle as po
from olfgpu import _native as N
from olfgpu import gridgen

pytestmark = pytest.mark.gpu

BANK = 16


@pytest.fixture(scope="module")
def g2000(grid):
    if N.device_count() == 0:
        pytest.fail("no GPU visible: the gpu-marked tests need an MI355X")
    net = grid("grid2000")
    osys = po.OracleSystem(net)
    x0, tg, md, _ = net.system()
    xb, mb, tb, ltb, res = osys.solve(x0, md, tg)
    assert po.STATUS[res.status] == "CONVERGED"
    return net, N.Structure.from_network(net), (xb, mb, tb, ltb)


def inputs(base, ctg, pause=0):
    xb, mb, tb, ltb = base
    return N.BatchContext.make_inputs(len(ctg), xb, tb, mb, outaged_branches=ctg, reactive_limits_seed_state=ltb,
                                      pause_at_ended=pause)


def run_bank(ctx, base, ctg, refill_at):
    """The reference's refill loop: the free slots take the next contingencies; a solve with contingencies still
    queued returns once refill_at (or fewer, the queue's rest) of its scenarios have ended."""
    m = len(ctg)
    free = list(range(BANK))
    occupant = {}
    out = {}
    calls, paused_with_running = 0, 0
    start = 0
    while len(out) < m:
        take = min(len(free), m - start)
        slots = [free.pop(0) for _ in range(take)]
        ctx.begin_install(None if take == BANK else slots)
        for j, s in enumerate(slots):
            occupant[s] = start + j
        queued = m - start - take
        pause = min(queued, refill_at) if queued > 0 else 0
        inp, keep = inputs(base, ctg[start:start + take], pause)
        r = ctx.solve(inp)
        calls += 1
        start += take
        handed = r["slots"]
        assert len(handed) == len(r["status"])
        # rows in install order
        assert [occupant[s] for s in handed] == sorted(occupant[s] for s in handed)
        assert np.array_equal(handed, ctx.read_handed_back_slots())
        if pause:
            assert len(handed) >= pause
        running = len(occupant) - len(handed)
        paused_with_running += pause > 0 and running > 0
        for j, s in enumerate(handed):
            i = occupant.pop(s)
            out[i] = (r["x"][j].copy(), int(r["status"][j]), int(r["total_iterations"][j]), int(r["iterations"][j]))
            free.append(s)
        free.sort()
        if not pause:
            assert not occupant, "a solve without a pause hands back every scenario"
    return out, calls, paused_with_running


@pytest.mark.parametrize("mode", [N.LU_FP64, N.LU_FP32_REFINED])
def test_refilled_bank_equals_one_solve(g2000, mode):
    net, st, base = g2000
    print(f"[bank] grid2000 64 contingencies through a bank of {BANK}, mode {mode}", flush=True)
    ctg = gridgen.sample_contingencies(net, 64, 0.1, 11)
    one = N.BatchContext(st, len(ctg), mode, net.reactive_limits())
    inp, keep = inputs(base, ctg)
    ref = one.solve(inp)
    assert (ref["status"] == N.CONVERGED).sum() >= 50
    bank = N.BatchContext(st, BANK, mode, net.reactive_limits())
    out, calls, paused = run_bank(bank, base, ctg, refill_at=3)
    assert calls >= 4 and paused >= 1
    for i in range(len(ctg)):
        x, status, total, last = out[i]
        assert status == ref["status"][i], i
        assert total == ref["total_iterations"][i] and last == ref["iterations"][i], i
        assert np.array_equal(x, ref["x"][i]), i
    # the same bank, driven again, gives the same answers (nothing left behind by the first run)
    out2, _, _ = run_bank(bank, base, ctg, refill_at=5)
    for i in range(len(ctg)):
        assert np.array_equal(out2[i][0], out[i][0]) and out2[i][1:] == out[i][1:], i


def test_bank_slot_choice_is_checked(g2000):
    net, st, base = g2000
    print("[bank] slot checks", flush=True)
    ctx = N.BatchContext(st, BANK, N.LU_FP64, net.reactive_limits())
    for bad in ([0, 0], [BANK], [-1]):
        with pytest.raises(N.OlfHipError):
            ctx.begin_install(bad)
    ctg = gridgen.sample_contingencies(net, 48, 0.1, 12)
    # a bank whose scenarios end at different steps: the fastest first, the slowest among them
    probe = N.BatchContext(st, len(ctg), N.LU_FP64, net.reactive_limits())
    inp, keep = inputs(base, ctg)
    ref = probe.solve(inp)
    tot = ref["total_iterations"]
    order = np.argsort(tot, kind="stable")
    assert tot[order[-1]] - tot[order[0]] >= 3
    first = np.concatenate([order[:BANK - 1], order[-1:]])
    inp, keep = inputs(base, ctg[first], pause=1)
    r = ctx.solve(inp)
    held = sorted(set(range(BANK)) - set(r["slots"].tolist()))
    assert held, "the pause left scenarios solving"
    with pytest.raises(N.OlfHipError):
        ctx.begin_install([held[0]])          # still holds a scenario not handed back
    # a solve whose count differs from the chosen slots is refused, and the choice is spent
    ctx.begin_install(r["slots"][:2])
    inp, keep = inputs(base, ctg[order[BANK:BANK + 1]])
    with pytest.raises(N.OlfHipError):
        ctx.solve(inp)
    # the rest goes on: the remaining scenarios end at the next call, none lost
    ctx.begin_install(r["slots"][:1])
    last = ctx.solve(inp)
    assert sorted(last["slots"].tolist()) == sorted(held + [int(r["slots"][0])])
    holder = {j: int(first[j]) for j in held}
    holder[int(r["slots"][0])] = int(order[BANK])
    for row, slot in enumerate(last["slots"]):
        i = holder[int(slot)]
        assert last["status"][row] == ref["status"][i] and last["total_iterations"][row] == tot[i]
        assert np.array_equal(last["x"][row], ref["x"][i])


def _by_scenario(v, scenarios):
    """violation records of the given scenarios, as comparable tuples"""
    return sorted((int(r["scenario"]), int(r["kind"]), int(r["element"]), int(r["side"]), int(r["limit"]),
                   float(r["value"])) for r in v if int(r["scenario"]) in scenarios)


def test_readbacks_after_a_paused_solve_are_by_slot(g2000):
    """After a bank or paused solve the readbacks index scenarios by slot: the violation detection, both screens and
    the near-threshold flags report exactly the slots the solve handed back, with what one solve of the same
    contingencies reports for them; a slot still solving, or one handed back by an earlier call, reports nothing."""
    net, st, base = g2000
    xb = base[0]
    print("[bank] readbacks by slot", flush=True)
    ctg = gridgen.sample_contingencies(net, 24, 0.1, 13)
    lt = N.LimitTable(st, gridgen.limit_table(net, xb))
    one = N.BatchContext(st, 24, N.LU_FP64, net.reactive_limits())
    inp, keep = inputs(base, ctg)
    ref = one.solve(inp)
    ref_v = lt.detect(one)
    ref_counts, ref_rec = lt.screen(one, 24)
    ref_amb = lt.ambiguous(one)
    bank = N.BatchContext(st, BANK, N.LU_FP64, net.reactive_limits())
    inp, keep = inputs(base, ctg[:BANK], pause=3)
    r = bank.solve(inp)
    holder = {j: j for j in range(BANK)}   # the default install: contingency j in slot j

    def check(handed):
        v = lt.detect(bank)
        assert {int(x["scenario"]) for x in v} <= set(handed)
        got = [(holder[s], *rest) for s, *rest in _by_scenario(v, set(handed))]
        want = _by_scenario(ref_v, {holder[s] for s in handed})
        assert sorted(got) == want
        for screen in (lt.screen, lt.screen_no_alloc):
            counts, rec = screen(bank, BANK)
            for s in range(BANK):
                if s in handed:
                    i = holder[s]
                    assert counts[s] == ref_counts[i]
                    n = min(counts[s], rec.shape[1])
                    assert sorted(rec[s, :n]) == sorted(ref_rec[i, :n])
                else:
                    assert counts[s] == 0, s
        amb = lt.ambiguous(bank)
        for s in range(BANK):
            assert amb[s] == (s in handed and ref_amb[holder[s]]), s

    first = r["slots"].tolist()
    assert len(first) >= 3
    check(first)
    assert np.array_equal(r["status"], ref["status"][[holder[s] for s in first]])
    # refill two of the handed-back slots with the next contingencies and run the bank to its end
    refill = first[:2]
    bank.begin_install(refill)
    for j, s in enumerate(refill):
        holder[s] = BANK + j
    inp, keep = inputs(base, ctg[BANK:BANK + 2])
    r = bank.solve(inp)
    second = r["slots"].tolist()
    assert set(second) == (set(range(BANK)) - set(first)) | set(refill)
    check(second)
