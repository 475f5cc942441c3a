"""CPU checks that pin the price-epoch oracle: the reference's known answers
and the order of every bound on hand cases."""
from oracle import prices as op


def test_known_answers_two_epochs_and_cold_epoch():
    # One overused unit on each of three cells (capacity 1, usage 2), as the
    # reference's microcase: epoch 0 -> 100 / 10 / 110, epoch 1 -> 200 / 20 / 220.
    usage, cap = [2, 1, 2, 0, 2], [1, 1, 1, 1, 1]
    oc, _, r0 = op.price_epoch(0, usage, cap, [0] * 5)
    assert oc == op.OK and r0["present_factor"] == 100
    assert [r0["present"][c] for c in (0, 2, 4)] == [100] * 3
    assert [r0["history"][c] for c in (0, 2, 4)] == [10] * 3
    assert [r0["total"][c] for c in (0, 2, 4)] == [110] * 3
    assert r0["price_entries"] == 3 and r0["aggregate_price"] == 330
    oc, _, r1 = op.price_epoch(1, usage, cap, r0["history"])
    assert oc == op.OK and r1["present_factor"] == 200
    assert [(r1["present"][c], r1["history"][c], r1["total"][c]) for c in (0, 2, 4)] == [(200, 20, 220)] * 3
    # overuse gone: history persists, present resets
    oc, _, cold = op.price_epoch(1, [1] * 5, cap, r0["history"])
    assert oc == op.OK and cold["hot_resources"] == 0
    assert [(cold["present"][c], cold["history"][c], cold["total"][c]) for c in (0, 2, 4)] == [(0, 10, 10)] * 3


def test_bound_order():
    usage, cap = [3, 0, 2, 2], [1, 1, 1, 1]
    hist = [0, 5, 0, 0]
    # hot bound is checked before anything else
    assert op.price_epoch(0, usage, cap, hist, limits=dict(maximum_hot_resources=2))[0] == op.HOT_BOUND
    # entry bound: prior entry {1}; new entries 0, 2, 3; bound 3 trips at the
    # third new entry (cell 3), bound 1 at the first (cell 0)
    assert op.price_epoch(0, usage, cap, hist, limits=dict(maximum_price_entries=3))[:2] == (op.ENTRY_BOUND, 3)
    assert op.price_epoch(0, usage, cap, hist, limits=dict(maximum_price_entries=1))[:2] == (op.ENTRY_BOUND, 0)
    assert op.price_epoch(0, usage, cap, hist, limits=dict(maximum_price_entries=4))[0] == op.OK
    # value bound in cell order: cell 0 has present 200 (overuse 2)
    assert op.price_epoch(0, usage, cap, hist, limits=dict(maximum_price_value=150))[:2] == (op.VALUE_BOUND, 0)
    # aggregate: totals 220, 5, 110, 110 -> running 220, 225, 335, 445
    assert op.price_epoch(0, usage, cap, hist, limits=dict(maximum_aggregate_price=334))[:2] == (op.AGG_BOUND, 2)
    assert op.price_epoch(0, usage, cap, hist, limits=dict(maximum_aggregate_price=445))[0] == op.OK
    # arithmetic: factor overflow, multiply, history and total overflows
    big = dict(initial_present_factor=1 << 63, present_factor_increment=1 << 63)
    assert op.price_epoch(2, usage, cap, hist, policy=big)[0] == op.FACTOR_OVERFLOW
    assert op.price_epoch(0, usage, cap, hist, policy=dict(initial_present_factor=1 << 63))[:2] == \
        (op.MUL_OVERFLOW, 0)
    assert op.price_epoch(0, [2], [1], [op.U64 - 5], policy=dict(historical_price_increment=10))[0] == \
        op.HIST_OVERFLOW
    assert op.price_epoch(0, [2], [1], [op.U64 - 20], policy=dict(historical_price_increment=10))[0] == \
        op.TOTAL_OVERFLOW
