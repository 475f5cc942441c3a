"""CPU oracle of the candidate-allocator price epoch (test infrastructure only).

Only tests/ use this module, to check the library.  It restates, loop for
loop, the price update of PlanNegotiatedRegeneration
(src/allocator/negotiated_regeneration_plan.cc:697-790) and NextPresentFactor
(:430-442) on cell resources: the resource map is a dict keyed by cell index
(the canonical resource order), Python integers with explicit uint64 overflow
checks.  Pinned by the reference's known answers
(tests/allocator/negotiated_regeneration_plan_test.cc:815-872: epoch 0 prices
100 / 10 / 110, epoch 1 prices 200 / 20 / 220, historical prices persisting as
0 / 10 / 10 once overuse clears) and hand cases for every bound's order.
"""
from __future__ import annotations

U64 = (1 << 64) - 1

OK, FACTOR_OVERFLOW, HOT_BOUND, ENTRY_BOUND, MUL_OVERFLOW, HIST_OVERFLOW, TOTAL_OVERFLOW, VALUE_BOUND, \
    AGG_OVERFLOW, AGG_BOUND = range(10)

DEFAULT_POLICY = dict(initial_present_factor=100, present_factor_increment=100, historical_price_increment=10)
DEFAULT_LIMITS = dict(maximum_price_entries=1_000_000, maximum_hot_resources=1_000_000,
                      maximum_price_value=10_000_000, maximum_aggregate_price=1_000_000_000_000)


def next_present_factor(epoch, policy):
    steps = epoch * policy["present_factor_increment"]
    if steps > U64:
        return None
    f = policy["initial_present_factor"] + steps
    return None if f > U64 else f


def price_epoch(epoch, usage, capacity, history, policy=None, limits=None):
    """usage, capacity, history: sequences over cells.  Returns (outcome,
    cell or None, result dict); result has present / history / total lists
    and the epoch summary when the outcome is OK."""
    policy = dict(DEFAULT_POLICY, **(policy or {}))
    limits = dict(DEFAULT_LIMITS, **(limits or {}))
    factor = next_present_factor(epoch, policy)
    if factor is None:
        return FACTOR_OVERFLOW, None, None
    n = len(usage)
    over = [max(0, int(usage[c]) - int(capacity[c])) for c in range(n)]
    hot = sum(1 for v in over if v)
    if hot > limits["maximum_hot_resources"]:
        return HOT_BOUND, None, None
    # prior entries with non-zero history: present 0, total = history
    prices = {c: [0, int(history[c]), int(history[c])] for c in range(n) if history[c] != 0}
    for c in range(n):  # overused resources in canonical order
        if not over[c]:
            continue
        if c not in prices and len(prices) >= limits["maximum_price_entries"]:
            return ENTRY_BOUND, c, None
        present = factor * over[c]
        inc = policy["historical_price_increment"] * over[c]
        if present > U64 or inc > U64:
            return MUL_OVERFLOW, c, None
        p = prices.setdefault(c, [0, 0, 0])
        p[0] = present
        if p[1] + inc > U64:
            return HIST_OVERFLOW, c, None
        p[1] += inc
        if p[0] + p[1] > U64:
            return TOTAL_OVERFLOW, c, None
        p[2] = p[0] + p[1]
    aggregate = 0
    for c in sorted(prices):
        present, hist, total = prices[c]
        mv = limits["maximum_price_value"]
        if present > mv or hist > mv or total > mv:
            return VALUE_BOUND, c, None
        if aggregate + total > U64:
            return AGG_OVERFLOW, c, None
        aggregate += total
        if aggregate > limits["maximum_aggregate_price"]:
            return AGG_BOUND, c, None
    out_present = [0] * n
    out_hist = [int(h) for h in history]
    out_total = [0] * n
    for c, (p, h, t) in prices.items():
        out_present[c], out_hist[c], out_total[c] = p, h, t
    return OK, None, {"present_factor": factor, "hot_resources": hot, "price_entries": len(prices),
                      "aggregate_price": aggregate, "present": out_present, "history": out_hist,
                      "total": out_total}
