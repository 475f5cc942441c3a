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


# --------------------------------------------------------------------------
# Price snapshots on cell resources.  The entry map is the nonzero cells of
# the present and history planes in cell order; the identity follows
# SnapshotIdentity (src/allocator/negotiated_regeneration_plan.cc:108-127)
# field for field with the cell resource (layer, x, y) in place of the edge
# resource (layer, x, y, direction), under its own domain string, and the
# lattice shape after the compiler version.  PricePolicyIdentity (:98-106) is
# the reference's exactly.
# --------------------------------------------------------------------------
class _Fnv:
    def __init__(self):
        self.v = 14695981039346656037

    def byte(self, b):
        self.v = ((self.v ^ b) * 1099511628211) & U64

    def u32(self, x):
        for i in range(4):
            self.byte((x >> (8 * i)) & 0xFF)

    def u64(self, x):
        x &= U64
        for i in range(8):
            self.byte((x >> (8 * i)) & 0xFF)

    def string(self, s):
        self.u64(len(s))
        for c in s.encode():
            self.byte(c)

    def nonzero(self):
        return self.v or 1


def policy_identity(policy=None):
    pol = dict(DEFAULT_POLICY, **(policy or {}))
    h = _Fnv()
    h.string("APGAR-P4R06-NEGOTIATED-PRICE-POLICY-V1")
    h.u64(pol["initial_present_factor"])
    h.u64(pol["present_factor_increment"])
    h.u64(pol["historical_price_increment"])
    return h.nonzero()


def snapshot_entries(present, history):
    """[(cell, present, history, total)] of the nonzero cells, in cell order."""
    import numpy as np
    p = np.asarray(present, dtype=np.uint64).reshape(-1)
    hh = np.asarray(history, dtype=np.uint64).reshape(-1)
    idx = np.flatnonzero((p | hh) != 0)
    return [(int(c), int(p[c]), int(hh[c]), int(p[c]) + int(hh[c])) for c in idx]


def snapshot_identity(assoc, capacity_units, policy_ident, prior_identity, epoch, present_factor, entries):
    """assoc: dict(board_content_hash, compiler_profile_fingerprint,
    geometry_compiler_version, width, height, layers)."""
    h = _Fnv()
    h.string("APGX-CELL-PRICE-SNAPSHOT-V1")
    h.u64(assoc["board_content_hash"])
    h.u64(assoc["compiler_profile_fingerprint"])
    h.u32(assoc["geometry_compiler_version"])
    h.u32(assoc["width"]); h.u32(assoc["height"]); h.u32(assoc["layers"])
    h.u32(capacity_units)
    h.u64(policy_ident)
    h.u64(prior_identity)
    h.u64(epoch)
    h.u64(present_factor)
    h.u64(len(entries))
    plane = assoc["width"] * assoc["height"]
    for cell, pres, hist, tot in entries:
        h.u32(cell // plane)
        h.u64(cell % assoc["width"])          # x as i64 (non-negative)
        h.u64((cell % plane) // assoc["width"])
        h.u64(pres); h.u64(hist); h.u64(tot)
    return h.nonzero()
