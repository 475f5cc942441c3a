"""CPU checks of the negotiated-allocator oracle (oracle/negotiate.c).

Pins the restated CS-RR-v1 arithmetic with the reference's zero-capacity
known answer, a hand-derived two-net bottleneck, and invariants that hold for
any board (occupancy and overuse recomputed from the returned routes)."""
import numpy as np
import pytest

from tests.test_oracle_dense import random_lattice


def cell(W, H, x, y, l=0):
    return (l * H + y) * W + x


def test_zero_capacity_known_answer(oracle):
    # Reference known answer: congestion cost 12 = base 5 + initial present 7 *
    # max(0, 0 + 1 - 0); a bound of 12 is accepted, 11 rejected
    # (tests/allocator/cpu_sequential_negotiated_routing_test.cc:787-831).
    cost = np.array([[[1, 5, 1]]], dtype=np.uint32)
    cap = np.zeros(3, dtype=np.uint8)
    nets = [(0, 2, None)]
    pins = np.array([0, 2], dtype=np.uint32)
    r = oracle.negotiate(cost, 0, nets, pins, capacity=cap, initial_present_cost=7,
                         maximum_cell_cost=12, maximum_passes=1)
    assert r["status"] == 0
    assert r["termination"] == 2 and r["passes"] == 1 and r["overuse_per_pass"] == [3]
    assert not r["history"].any()  # no history update after the terminal pass
    r = oracle.negotiate(cost, 0, nets, pins, capacity=cap, initial_present_cost=7,
                         maximum_cell_cost=11, maximum_passes=1)
    assert r["status"] == 1  # the cell of cost 12 exceeds the bound
    r = oracle.negotiate(cost, 0, nets, pins, capacity=cap, initial_present_cost=7,
                         maximum_congestion_cost_value=6, maximum_passes=1)
    assert r["status"] == 1  # congestion 7 exceeds the congestion bound
    # Default prices: the only route never changes, so pass 1 stalls; history
    # is updated once, after the non-terminal pass 0: 10 x overuse 1 per cell.
    r = oracle.negotiate(cost, 0, nets, pins, capacity=cap, maximum_passes=4)
    assert r["status"] == 0 and r["termination"] == 1 and r["passes"] == 2
    assert r["overuse_per_pass"] == [3, 3]
    assert list(r["history"].reshape(-1)) == [10, 10, 10]


def bottleneck_board():
    """7 x 5 single layer, unit costs; a wall at x = 3 with gaps at y = 1 and 3.
    Net A runs (0,1) -> (6,1) straight through gap 1 (cost 6; detour via gap 3
    costs 10).  Net B runs (1,0) -> (5,0) and must pass gap 1 (cost 6) or
    detour via gap 3 (cost 10)."""
    W, H = 7, 5
    cost = np.ones((1, H, W), dtype=np.uint32)
    for y in (0, 2, 4):
        cost[0, y, 3] = 0
    pins = np.array([cell(W, H, 0, 1), cell(W, H, 6, 1), cell(W, H, 1, 0), cell(W, H, 5, 0)],
                    dtype=np.uint32)
    nets = [(0, 2, None), (2, 2, None)]
    return W, H, cost, pins, nets


def test_two_nets_negotiate_a_bottleneck_sequentially(oracle):
    W, H, cost, pins, nets = bottleneck_board()
    r = oracle.negotiate(cost, 0, nets, pins, maximum_passes=8)
    assert r["status"] == 0
    # Pass 0 (no prices): both use gap 1 and share at least (2,1), (3,1), (4,1).
    # Pass 1 (present factor 100): A, rerouted first against B's route, pays
    # 300 through gap 1 and detours (cost 10); B then keeps gap 1.
    assert r["termination"] == 0 and r["passes"] == 2
    assert r["overuse_per_pass"][0] >= 3 and r["overuse_per_pass"][1] == 0
    a, b = r["routes"]
    gap1, gap3 = cell(W, H, 3, 1), cell(W, H, 3, 3)
    assert gap3 in a and gap1 not in a and len(a) == 11
    assert gap1 in b and len(b) == 7
    hist = r["history"].reshape(-1)
    shared = {cell(W, H, x, 1) for x in (2, 3, 4)}
    assert all(hist[c] == 10 for c in shared)
    assert set(np.flatnonzero(hist)) <= shared | {cell(W, H, 1, 1), cell(W, H, 5, 1)}
    assert np.array_equal(r["occupancy"].reshape(-1) > 0,
                          np.isin(np.arange(W * H), np.concatenate([a, b])))


def test_batched_snapshot_moves_both_nets_together(oracle):
    W, H, cost, pins, nets = bottleneck_board()
    r = oracle.negotiate(cost, 0, nets, pins, maximum_passes=2, batch_nets=2)
    # Both nets see the same snapshot in pass 1: no present cost (each sees
    # the other's route ripped up), history 10 on each shared gap-1 cell, so
    # both detour through gap 3 together.
    a, b = r["routes"]
    gap3 = cell(W, H, 3, 3)
    assert gap3 in a and gap3 in b
    assert r["termination"] == 2 and r["overuse_per_pass"][1] >= 3


@pytest.mark.parametrize("seed", range(4))
def test_occupancy_and_overuse_follow_from_routes(oracle, seed):
    rng = np.random.default_rng(300 + seed)
    L, H, W = 2, 16, 16
    cost = random_lattice(rng, L, H, W, blocked=0.1, cmax=6)
    free = np.flatnonzero(cost.reshape(-1))
    nets, flat = [], []
    for i in range(8):
        k = int(rng.integers(2, 4))
        p = rng.choice(free, size=k, replace=False)
        p = [int(v) for v in p if int(v) not in flat][:k]
        if len(p) < 2:
            continue
        nets.append((len(flat), len(p), None))
        flat.extend(p)
    pins = np.array(flat, dtype=np.uint32)
    r = oracle.negotiate(cost, 5, nets, pins, maximum_passes=6, batch_nets=[1, 3, 8, 2][seed])
    assert r["status"] == 0
    occ = np.zeros(L * H * W, dtype=np.int64)
    for t in r["routes"]:
        if t is not None:
            assert len(set(t.tolist())) == len(t)  # tree cells are distinct
            occ[t] += 1
    assert np.array_equal(occ, r["occupancy"].reshape(-1))
    assert int(np.maximum(occ - 1, 0).sum()) == r["overuse"]
    assert r["overuse_per_pass"][-1] == r["overuse"]
    hist = r["history"].reshape(-1)
    assert (hist % 10 == 0).all()
    # Every non-terminal pass adds 10 x its overuse, summed over cells.
    assert int(hist.sum()) == 10 * sum(r["overuse_per_pass"][:-1])
    if r["termination"] == 0:
        assert r["overuse"] == 0 and r["routed"] == len(nets)


@pytest.mark.parametrize("seed", range(3))
def test_zero_present_factor_pass_is_order_free(oracle, seed):
    """A pass with present factor 0 prices cells by base + history only, so its
    routes do not depend on the order or batching of its nets: one pass in
    batches of 1 (the reference's order) equals one batch of every net.  (The
    library routes such passes as one batch.)"""
    rng = np.random.default_rng(500 + seed)
    cost = random_lattice(rng, 2, 24, 30, blocked=0.15, cmax=6)
    free = np.flatnonzero(cost.reshape(-1))
    pins = rng.choice(free, size=24, replace=False).astype(np.uint32)
    nets = [(2 * i, 2, None) for i in range(12)]
    history = rng.integers(0, 40, cost.size).astype(np.uint32)
    a = oracle.negotiate(cost, 5, nets, pins, maximum_passes=1, batch_nets=1, initial_history=history)
    b = oracle.negotiate(cost, 5, nets, pins, maximum_passes=1, batch_nets=len(nets), initial_history=history)
    assert a["status"] == b["status"] == 0
    assert a["overuse_per_pass"] == b["overuse_per_pass"]
    assert [list(x) if x is not None else None for x in a["routes"]] == \
           [list(x) if x is not None else None for x in b["routes"]]
    # With a nonzero present factor the order matters: nets routed one at a
    # time see each other's occupancy and avoid it; one batch does not.
    kw = dict(maximum_passes=1, initial_history=history, initial_present_cost=50)
    c = oracle.negotiate(cost, 5, nets, pins, batch_nets=1, **kw)
    d = oracle.negotiate(cost, 5, nets, pins, batch_nets=len(nets), **kw)
    assert c["overuse_per_pass"][0] < d["overuse_per_pass"][0]


@pytest.mark.parametrize("seed", range(3))
def test_replayed_pass_equals_the_oracle_pass(oracle, seed):
    """replay_exact_pass (one exact-order pass from a given state, used at
    full size where the whole oracle run is too slow) started from the state
    after pass 0 gives the trees of pass 1 of a two-pass oracle run, net for
    net, with windows; stopped early, the trees it routed are the first ones."""
    rng = np.random.default_rng(700 + seed)
    L, H, W = 2, 20, 24
    cost = random_lattice(rng, L, H, W, blocked=0.1, cmax=6)
    free = np.flatnonzero(cost.reshape(-1))
    pins = rng.choice(free, size=48, replace=False).astype(np.uint32)
    nets = []
    for i in range(16):
        p = pins[3 * i:3 * i + 3].astype(np.int64)
        xs, ys = p % W, (p // W) % H
        x0, y0 = max(int(xs.min()) - 2, 0), max(int(ys.min()) - 2, 0)
        nets.append((3 * i, 3, (x0, y0, min(int(xs.max()) + 3, W) - x0, min(int(ys.max()) + 3, H) - y0)))
    a = oracle.negotiate(cost, 5, nets, pins, maximum_passes=1)  # state after pass 0
    b = oracle.negotiate(cost, 5, nets, pins, maximum_passes=2)
    assert a["status"] == b["status"] == 0
    assert a["overuse_per_pass"][0] > 0 and b["passes"] == 2
    hist = 10 * np.maximum(a["occupancy"].astype(np.int64) - 1, 0)
    assert np.array_equal(hist, b["history"])  # the update after the non-terminal pass 0
    trees, _ = oracle.replay_exact_pass(cost, 5, nets, pins, a["routes"], a["occupancy"], hist, 1)
    assert len(trees) == len(nets)
    for x, y in zip(trees, b["routes"]):
        assert (x is None) == (y is None)
        if x is not None:
            assert np.array_equal(x, y)
    head, _ = oracle.replay_exact_pass(cost, 5, nets, pins, a["routes"], a["occupancy"], hist, 1, first=5)
    assert len(head) == 5 and all(np.array_equal(x, y) for x, y in zip(head, trees))
    # Planted mistakes change the trees: history left out, incumbents not
    # ripped up.
    def differs(other):
        return any((x is None) != (y is None) or (x is not None and not np.array_equal(x, y))
                   for x, y in zip(other, trees))

    assert differs(oracle.replay_exact_pass(cost, 5, nets, pins, a["routes"], a["occupancy"], 0 * hist, 1)[0])
    assert differs(oracle.replay_exact_pass(cost, 5, nets, pins, [None] * len(nets), a["occupancy"], hist, 1)[0])


def test_replayed_pass_with_capacity_plane_and_seeded_history(oracle):
    """The replay with a capacity plane and a seeded history equals pass 1 of
    the oracle run with the same planes."""
    rng = np.random.default_rng(731)
    L, H, W = 2, 20, 24
    cost = random_lattice(rng, L, H, W, blocked=0.1, cmax=6)
    free = np.flatnonzero(cost.reshape(-1))
    pins = rng.choice(free, size=36, replace=False).astype(np.uint32)
    nets = [(2 * i, 2, None) for i in range(18)]
    cap = rng.integers(0, 3, cost.size).astype(np.uint8).reshape(cost.shape)
    cap.reshape(-1)[pins] = 1  # pins keep room for their own net
    seed = rng.integers(0, 30, cost.size).astype(np.uint32)
    kw = dict(capacity=cap, initial_history=seed)
    a = oracle.negotiate(cost, 5, nets, pins, maximum_passes=1, **kw)
    b = oracle.negotiate(cost, 5, nets, pins, maximum_passes=2, **kw)
    assert a["status"] == b["status"] == 0 and b["passes"] == 2 and a["overuse_per_pass"][0] > 0
    over = np.maximum(a["occupancy"].astype(np.int64) - cap.astype(np.int64), 0)
    hist = seed.reshape(cost.shape).astype(np.int64) + 10 * over
    assert np.array_equal(hist, b["history"])
    trees, _ = oracle.replay_exact_pass(cost, 5, nets, pins, a["routes"], a["occupancy"], hist, 1, capacity=cap)
    for x, y in zip(trees, b["routes"]):
        assert (x is None) == (y is None)
        if x is not None:
            assert np.array_equal(x, y)
    # Capacity 1 everywhere instead of the plane gives other trees.
    other, _ = oracle.replay_exact_pass(cost, 5, nets, pins, a["routes"], a["occupancy"], hist, 1)
    assert any((x is None) != (y is None) or (x is not None and not np.array_equal(x, y))
               for x, y in zip(other, trees))
