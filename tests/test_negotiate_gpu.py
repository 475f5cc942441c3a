"""GPU parity of the negotiated-congestion allocator against the CPU oracle.

Bar: bit-exact (termination, passes, overuse per pass, history and usage
planes, every net's tree cells) on scaled boards of the six_layer_1536 family
and on the hand-derived bottleneck board; at full size, legality and the
accounting identities for every net."""
import numpy as np
import pytest

from tests.lattice_cases import SIX_LAYER_1536, make_board
from tests.test_oracle_negotiate import bottleneck_board

pytestmark = pytest.mark.gpu

SMALL = dict(width=96, height=96, layers=3, nets=30, bga_fields=3, bga_pads=8, bga_pitch=3,
             scatter_radius=40, window_margin=6, keepout_max_side=16)


def _say(msg):
    print(msg, flush=True)


def oracle_nets(nets):
    return [(n.pin_offset, n.pin_count, (n.window.x0, n.window.y0, n.window.width, n.window.height))
            for n in nets]


def assert_same(gpu, orc):
    assert gpu["termination"] == orc["termination"]
    assert gpu["passes"] == orc["passes"]
    assert gpu["overuse_per_pass"] == orc["overuse_per_pass"]
    assert gpu["routed"] == orc["routed"]
    assert np.array_equal(gpu["history"], orc["history"])
    assert np.array_equal(gpu["occupancy"], orc["occupancy"])
    for a, b in zip(gpu["routes"], orc["routes"]):
        assert (a is None) == (b is None)
        if a is not None:
            assert np.array_equal(a, b)


@pytest.mark.parametrize("seed,batch,group", [(0, 1, 0), (1, 4, 0), (2, 16, 0), (3, 30, 0), (4, 0, 0), (5, 0, 0),
                                              (4, 0, 32), (5, 0, 32)])
def test_small_boards_match_oracle(apgx, oracle, gpu, seed, batch, group):
    """batch 0 must equal the oracle's one-net-at-a-time order, both as
    window-disjoint levels (group 0) and as speculative groups."""
    _say(f"[case] allocator parity seed={seed} batch={batch} group={group}")
    cost, nets, pins = make_board(apgx, SIX_LAYER_1536, 1000 + seed, **SMALL)
    via = SIX_LAYER_1536["via"]
    lat = apgx.Lattice(cost, via_cost=via, cost_bytes=2)
    g = lat.negotiate(nets, pins, maximum_passes=12, batch_nets=batch, group_nets=group)
    o = oracle.negotiate(cost, via, oracle_nets(nets), pins, maximum_passes=12, batch_nets=max(batch, 1))
    if batch == 0:
        assert g["batches"] < 12 * len(nets)  # levels group nets
    assert o["status"] == 0
    assert g["overuse_per_pass"][0] > 0  # the board is congested to begin with
    assert_same(g, o)
    lat.close()


@pytest.mark.parametrize("seed", [7, 8])
def test_high_present_costs_match_oracle(apgx, oracle, gpu, seed):
    """Present factors of 6,000 + 3,000 per pass: snapshot costs cross the
    register tiles' narrow limit (4,095, costs kept pre-shifted) into the wide
    range (up to 32,767, shifted at each relaxation) and, where a cell is
    overused twice, beyond it (LDS tiles); every pass bit-exact."""
    _say(f"[case] allocator parity under high present costs seed={seed}")
    cost, nets, pins = make_board(apgx, SIX_LAYER_1536, 1000 + seed, **SMALL)
    via = SIX_LAYER_1536["via"]
    lat = apgx.Lattice(cost, via_cost=via, cost_bytes=2)
    kw = dict(maximum_passes=6, initial_present_cost=6000, present_cost_increment=3000)
    g = lat.negotiate(nets, pins, batch_nets=0, **kw)
    o = oracle.negotiate(cost, via, oracle_nets(nets), pins, batch_nets=1, **kw)
    assert o["status"] == 0
    assert_same(g, o)
    lat.close()


REDONE = []


@pytest.mark.parametrize("seed,group", [(10, 1), (11, 4), (12, 8), (13, 30), (14, 16)])
def test_speculative_groups_match_oracle(apgx, oracle, gpu, seed, group):
    """Speculative groups of 1 to 30 nets (a group searches its nets at once,
    each on its own snapshot, commits them in order and searches a net again
    on the exact snapshot when its tree meets a cell an earlier member
    installed): every pass bit-exact with the oracle's one-net order."""
    _say(f"[case] allocator speculative groups seed={seed} group={group}")
    cost, nets, pins = make_board(apgx, SIX_LAYER_1536, 1000 + seed, **SMALL)
    via = SIX_LAYER_1536["via"]
    lat = apgx.Lattice(cost, via_cost=via, cost_bytes=2)
    g = lat.negotiate(nets, pins, maximum_passes=12, batch_nets=0, group_nets=group)
    o = oracle.negotiate(cost, via, oracle_nets(nets), pins, maximum_passes=12, batch_nets=1)
    assert o["status"] == 0
    assert_same(g, o)
    # Every pass after the first (zero present factor: one batch) runs in groups.
    assert g["group_routed_nets"] == (g["passes"] - 1) * len(nets)
    if group == 1:
        assert g["redone_nets"] == 0  # a group of one is never speculative
    REDONE.append(g["redone_nets"])
    _say(f"[info] passes={g['passes']} redone={g['redone_nets']} of {g['group_routed_nets']}")
    lat.close()


def test_speculation_was_exercised(apgx, gpu):
    """The parametrised cases above must have searched some nets again."""
    if len(REDONE) < 5:
        pytest.skip("runs after the speculative group cases")
    assert sum(REDONE) > 0


def test_seeded_history_and_capacity_match_oracle(apgx, oracle, gpu):
    _say("[case] allocator parity with seeded history and capacity plane")
    cost, nets, pins = make_board(apgx, SIX_LAYER_1536, 1010, **SMALL)
    rng = np.random.default_rng(4)
    hist0 = (rng.integers(0, 6, size=cost.shape) * 10).astype(np.uint32)
    cap = np.where(rng.random(cost.shape) < 0.05, 2, 1).astype(np.uint8)
    via = SIX_LAYER_1536["via"]
    lat = apgx.Lattice(cost, via_cost=via, cost_bytes=2)
    g = lat.negotiate(nets, pins, capacity=cap, initial_history=hist0, maximum_passes=10, batch_nets=8,
                      present_cost_increment=50, historical_cost_increment=20)
    o = oracle.negotiate(cost, via, oracle_nets(nets), pins, capacity=cap, initial_history=hist0,
                         maximum_passes=10, batch_nets=8, present_cost_increment=50,
                         historical_cost_increment=20)
    assert_same(g, o)
    lat.close()


def test_seeded_history_and_capacity_groups_match_oracle(apgx, oracle, gpu):
    _say("[case] allocator speculative groups with seeded history and capacity plane")
    cost, nets, pins = make_board(apgx, SIX_LAYER_1536, 1011, **SMALL)
    rng = np.random.default_rng(5)
    hist0 = (rng.integers(0, 6, size=cost.shape) * 10).astype(np.uint32)
    cap = np.where(rng.random(cost.shape) < 0.05, 2, 1).astype(np.uint8)
    via = SIX_LAYER_1536["via"]
    lat = apgx.Lattice(cost, via_cost=via, cost_bytes=2)
    kw = dict(maximum_passes=10, present_cost_increment=50, historical_cost_increment=20)
    g = lat.negotiate(nets, pins, capacity=cap, initial_history=hist0, batch_nets=0, group_nets=12, **kw)
    o = oracle.negotiate(cost, via, oracle_nets(nets), pins, capacity=cap, initial_history=hist0,
                         batch_nets=1, **kw)
    assert o["status"] == 0
    assert_same(g, o)
    lat.close()


def test_bottleneck_board(apgx, oracle, gpu):
    _say("[case] allocator bottleneck board")
    W, H, cost, pins, onets = bottleneck_board()
    nets, p = apgx.nets_from_lists([pins[:2], pins[2:]])
    lat = apgx.Lattice(cost, via_cost=0)
    for batch in (1, 2):
        g = lat.negotiate(nets, p, maximum_passes=8 if batch == 1 else 2, batch_nets=batch)
        o = oracle.negotiate(cost, 0, onets, pins, maximum_passes=8 if batch == 1 else 2, batch_nets=batch)
        assert_same(g, o)
    lat.close()


def test_pass_callback_reports_every_pass(apgx, gpu):
    _say("[case] allocator per-pass callback")
    cost, nets, pins = make_board(apgx, SIX_LAYER_1536, 1003, **SMALL)
    lat = apgx.Lattice(cost, via_cost=SIX_LAYER_1536["via"], cost_bytes=2)
    seen = []
    g = lat.negotiate(nets, pins, maximum_passes=6, batch_nets=0,
                      on_pass=lambda p, o, r, ms: seen.append((p, o, r, ms)))
    assert [p for p, _, _, _ in seen] == list(range(g["passes"]))
    assert [o for _, o, _, _ in seen] == g["overuse_per_pass"]
    assert seen[-1][2] == g["routed"]
    assert all(ms > 0 for *_, ms in seen) and seen[-1][3] <= g["wall_ms"]
    lat.close()


def test_value_bound_is_reported(apgx, gpu):
    cost = np.array([[[1, 5, 1]]], dtype=np.uint32)
    nets, p = apgx.nets_from_lists([[0, 2]])
    lat = apgx.Lattice(cost, via_cost=0)
    with pytest.raises(apgx.ApgxError) as e:
        lat.negotiate(nets, p, capacity=np.zeros(3, dtype=np.uint8), initial_present_cost=7,
                      maximum_congestion_cost_value=6, maximum_passes=1)
    assert e.value.code == 2
    lat.close()


@pytest.mark.parametrize("group", [0, 2])
def test_value_bound_at_a_later_net(apgx, oracle, gpu, group):
    """The first net's snapshot is within the bound; the second net's is not,
    because the first net installed its route (congestion 0 + 7 * 1 > 6).
    The oracle's one-net order reports the bound; so must both exact modes."""
    cost = np.ones((1, 1, 5), dtype=np.uint32)
    nets, p = apgx.nets_from_lists([[0, 2], [1, 3]])
    o = oracle.negotiate(cost, 0, [(0, 2, None), (2, 2, None)], p, initial_present_cost=7,
                         maximum_congestion_cost_value=6, maximum_passes=1, batch_nets=1)
    assert o["status"] != 0
    lat = apgx.Lattice(cost, via_cost=0)
    with pytest.raises(apgx.ApgxError) as e:
        lat.negotiate(nets, p, initial_present_cost=7, maximum_congestion_cost_value=6, maximum_passes=1,
                      batch_nets=0, group_nets=group)
    assert e.value.code == 2
    # Within the bound (7 <= 7) both exact modes route.
    g = lat.negotiate(nets, p, initial_present_cost=7, maximum_congestion_cost_value=7, maximum_passes=1,
                      batch_nets=0, group_nets=group)
    o = oracle.negotiate(cost, 0, [(0, 2, None), (2, 2, None)], p, initial_present_cost=7,
                         maximum_congestion_cost_value=7, maximum_passes=1, batch_nets=1)
    assert o["status"] == 0
    assert_same(g, o)
    lat.close()


def test_groups_equal_levels_on_the_full_board(apgx, gpu):
    """six_layer_1536 seed 200 (the full 6 x 1536^2 lattice, its first 400
    nets), three passes: speculative groups of 32 give the same routes, planes
    and overuse as window-disjoint levels, both exact orders."""
    _say("[case] six_layer_1536 seed 200, 400 nets, groups vs levels")
    spec = SIX_LAYER_1536
    cost, nets, pins = make_board(apgx, spec, 200)
    sub = (apgx.NetQuery * 400)(*nets[:400])
    lat = apgx.Lattice(cost, via_cost=spec["via"], cost_bytes=2)
    a = lat.negotiate(sub, pins, maximum_passes=3, batch_nets=0, group_nets=0)
    b = lat.negotiate(sub, pins, maximum_passes=3, batch_nets=0, group_nets=32)
    _say(f"[info] levels {a['wall_ms']:.0f} ms, groups {b['wall_ms']:.0f} ms, redone {b['redone_nets']} "
         f"of {b['group_routed_nets']}, overuse {b['overuse_per_pass']}")
    assert_same(b, a)
    assert b["redone_nets"] < b["group_routed_nets"]
    lat.close()


def test_full_board_exact_passes_match_the_oracle_replay(apgx, oracle, gpu):
    """six_layer_1536 seed 200 at full size (the whole oracle run is far too
    slow here): the first nets of pass 0 (base costs) and of pass 1 (exact
    order, present factor 100, history from pass 0's overuse) are the oracle's
    trees, bit for bit, when the oracle replays each pass from the GPU's state
    at its start."""
    spec = SIX_LAYER_1536
    _say("[case] six_layer_1536 seed 200: first nets of passes 0 and 1 vs the oracle replay")
    cost, nets, pins = make_board(apgx, spec, 200)
    lat = apgx.Lattice(cost, via_cost=spec["via"], cost_bytes=2)
    a = lat.negotiate(nets, pins, maximum_passes=1, batch_nets=0, group_nets=0)
    b = lat.negotiate(nets, pins, maximum_passes=2, batch_nets=0, group_nets=0)
    lat.close()
    hist = 10 * np.maximum(a["occupancy"].astype(np.int64) - 1, 0)
    assert np.array_equal(hist, b["history"])  # the update after pass 0
    onets = oracle_nets(nets)
    zeros = np.zeros(cost.shape, dtype=np.int64)
    first0, t0 = oracle.replay_exact_pass(cost, spec["via"], onets, pins, [None] * len(nets), zeros, zeros, 0,
                                          first=3)
    first1, t1 = oracle.replay_exact_pass(cost, spec["via"], onets, pins, a["routes"], a["occupancy"], hist, 1,
                                          first=6)
    _say(f"[info] oracle replay: 3 nets of pass 0 in {t0:.1f} s, 6 nets of pass 1 in {t1:.1f} s")
    for orc, gpu_trees in ((first0, a["routes"]), (first1, b["routes"])):
        for i, t in enumerate(orc):
            assert t is not None and gpu_trees[i] is not None, i
            assert np.array_equal(t, gpu_trees[i]), i
    # Pass 1 moved some of these nets (else the check would only repeat pass 0).
    assert any(not np.array_equal(a["routes"][i], b["routes"][i]) for i in range(6))


def tree_is_legal(route, pins, flat_cost, W, H, L, window):
    cells = set(int(c) for c in route)
    assert len(cells) == len(route)
    assert set(int(p) for p in pins) <= cells
    assert (flat_cost[route] > 0).all()
    x0, y0, w, h = window
    xs, ys = route % W, (route // W) % H
    assert ((xs >= x0) & (xs < x0 + w) & (ys >= y0) & (ys < y0 + h)).all()
    plane = W * H
    seen, stack = {int(pins[0])}, [int(pins[0])]
    while stack:
        c = stack.pop()
        x, y, l = c % W, (c // W) % H, c // plane
        for nb, ok in ((c - 1, x > 0), (c + 1, x + 1 < W), (c - W, y > 0), (c + W, y + 1 < H),
                       (c - plane, l > 0), (c + plane, l + 1 < L)):
            if ok and nb in cells and nb not in seen:
                seen.add(nb)
                stack.append(nb)
    assert seen == cells  # one connected tree over lattice edges


def test_six_layer_1536_full_board(apgx, gpu):
    """Full-size board, three passes in batches of 250: legality of every route
    and the usage / overuse / history identities (the oracle is too slow here)."""
    spec = SIX_LAYER_1536
    _say("[case] six_layer_1536 seed 200: 2000 nets, 3 passes of batched negotiation")
    cost, nets, pins = make_board(apgx, spec, 200)
    L, H, W = cost.shape
    lat = apgx.Lattice(cost, via_cost=spec["via"], cost_bytes=2)
    r = lat.negotiate(nets, pins, maximum_passes=3, batch_nets=250)
    _say(f"[info] termination={r['termination']} passes={r['passes']} overuse={r['overuse_per_pass']} "
         f"wall_ms={r['wall_ms']:.0f}")
    flat = cost.reshape(-1)
    usage = np.zeros(flat.size, dtype=np.int64)
    for i, route in enumerate(r["routes"]):
        n = nets[i]
        if route is None:
            continue
        p = pins[n.pin_offset:n.pin_offset + n.pin_count]
        tree_is_legal(route.astype(np.int64), p, flat, W, H, L,
                      (n.window.x0, n.window.y0, n.window.width, n.window.height))
        usage[route] += 1
    assert np.array_equal(usage, r["occupancy"].reshape(-1))
    assert int(np.maximum(usage - 1, 0).sum()) == r["overuse"]
    assert int(r["history"].sum()) == 10 * sum(r["overuse_per_pass"][:-1])
    assert r["passes"] == 3 and r["routed"] > 0.95 * len(nets)
    lat.close()
