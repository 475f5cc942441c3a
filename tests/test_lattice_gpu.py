"""GPU parity of the dense lattice search against the CPU oracle.

Bar: distances bit-exact and every connection path identical to the oracle's
canonical path; at full size, size-independent properties (adjacency, exact
path cost, canonical-predecessor condition) for every query.
"""
import numpy as np
import pytest

from tests.lattice_cases import EIGHT_LAYER_2048, FOUR_LAYER_1024, SINGLE_LAYER_256, make_lattice
from tests.test_oracle_dense import check_canonical, random_lattice

pytestmark = pytest.mark.gpu
UNR = 0xFFFFFFFF


def _say(msg):
    print(msg, flush=True)


def test_single_layer_256_full_workload(apgx, oracle, gpu):
    spec = SINGLE_LAYER_256
    _say("[case] single_layer_256: 64 pairs, seed 1")
    cost = make_lattice(apgx, spec)
    pairs = apgx.generate_pairs(cost, spec["queries"], 101)
    lat = apgx.Lattice(cost, via_cost=0, cost_bytes=4)
    routes, stats = lat.route_pairs(pairs)
    o = oracle.Dense(cost, 0)
    for (s, t), r in zip(pairs, routes):
        d_astar, _, _ = o.astar(int(s), int(t))
        dist = o.dijkstra([int(s)])
        assert d_astar == dist[t]
        if d_astar == UNR:
            assert r.status == apgx.NET_DISCONNECTED
            continue
        assert r.status == apgx.NET_REACHED
        assert r.cost == d_astar
        pin, d, cells = r.connections[0]
        assert pin == 1 and d == d_astar
        assert np.array_equal(cells, o.canonical_path(dist, int(t)))
    assert stats.tile_visits > 0 and stats.device_ms > 0
    lat.close()


@pytest.mark.parametrize("seed", range(8))
def test_random_small_lattices(apgx, oracle, gpu, seed):
    rng = np.random.default_rng(1000 + seed)
    L = [1, 2, 4, 3][seed % 4]
    H, W = [(37, 53), (64, 64), (70, 129), (100, 33)][seed % 4]
    cost = random_lattice(rng, L, H, W, blocked=0.2, cmax=[3, 8, 16, 2][seed % 4])
    via = 0 if L == 1 else [20, 1, 7, 3][seed % 4]
    _say(f"[case] random lattice seed={seed} shape={(L, H, W)} via={via}")
    free = np.flatnonzero(cost.reshape(-1))
    pairs = rng.choice(free, size=(12, 2), replace=True)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    tile = [(0, 0), (32, 16), (32, 32), (64, 32)][seed % 4]
    delta = [0, 0xFFFFFFFF, 1, 50][(seed // 2) % 4]
    lat = apgx.Lattice(cost, via_cost=via)
    routes, _ = lat.route_pairs(pairs, tile=tile, frontier_delta=delta, goal_bound=seed % 3 != 2)
    o = oracle.Dense(cost, via)
    for (s, t), r in zip(pairs, routes):
        dist = o.dijkstra([int(s)])
        if dist[t] == UNR:
            assert r.status == apgx.NET_DISCONNECTED
            continue
        assert r.status == apgx.NET_REACHED and r.cost == dist[t]
        assert np.array_equal(r.connections[0][2], o.canonical_path(dist, int(t)))
    lat.close()


@pytest.mark.parametrize("seed", range(4))
def test_windows_and_multi_pin_nets(apgx, oracle, gpu, seed):
    rng = np.random.default_rng(2000 + seed)
    L, H, W = 3, 90, 110
    cost = random_lattice(rng, L, H, W, blocked=0.15, cmax=8)
    via = 20
    _say(f"[case] multi-pin nets seed={seed}")
    pin_lists, windows = [], []
    for _ in range(10):
        w, h = int(rng.integers(8, W)), int(rng.integers(8, H))
        x0, y0 = int(rng.integers(0, W - w + 1)), int(rng.integers(0, H - h + 1))
        ll, yy, xx = np.meshgrid(np.arange(L), np.arange(y0, y0 + h), np.arange(x0, x0 + w), indexing="ij")
        cells = ((ll * H + yy) * W + xx).reshape(-1)
        cells = cells[cost.reshape(-1)[cells] > 0]
        k = int(rng.integers(2, 7))
        pin_lists.append(rng.choice(cells, size=k, replace=False))
        windows.append((x0, y0, w, h))
    nets, pins = apgx.nets_from_lists(pin_lists, windows)
    lat = apgx.Lattice(cost, via_cost=via)
    routes, stats = lat.route(nets, pins, tile=[(0, 0), (32, 16), (32, 32), (64, 32)][seed],
                              frontier_delta=[0, 0xFFFFFFFF, 7, 0][seed], goal_bound=seed != 3,
                              use_graph=seed % 2, rounds_per_sync=[8, 2, 8, 3][seed])
    assert stats.graph_launches > 0 or seed % 2 == 0
    for pl, win, r in zip(pin_lists, windows, routes):
        rc, conns = oracle.Dense(cost, via, win).route_net(pl)
        if rc == 1:
            assert r.status == apgx.NET_DISCONNECTED
            continue
        assert r.status == apgx.NET_REACHED
        assert len(r.connections) == len(conns)
        for (gp, gd, gc), (op, od, oc) in zip(r.connections, conns):
            assert (gp, gd) == (op, od)
            assert np.array_equal(gc, oc)
        assert r.cost == sum(d for _, d, _ in conns)
    lat.close()


@pytest.mark.parametrize("layers", [1, 3])
def test_costs_above_register_tile_limit(apgx, oracle, gpu, layers):
    """Cell costs above 4095 do not fit the register tile's packed halves; the
    search must fall back to the LDS tile kernel and stay exact."""
    _say(f"[case] large costs layers={layers}")
    rng = np.random.default_rng(77 + layers)
    cost = random_lattice(rng, layers, 70, 90, blocked=0.1, cmax=8)
    cost = (cost.astype(np.int64) * 9000).astype(np.uint32)  # up to 72000
    via = 0 if layers == 1 else 40000
    free = np.flatnonzero(cost.reshape(-1))
    pairs = rng.choice(free, size=(10, 2), replace=False)
    lat = apgx.Lattice(cost, via_cost=via)
    routes, stats = lat.route_pairs(pairs)
    o = oracle.Dense(cost, via)
    for (s, t), r in zip(pairs, routes):
        dist = o.dijkstra([int(s)])
        if dist[t] == UNR:
            assert r.status == apgx.NET_DISCONNECTED
            continue
        assert r.status == apgx.NET_REACHED and r.cost == dist[t]
        assert np.array_equal(r.connections[0][2], o.canonical_path(dist, int(t)))
    lat.close()


def test_edge_cases(apgx, gpu):
    _say("[case] edge cases")
    cost = np.ones((1, 16, 16), dtype=np.uint32)
    cost[0, :, 8] = 0  # wall: left and right halves disconnected
    lat = apgx.Lattice(cost, via_cost=0)
    # empty batch
    routes, stats = lat.route_pairs(np.zeros((0, 2), dtype=np.uint32))
    assert routes == []
    # disconnected, blocked pin (invalid), identical pins (invalid), neighbour pair
    pairs = np.array([[0, 15], [8, 0], [3, 3], [0, 1]], dtype=np.uint32)
    routes, _ = lat.route_pairs(pairs)
    assert [r.status for r in routes] == [apgx.NET_DISCONNECTED, apgx.NET_INVALID,
                                          apgx.NET_INVALID, apgx.NET_REACHED]
    assert routes[3].cost == 1 and list(routes[3].connections[0][2]) == [1, 0]
    lat.close()
    # A one-round budget cannot finish a route that crosses six 32 x 16 tiles:
    # round 1 visits only the source tile.  Without the budget it is reached.
    lat = apgx.Lattice(np.ones((1, 16, 200), dtype=np.uint32), via_cost=0)
    pairs = np.array([[0, 199]], dtype=np.uint32)
    routes, st = lat.route_pairs(pairs, max_rounds=1, rounds_per_sync=1, tile=(32, 16))
    assert routes[0].status == apgx.NET_BUDGET_EXHAUSTED and routes[0].cost == UNR and st.rounds == 1
    routes, _ = lat.route_pairs(pairs, tile=(32, 16))
    assert routes[0].status == apgx.NET_REACHED and routes[0].cost == 199
    lat.close()


def test_repeatable(apgx, gpu):
    _say("[case] repeatability")
    rng = np.random.default_rng(5)
    cost = random_lattice(rng, 2, 80, 80, blocked=0.1, cmax=4)
    free = np.flatnonzero(cost.reshape(-1))
    pairs = rng.choice(free, size=(20, 2), replace=False)
    lat = apgx.Lattice(cost, via_cost=3)
    a, _ = lat.route_pairs(pairs)
    b, _ = lat.route_pairs(pairs)
    for x, y in zip(a, b):
        assert (x.status, x.cost) == (y.status, y.cost)
        assert all(np.array_equal(p[2], q[2]) for p, q in zip(x.connections, y.connections))
    lat.close()


def _oracle_pool(fn, items, threads=16):
    """Runs oracle checks in threads (the oracle's C calls release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=threads) as pool:
        return list(pool.map(fn, items))


def test_four_layer_1024_full_size(apgx, oracle, gpu):
    """The benchmark's configuration exactly: 16 landmark fields, default
    ordering weight, bucket and sweeps per visit, 1024 pairs of seed 202.
    Every query: adjacency and exact path cost; every 8th query (128): distance
    and canonical path against the oracle's Dijkstra."""
    spec = FOUR_LAYER_1024
    _say("[case] four_layer_1024: 1024 pairs, 16 landmark fields (bench options), 128 oracle-checked")
    cost = make_lattice(apgx, spec)
    pairs = apgx.generate_pairs(cost, spec["queries"], 202)
    lat = apgx.Lattice(cost, via_cost=spec["via"], cost_bytes=2)
    lat.build_landmarks(16)
    routes, stats = lat.route_pairs(pairs, landmarks=True)
    _say(f"[info] device_ms={stats.device_ms:.2f} rounds={stats.rounds} tiles={stats.tile_visits}")
    L, H, W = cost.shape
    plane = H * W
    flat = cost.reshape(-1)
    for (s, t), r in zip(pairs, routes):
        assert r.status in (apgx.NET_REACHED, apgx.NET_DISCONNECTED)
        if r.status != apgx.NET_REACHED:
            continue
        cells = r.connections[0][2].astype(np.int64)
        assert cells[0] == t and cells[-1] == s
        steps = np.abs(np.diff(cells))
        assert np.isin(steps, [1, W, plane]).all()
        w = np.where(steps == plane, spec["via"], flat[cells[:-1]].astype(np.int64))
        assert int(w.sum()) == r.cost
    o = oracle.Dense(cost, spec["via"])

    def expect(i):
        s, t = (int(v) for v in pairs[i])
        dist = o.dijkstra([s])
        if dist[t] == UNR:
            return i, UNR, None
        return i, int(dist[t]), o.canonical_path(dist, t)

    checked = _oracle_pool(expect, range(0, spec["queries"], 8))
    assert len(checked) == 128
    for i, d, path in checked:
        r = routes[i]
        if d == UNR:
            assert r.status == apgx.NET_DISCONNECTED, i
            continue
        assert r.status == apgx.NET_REACHED and r.cost == d, (i, r.cost, d)
        assert np.array_equal(r.connections[0][2], path), i
    _say(f"[info] 128 queries bit-exact vs oracle Dijkstra")
    lat.close()


def check_net_structure(flat_cost, W, plane, via, pins, window, route):
    """Every connection of a reached net: starts at its pin, ends on the tree
    built so far, steps to lattice neighbours inside the window through free
    cells, and its cost equals its distance."""
    x0, y0, w, h = window
    tree = {int(pins[0])}
    total = 0
    for pin, d, cells in route.connections:
        cells = cells.astype(np.int64)
        assert cells[0] == pins[pin] and int(cells[-1]) in tree
        steps = np.abs(np.diff(cells))
        assert np.isin(steps, [1, W, plane]).all()
        xs, ys = cells % W, (cells // W) % (plane // W)
        assert ((xs >= x0) & (xs < x0 + w) & (ys >= y0) & (ys < y0 + h)).all()
        assert (flat_cost[cells] > 0).all()
        cw = np.where(steps == plane, via, flat_cost[cells[:-1]].astype(np.int64))
        assert int(cw.sum()) == d
        total += d
        tree.update(int(c) for c in cells)
    assert total == route.cost


def test_eight_layer_2048_full_size(apgx, oracle, gpu):
    """eight_layer_2048 with the benchmark's search options (16 landmark
    fields, graph-replayed round chunks, 32 sweeps per visit as the autotuner
    picks on this workload).  Every reached net: connection structure and exact
    costs; against the oracle: six nets spread over the lower half of window
    sides and four from the top decile, the largest window (side 2046) among
    them."""
    spec = EIGHT_LAYER_2048
    _say("[case] eight_layer_2048: 4096 nets of 2-8 pins, log-uniform windows, 16 landmarks, graph launches")
    cost = make_lattice(apgx, spec)
    nets, pins = apgx.generate_nets(cost, spec["nets"], spec["min_pins"], spec["max_pins"],
                                    spec["min_side"], spec["max_side"], spec["seed"])
    lat = apgx.Lattice(cost, via_cost=spec["via"], cost_bytes=2)
    lat.build_landmarks(16)
    routes, stats = lat.route(nets, pins, use_graph=1, landmarks=True, max_sweeps=32)
    _say(f"[info] device_ms={stats.device_ms:.1f} rounds={stats.rounds} tiles={stats.tile_visits} "
         f"graph_launches={stats.graph_launches}")
    assert stats.graph_launches > 0
    L, H, W = cost.shape
    plane = H * W
    flat = cost.reshape(-1)
    reached = 0
    for i, r in enumerate(routes):
        n = nets[i]
        assert r.status in (apgx.NET_REACHED, apgx.NET_DISCONNECTED)
        if r.status == apgx.NET_REACHED:
            reached += 1
            assert len(r.connections) == n.pin_count - 1
            p = pins[n.pin_offset:n.pin_offset + n.pin_count]
            win = (n.window.x0, n.window.y0, n.window.width, n.window.height)
            check_net_structure(flat, W, plane, spec["via"], p, win, r)
    assert reached > 0.9 * len(routes)
    sides = np.array([n.window.width for n in nets])
    counts = np.array([n.pin_count for n in nets])
    order = [int(i) for i in np.argsort(sides, kind="stable")]
    sample = [order[q * len(order) // 12] for q in range(6)]
    top = order[int(0.9 * len(order)):]
    multi = [i for i in top if counts[i] >= 4]
    picks = []
    for i in (top[-1], min(multi, key=lambda i: sides[i]), top[len(top) // 2], top[0], top[len(top) // 4]):
        if i not in picks:
            picks.append(i)
    picks = picks[:4]
    assert len(picks) == 4 and all(sides[i] >= sides[top[0]] for i in picks)
    sample += picks
    _say(f"[info] oracle nets: sides {[int(sides[i]) for i in sample]}, pins {[int(counts[i]) for i in sample]}")

    def expect(i):
        n = nets[i]
        win = (n.window.x0, n.window.y0, n.window.width, n.window.height)
        p = pins[n.pin_offset:n.pin_offset + n.pin_count]
        return i, oracle.Dense(cost, spec["via"], win).route_net(p)

    for i, (rc, conns) in _oracle_pool(expect, sample):
        r = routes[i]
        if rc == 1:
            assert r.status == apgx.NET_DISCONNECTED, i
            continue
        assert r.status == apgx.NET_REACHED and len(r.connections) == len(conns), i
        for (gp, gd, gc), (op, od, oc) in zip(r.connections, conns):
            assert (gp, gd) == (op, od) and np.array_equal(gc, oc), i
    lat.close()


def test_autotune_picks_a_measured_candidate(apgx, gpu):
    _say("[case] autotune on 96 nets of a 2x300x300 lattice")
    rng = np.random.default_rng(9)
    cost = random_lattice(rng, 2, 300, 300, blocked=0.1, cmax=8)
    nets, pins = apgx.generate_nets(cost, 96, 2, 6, 32, 300, 11)
    lat = apgx.Lattice(cost, via_cost=10)
    opts, report = lat.autotune(nets, pins, sample_nets=32)
    assert report["sample_nets"] == 32 and len(report["candidates"]) >= 8
    best = report["candidates"][report["best"]]
    assert all(best[4] <= c[4] for c in report["candidates"])
    assert opts["tile"] == best[:2]
    a, _ = lat.route(nets, pins, **opts)
    b, _ = lat.route(nets, pins)
    for x, y in zip(a, b):
        assert (x.status, x.cost) == (y.status, y.cost)
        assert all(np.array_equal(p[2], q[2]) for p, q in zip(x.connections, y.connections))
    lat.close()


def test_autotune_candidates_agree_with_the_oracle(apgx, oracle, gpu):
    """Every candidate the autotuner times (tile shapes 64x64 register / 64x32
    and 32x32 LDS, bucket widths, sweep caps; with landmark bounds, as the
    eight_layer_2048 benchmark runs it) routes the whole batch to the same
    distances and canonical paths, and those equal the oracle's: the choice
    changes speed only.  The reported choice is the fastest timing."""
    _say("[case] autotune with landmarks on 64 nets of a 3x256x256 lattice: every candidate vs the oracle")
    rng = np.random.default_rng(21)
    cost = random_lattice(rng, 3, 256, 256, blocked=0.12, cmax=8)
    via = 20
    nets, pins = apgx.generate_nets(cost, 64, 2, 6, 32, 256, 13)
    lat = apgx.Lattice(cost, via_cost=via)
    lat.build_landmarks(8)
    opts, report = lat.autotune(nets, pins, sample_nets=32, landmarks=True)
    cands = report["candidates"]
    assert len(cands) >= 12 and any(c[3] == 32 for c in cands)  # the landmark schedule is tried
    best = cands[report["best"]]
    assert all(best[4] <= c[4] for c in cands)
    assert opts["tile"] == best[:2] and opts["max_sweeps"] == best[3]
    expected = []
    for n in nets:
        win = (n.window.x0, n.window.y0, n.window.width, n.window.height)
        expected.append(oracle.Dense(cost, via, win).route_net(pins[n.pin_offset:n.pin_offset + n.pin_count]))
    for tw, th, delta, sweeps, _ in cands:
        routes, _ = lat.route(nets, pins, tile=(tw, th), frontier_delta=delta, max_sweeps=sweeps, landmarks=True)
        for i, (r, (rc, conns)) in enumerate(zip(routes, expected)):
            if rc == 1:
                assert r.status == apgx.NET_DISCONNECTED, (i, tw, th, delta, sweeps)
                continue
            assert r.status == apgx.NET_REACHED and len(r.connections) == len(conns), (i, tw, th, delta, sweeps)
            for (gp, gd, gc), (op, od, oc) in zip(r.connections, conns):
                assert (gp, gd) == (op, od) and np.array_equal(gc, oc), (i, tw, th, delta, sweeps)
    lat.close()


@pytest.mark.parametrize("schedule,max_sweeps,seed", [(1, 0, 0), (2, 0, 0), (2, 0, 1), (2, 3, 2), (2, 1, 3),
                                                      (1, 2, 4), (2, 0, 5), (3, 0, 1), (4, 5, 3), (3, 2, 0)])
def test_register_tiles_sweep_schedules(apgx, oracle, gpu, schedule, max_sweeps, seed):
    """Register-resident 64 x 64 x L tiles under both sweep schedules (every
    sweep full, or in-layer-only sweeps alternating with full ones) and small
    sweep caps that end visits on either kind of sweep: distances and canonical
    paths bit-exact against the oracle, vias on."""
    rng = np.random.default_rng(3000 + seed)
    L = [2, 3, 4, 6, 4, 5][seed % 6]
    H, W = [(130, 70), (64, 128), (100, 100), (70, 129), (128, 64), (65, 65)][seed % 6]
    cost = random_lattice(rng, L, H, W, blocked=0.2, cmax=[8, 3, 16, 8, 5, 2][seed % 6])
    via = [20, 1, 7, 3, 20, 5][seed % 6]
    _say(f"[case] register tiles schedule={schedule} max_sweeps={max_sweeps} shape={(L, H, W)} via={via}")
    free = np.flatnonzero(cost.reshape(-1))
    pairs = rng.choice(free, size=(10, 2), replace=True)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    lat = apgx.Lattice(cost, via_cost=via)
    routes, st = lat.route_pairs(pairs, tile=(64, 64), max_sweeps=max_sweeps, sweep_schedule=schedule)
    assert (st.tile_w, st.tile_h) == (64, 64)
    o = oracle.Dense(cost, via)
    for (s, t), r in zip(pairs, routes):
        dist = o.dijkstra([int(s)])
        if dist[t] == UNR:
            assert r.status == apgx.NET_DISCONNECTED
            continue
        assert r.status == apgx.NET_REACHED and r.cost == dist[t]
        assert np.array_equal(r.connections[0][2], o.canonical_path(dist, int(t)))
    lat.close()
