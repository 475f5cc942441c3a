"""ALT landmark goal bounds (apgx_lattice_build_landmarks, route option
`landmarks`).

Bar: each landmark field equals the oracle's Dijkstra distances from its
landmark cell (unreachable and blocked cells read 0x1FFFFFFF); routes with the
bound are bit-identical to the oracle (distances and canonical paths) on random
lattices with disconnected regions, windows and multi-pin nets, on register and
LDS tiles and under every schedule; the allocator's override planes give the
oracle's allocation; at four_layer_1024 size every query equals the search
without the bound while visiting fewer tiles."""
import numpy as np
import pytest

from tests.lattice_cases import FOUR_LAYER_1024, SIX_LAYER_1536, make_board, make_lattice
from tests.test_negotiate_gpu import SMALL, assert_same, oracle_nets
from tests.test_oracle_dense import random_lattice

pytestmark = pytest.mark.gpu
UNR = 0xFFFFFFFF
INF = 0x1FFFFFFF


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("shape, via, count", [((1, 70, 90), 0, 8), ((3, 130, 100), 20, 8), ((2, 64, 200), 5, 3),
                                               ((2, 96, 96), 4, 16)])
def test_fields_match_oracle(apgx, oracle, gpu, shape, via, count):
    _say(f"[case] landmark fields {shape} via={via} count={count}")
    rng = np.random.default_rng(sum(shape) + count)
    cost = random_lattice(rng, *shape, blocked=0.25, cmax=8)
    lat = apgx.Lattice(cost, via_cost=via)
    ms = lat.build_landmarks(count)
    assert lat.landmark_count == count and ms > 0
    o = oracle.Dense(cost, via)
    flat = cost.reshape(-1)
    sources = set()
    for k in range(count):
        f = lat.landmark_field(k).reshape(-1)
        src = np.flatnonzero(f == 0)
        assert src.size == 1 and flat[src[0]] > 0
        sources.add(int(src[0]))
        want = o.dijkstra([int(src[0])])
        assert np.array_equal(f, np.where(want == UNR, INF, want))
    assert len(sources) == count
    # The first landmark is the free cell nearest the corner (0, 0) on layer 0.
    if flat[0] > 0:
        assert lat.landmark_field(0).reshape(-1)[0] == 0
    lat.build_landmarks(0)
    assert lat.landmark_count == 0
    lat.close()


def test_route_without_fields_is_refused(apgx, gpu):
    _say("[case] landmarks requested before any are built")
    cost = np.ones((1, 16, 16), np.uint32)
    lat = apgx.Lattice(cost)
    with pytest.raises(apgx.ApgxError) as e:
        lat.route_pairs(np.array([[0, 255]]), landmarks=True)
    assert "none are built" in e.value.detail
    with pytest.raises(apgx.ApgxError):
        lat.build_landmarks(17)
    lat.close()


@pytest.mark.parametrize("seed", range(8))
def test_pairs_match_oracle(apgx, oracle, gpu, seed):
    rng = np.random.default_rng(3000 + seed)
    L = [1, 2, 4, 3][seed % 4]
    H, W = [(37, 53), (96, 64), (70, 129), (100, 160)][seed % 4]
    cost = random_lattice(rng, L, H, W, blocked=[0.2, 0.3][seed % 2], cmax=[3, 8, 16, 8][seed % 4])
    via = 0 if L == 1 else [20, 1, 7, 3][seed % 4]
    tile = [(0, 0), (32, 16), (32, 32), (64, 32)][seed % 4]
    delta = [0, 0xFFFFFFFF, 1, 50][(seed // 2) % 4]
    count = [8, 1, 4, 16][seed % 4]
    weight = [0, 16, 1, 8, 16, 0, 12, 1][seed]
    _say(f"[case] landmark pairs seed={seed} shape={(L, H, W)} via={via} tile={tile} delta={delta} "
         f"count={count} weight={weight}")
    free = np.flatnonzero(cost.reshape(-1))
    pairs = rng.choice(free, size=(16, 2), replace=True)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    lat = apgx.Lattice(cost, via_cost=via)
    lat.build_landmarks(count)
    routes, _ = lat.route_pairs(pairs, tile=tile, frontier_delta=delta, landmarks=True, landmark_weight=weight)
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
    rng = np.random.default_rng(4000 + seed)
    L, H, W = 3, 90, 150
    cost = random_lattice(rng, L, H, W, blocked=0.15, cmax=8)
    via = 20
    _say(f"[case] landmark multi-pin nets seed={seed}")
    pin_lists, windows = [], []
    for _ in range(12):
        w, h = int(rng.integers(8, W)), int(rng.integers(8, H))
        x0, y0 = int(rng.integers(0, W - w + 1)), int(rng.integers(0, H - h + 1))
        ll, yy, xx = np.meshgrid(np.arange(L), np.arange(y0, y0 + h), np.arange(x0, x0 + w), indexing="ij")
        cells = ((ll * H + yy) * W + xx).reshape(-1)
        cells = cells[cost.reshape(-1)[cells] > 0]
        pin_lists.append(rng.choice(cells, size=int(rng.integers(2, 7)), replace=False))
        windows.append((x0, y0, w, h))
    nets, pins = apgx.nets_from_lists(pin_lists, windows)
    lat = apgx.Lattice(cost, via_cost=via)
    lat.build_landmarks(8)
    routes, _ = lat.route(nets, pins, tile=[(0, 0), (32, 16), (32, 32), (64, 32)][seed],
                          frontier_delta=[0, 0xFFFFFFFF, 7, 0][seed], use_graph=seed % 2, landmarks=True)
    for pl, win, r in zip(pin_lists, windows, routes):
        rc, conns = oracle.Dense(cost, via, win).route_net(pl)
        if rc == 1:
            assert r.status == apgx.NET_DISCONNECTED
            continue
        assert r.status == apgx.NET_REACHED
        assert [(p, d) for p, d, _ in r.connections] == [(p, d) for p, d, _ in conns]
        for (_, _, gc), (_, _, oc) in zip(r.connections, conns):
            assert np.array_equal(gc, oc)
    lat.close()


def test_allocator_with_landmarks_matches_oracle(apgx, oracle, gpu):
    """Override planes (base + congestion, vias paying the entered cell) cost
    at least the fields' metric, so the bound stays admissible there."""
    _say("[case] allocator parity with landmark bounds")
    cost, nets, pins = make_board(apgx, SIX_LAYER_1536, 1003, **SMALL)
    via = SIX_LAYER_1536["via"]
    lat = apgx.Lattice(cost, via_cost=via, cost_bytes=2)
    lat.build_landmarks(8)
    g = lat.negotiate(nets, pins, maximum_passes=12, batch_nets=4, search_options={"landmarks": 1})
    o = oracle.negotiate(cost, via, oracle_nets(nets), pins, maximum_passes=12, batch_nets=4)
    assert o["status"] == 0
    assert_same(g, o)
    lat.close()


def test_four_layer_1024_same_routes_fewer_visits(apgx, gpu):
    spec = FOUR_LAYER_1024
    _say("[case] four_layer_1024 with and without landmarks")
    cost = make_lattice(apgx, spec)
    pairs = apgx.generate_pairs(cost, spec["queries"], 202)
    lat = apgx.Lattice(cost, via_cost=spec["via"], cost_bytes=2)
    plain, s0 = lat.route_pairs(pairs)
    ms = lat.build_landmarks(8)
    alt, s1 = lat.route_pairs(pairs, landmarks=True)
    _say(f"[info] build_ms={ms:.2f} visits {s0.tile_visits} -> {s1.tile_visits} "
         f"device_ms {s0.device_ms:.2f} -> {s1.device_ms:.2f}")
    for a, b in zip(plain, alt):
        assert (a.status, a.cost) == (b.status, b.cost)
        assert all(np.array_equal(p[2], q[2]) for p, q in zip(a.connections, b.connections))
    assert s1.tile_visits < s0.tile_visits
    lat.close()
