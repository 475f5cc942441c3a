"""Premises of the label-range cases, pinned by the oracle's 64-bit Dijkstra:
which shortest distances lie at or beyond 2^29 - 1 (where the device labels
saturate) and which lie below it, so that the GPU test's expectations (exact
result below the range, APGX_E_UNSUPPORTED above it) rest on computed
distances rather than on the case's construction."""
import numpy as np

from tests.label_range_cases import LABEL_LIMIT, MAX_COST, REG_MAX_COST, comb, corridor, serpentine

UNR = 0xFFFFFFFF


def test_corridors(oracle):
    d = oracle.Dense(corridor(600), 0).dijkstra([0])
    assert d[599] == 599 * MAX_COST and d[599] >= LABEL_LIMIT
    assert d[400] == 400 * MAX_COST < LABEL_LIMIT
    d = oracle.Dense(corridor(512), 0).dijkstra([0])
    assert d[511] == 511 * MAX_COST and d[511] < LABEL_LIMIT
    d = oracle.Dense(corridor(600, wall_at=300), 0).dijkstra([0])
    assert d[599] == UNR and d[:300].max() < LABEL_LIMIT


def test_serpentine_distances(oracle):
    c = serpentine()
    side = c.shape[1]
    d = oracle.Dense(c, 0).dijkstra([0])
    far, near = 638 * side, 200 * side + 320
    assert d[far] != UNR and d[far] >= LABEL_LIMIT
    assert d[near] < LABEL_LIMIT
    # A single path: the far end's distance is (cells on it - 1) * cost.
    assert d[far] % REG_MAX_COST == 0 and d[far] // REG_MAX_COST > LABEL_LIMIT // REG_MAX_COST


def test_comb_sums(oracle):
    for teeth, over in ((9, True), (8, False)):
        c, pins = comb(teeth)
        rc, conns = oracle.Dense(c, 0).route_net(pins)
        assert rc == 0 and len(conns) == teeth
        assert all(dist < LABEL_LIMIT for _, dist, _ in conns)
        total = sum(dist for _, dist, _ in conns)
        assert (total >= UNR) == over
