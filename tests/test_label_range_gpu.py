"""The dense search at the edge of its 29-bit label range.

Below 2^29 - 1 every distance and path is exact (against the oracle); a
distance at or beyond it is never reported as "disconnected": the search
fails with APGX_E_UNSUPPORTED.  The same holds for a net cost (the sum of its
connection distances) that exceeds the uint32 result field, and for landmark
fields whose exploring search saturates.  The premises (which distances lie
on which side) are pinned on the CPU by tests/test_label_range.py.
"""
import numpy as np
import pytest

from tests.label_range_cases import MAX_COST, comb, corridor, serpentine

pytestmark = pytest.mark.gpu
UNR = 0xFFFFFFFF


def _say(msg):
    print(msg, flush=True)


def _exact(apgx, oracle, cost, pairs, **kw):
    lat = apgx.Lattice(cost, via_cost=0)
    routes, _ = lat.route_pairs(np.asarray(pairs, dtype=np.uint32), **kw)
    o = oracle.Dense(cost, 0)
    for (s, t), r in zip(pairs, routes):
        dist = o.dijkstra([s])
        if dist[t] == UNR:
            assert r.status == apgx.NET_DISCONNECTED
            continue
        assert r.status == apgx.NET_REACHED and r.cost == dist[t]
        assert np.array_equal(r.connections[0][2], o.canonical_path(dist, t))
    lat.close()


def _refused(apgx, cost, pairs, what, **kw):
    lat = apgx.Lattice(cost, via_cost=0)
    with pytest.raises(apgx.ApgxError) as err:
        lat.route_pairs(np.asarray(pairs, dtype=np.uint32), **kw)
    lat.close()
    assert err.value.code == 1 and what in err.value.detail, err.value.detail


def test_corridor_lds_tiles(apgx, oracle, gpu):
    _say("[case] label range: corridors of cost 2^20 - 1 (LDS tiles)")
    _exact(apgx, oracle, corridor(512), [(0, 511)])          # 535,821,825: just below the range
    _exact(apgx, oracle, corridor(600), [(0, 400)])          # at-risk window, reached
    _exact(apgx, oracle, corridor(600, wall_at=300), [(0, 599)])  # at-risk window, truly disconnected
    _refused(apgx, corridor(600), [(0, 599)], "29-bit")       # 628,096,425: beyond it


def test_serpentine_register_tiles(apgx, oracle, gpu):
    _say("[case] label range: 640 x 640 serpentine of cost 4095 (register tiles)")
    c = serpentine()
    side = c.shape[1]
    _exact(apgx, oracle, c, [(0, 200 * side + 320)], tile=(64, 64))
    _refused(apgx, c, [(0, 638 * side)], "29-bit", tile=(64, 64))
    _refused(apgx, c, [(0, 638 * side)], "29-bit", tile=(32, 32))


def test_net_cost_sum(apgx, oracle, gpu):
    _say("[case] label range: comb nets whose cost sum exceeds uint32")
    for teeth in (8, 9):
        c, pins = comb(teeth)
        nets, flat = apgx.nets_from_lists([pins])
        lat = apgx.Lattice(c, via_cost=0)
        if teeth == 9:
            with pytest.raises(apgx.ApgxError) as err:
                lat.route(nets, flat)
            assert err.value.code == 1 and "uint32 net cost" in err.value.detail
        else:
            routes, _ = lat.route(nets, flat)
            rc, conns = oracle.Dense(c, 0).route_net(pins)
            r = routes[0]
            assert rc == 0 and r.status == apgx.NET_REACHED
            assert r.cost == sum(d for _, d, _ in conns)
            for (gp, gd, gc), (op, od, oc) in zip(r.connections, conns):
                assert (gp, gd) == (op, od) and np.array_equal(gc, oc)
        lat.close()


def test_landmarks_refuse_saturated_fields(apgx, gpu):
    _say("[case] label range: landmark build on the serpentine")
    lat = apgx.Lattice(serpentine(), via_cost=0)
    with pytest.raises(apgx.ApgxError) as err:
        lat.build_landmarks(1)
    assert err.value.code == 1 and "29-bit" in err.value.detail
    lat.close()
    # Below the range the fields build (and equal the oracle's, in test_landmarks_gpu).
    lat = apgx.Lattice(corridor(512, cost=MAX_COST), via_cost=0)
    lat.build_landmarks(1)
    lat.close()
