"""Pins the single-route generator restatement (oracle/v1.c orc_v1_route,
CPU only).

The reference published rounds and examined work of both single-route CUDA
generators for every corpus case (tests/golden/corpus_golden.json, from
benchmarks/results/phase2_planar_bakeoff_a2b73bd.json); the restatement must
reproduce them exactly.  At convergence every label must equal an independent
Dijkstra over the heading states written here, and every predecessor the
lowest tight predecessor recomputed here.  A hand-worked two-node corridor
fixes rounds, examined work, turn relaxations, labels and predecessors of
both generators, and a round budget that stops each early.
"""
import heapq
import json
from pathlib import Path

import numpy as np
import pytest

from tests import v1_corpus as vc

GOLDEN = json.loads((Path(__file__).parent / "golden" / "corpus_golden.json").read_text())
INF = 0xFFFFFFFFFFFFFFFF
NONE = 0xFFFFFFFF
DX = [1, 1, 0, -1, -1, -1, 0, 1]
DY = [0, 1, 1, 1, 0, -1, -1, -1]


def _oracle(oracle, profile):
    nodes = vc.compile_obstacle_free(profile)
    o = oracle.V1(nodes, profile.costs, profile.heading, profile.tile_w, profile.tile_h)
    order = o.flatten_order(profile.tile_w, profile.tile_h)
    flat = [nodes[i] for i in order]
    return o, flat


def _step(costs, d, inc):
    orth, diag, bend = costs
    return (diag if d & 1 else orth) + (bend if inc != 8 and inc != d else 0)


def _dijkstra(flat, costs, start):
    """Exact heading-state labels over the flattened nodes (written here)."""
    where = {(x, y): i for i, (x, y, _, _) in enumerate(flat)}
    dist = {start * 9 + 8: 0}
    heap = [(0, start * 9 + 8)]
    while heap:
        d0, s = heapq.heappop(heap)
        if d0 != dist[s]:
            continue
        u, inc = divmod(s, 9)
        x, y, _, m = flat[u]
        for d in range(8):
            if not m & (1 << d):
                continue
            t = where[(x + DX[d], y + DY[d])] * 9 + d
            c = d0 + _step(costs, d, inc)
            if c < dist.get(t, INF):
                dist[t] = c
                heapq.heappush(heap, (c, t))
    out = np.full(len(flat) * 9, INF, dtype=np.uint64)
    for s, v in dist.items():
        out[s] = v
    return out


def _lowest_tight(flat, costs, labels, start):
    where = {(x, y): i for i, (x, y, _, _) in enumerate(flat)}
    out = np.full(len(flat) * 9, NONE, dtype=np.uint32)
    for s in range(len(flat) * 9):
        u, inc = divmod(s, 9)
        if s == start * 9 + 8 or inc == 8 or labels[s] == INF:
            continue
        x, y = flat[u][0], flat[u][1]
        p = where.get((x - DX[inc], y - DY[inc]))
        if p is None:
            continue
        for ph in range(9):
            if ph == 8 and p != start:
                continue
            ps = p * 9 + ph
            if labels[ps] != INF and int(labels[ps]) + _step(costs, inc, ph) == int(labels[s]):
                out[s] = ps
                break
    return out


@pytest.mark.parametrize("name", sorted(vc.corpus()))
@pytest.mark.parametrize("generator", [0, 1])
def test_published_rounds_and_work(oracle, name, generator):
    profile, start, goal = vc.corpus()[name]
    o, flat = _oracle(oracle, profile)
    r = oracle.v1_route(o, vc.lattice(profile, start), vc.lattice(profile, goal), generator)
    g = GOLDEN[name]
    key = "cuda_frontier" if generator == 0 else "cuda_sweep"
    if f"{key}_rounds" in g:
        assert r["rounds"] == g[f"{key}_rounds"]
        assert r["examined_work"] == g[f"{key}_examined_work"]
    want = g.get(f"{key}_path_cost", g.get("base_policy_scalar_cost"))
    if name == "disconnected_fields":
        assert r["completion"] == 1
    else:
        assert r["completion"] == 0 and int(r["labels"][r["goal_state"]]) == want
    s = next(i for i, (x, y, _, _) in enumerate(flat) if (x, y) == vc.lattice(profile, start))
    exact = _dijkstra(flat, profile.costs, s)
    assert np.array_equal(r["labels"], exact)
    assert np.array_equal(r["predecessors"], _lowest_tight(flat, profile.costs, exact, s))


def _corridor(oracle):
    # Two nodes, (0, 0) -> (1, 0) east and back west; orth 10, diag 14, bend 3.
    nodes = [(0, 0, 0, 1 << 0), (1, 0, 0, 1 << 4)]
    return oracle.V1(nodes, (10, 14, 3), 1, 4, 4)


def test_corridor_hand_worked(oracle):
    o = _corridor(oracle)
    # Frontier: round 0 relaxes the start (1 edge) to (1,E)=10; round 1 (bucket
    # [10, 19]) relaxes it west to (0,W)=10+10+3=23; round 2 (bucket [20, 29])
    # relaxes (0,W) east, no improvement: converged after 3 rounds, 3 edges.
    f = oracle.v1_route(o, (0, 0), (1, 0), 0)
    assert (f["completion"], f["rounds"], f["examined_work"], f["turns"]) == (0, 3, 3, 0)
    # Sweep: round 1 sets (1,E)=10; round 2 departs node 1 in the 7 other
    # directions through the bend (7 turns) and sets (0,W)=23; round 3 repeats
    # those 7 turns and changes nothing.  2 legal edges per round.
    s = oracle.v1_route(o, (0, 0), (1, 0), 1)
    assert (s["completion"], s["rounds"], s["examined_work"], s["turns"]) == (0, 3, 6, 14)
    for r in (f, s):
        lab = r["labels"]
        assert int(lab[0 * 9 + 8]) == 0 and int(lab[1 * 9 + 0]) == 10 and int(lab[0 * 9 + 4]) == 23
        assert sum(int(v != INF) for v in lab) == 3
        assert r["predecessors"][1 * 9 + 0] == 8 and r["predecessors"][0 * 9 + 4] == 9
        assert r["goal_state"] == 9
    # A one-round budget stops both before convergence.
    for gen in (0, 1):
        b = oracle.v1_route(o, (0, 0), (1, 0), gen, max_rounds=1)
        assert b["completion"] == 2 and b["rounds"] == 1
