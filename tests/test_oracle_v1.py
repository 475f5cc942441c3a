"""Pins the planar heading-state oracle and the corpus restatement (CPU only).

Golden values come from the reference's own published results
(tests/golden/corpus_golden.json) and its CPU A* known-answer tests
(tests/routing/cpu_astar_test.cc).
"""
import json
from pathlib import Path

import pytest

from tests import v1_corpus as vc

GOLDEN = json.loads((Path(__file__).parent / "golden" / "corpus_golden.json").read_text())


def _oracle_for(oracle, profile):
    nodes = vc.compile_obstacle_free(profile)
    return nodes, oracle.V1(nodes, profile.costs, profile.heading, profile.tile_w, profile.tile_h)


@pytest.mark.parametrize("name", sorted(vc.corpus()))
def test_corpus_restatement_matches_published_descriptor(name):
    profile, _, _ = vc.corpus()[name]
    desc = GOLDEN[name]["descriptor"]
    nodes = vc.compile_obstacle_free(profile)
    assert profile.fingerprint() == int(desc["compiler_profile_fingerprint"])
    assert len(nodes) == int(desc["nodes"])
    legal = sum(bin(m).count("1") for *_, m in nodes)
    assert legal == int(desc["legal_edges"])
    assert desc["tile"] == f"{profile.tile_w}x{profile.tile_h}"
    assert desc["heading_mask"] == str(profile.heading)
    assert desc["base_costs"] == ",".join(map(str, profile.costs))


@pytest.mark.parametrize("name", sorted(vc.corpus()))
def test_astar_reproduces_published_costs_and_expansions(oracle, name):
    profile, start, goal = vc.corpus()[name]
    _, o = _oracle_for(oracle, profile)
    cost, path, expanded = o.astar(vc.lattice(profile, start), vc.lattice(profile, goal))
    g = GOLDEN[name]
    want = g.get("cpu_astar_path_cost", g.get("base_policy_scalar_cost"))
    if name == "disconnected_fields":
        assert cost is None
    else:
        assert cost == want
    if "cpu_astar_examined_work" in g:
        assert expanded == g["cpu_astar_examined_work"]
    ccost, states = o.canonical(vc.lattice(profile, start), vc.lattice(profile, goal))
    assert ccost == cost


def test_cheap_diagonal_known_answer(oracle):
    profile, start, goal = vc.known_answer_boards()["cheap_diagonal"]
    _, o = _oracle_for(oracle, profile)
    cost, path, _ = o.astar(vc.lattice(profile, start), vc.lattice(profile, goal))
    assert cost == 37
    assert path == [(0, 0), (1, 1), (2, 2), (3, 1), (4, 0), (3, -1), (2, -2), (3, -3)]


def test_ban_and_penalty_detours(oracle):
    profile, start, goal = vc.known_answer_boards()["horizontal_detour"]
    _, o = _oracle_for(oracle, profile)
    s, t = vc.lattice(profile, start), vc.lattice(profile, goal)
    middle = (4, 0, 0)
    assert o.astar(s, t)[0] == 100
    assert o.astar(s, t, oracle.V1.policy(bans=[middle]))[0] == 126
    assert o.astar(s, t, oracle.V1.policy(penalties=[(middle, 200)]))[0] == 126
    assert o.canonical(s, t, oracle.V1.policy(bans=[middle]))[0] == 126


def test_surcharges_and_merged_penalty_both_directions(oracle):
    profile, start, goal = vc.known_answer_boards()["horizontal_corridor"]
    _, o = _oracle_for(oracle, profile)
    s, t = vc.lattice(profile, start), vc.lattice(profile, goal)
    pol = oracle.V1.policy(2, 3, 5, penalties=[((4, 0, 0), 10), ((4, 0, 0), 7)])
    assert o.astar(s, t, pol)[0] == 137
    assert o.astar(t, s, pol)[0] == 137
    assert o.canonical(t, s, pol)[0] == 137
    assert o.astar(s, t, oracle.V1.policy(bans=[(4, 0, 0)]))[0] is None


def test_stable_hash_string_matches_replay_checksum():
    # A published replay trailer: FNV-1a of every byte before the checksum line.
    text = (Path(__file__).parent / "golden" / "reference_replays" / "batch_workspace_owner_v1.replay").read_bytes()
    payload, _, last = text.rstrip(b"\n").rpartition(b"\n")
    assert last.startswith(b"checksum_fnv1a64=")
    assert vc.stable_hash_string(payload + b"\n") == int(last.split(b"=")[1])


def test_frontier_restatement_hand_worked(oracle):
    """Horizontal corridor of 11 nodes (orth 10, bend 3), start x=0, goal x=10:
    the frontier closes the start state, the nine interior (node, E) states and
    the goal's (node, E) state in order of f = 100 (ties broken by lower g),
    examining 1 edge at the start and 2 (E and W) at each interior node:
    11 rounds, 19 examined edges; a 4-round budget stops after 1 + 2 + 2 + 2."""
    p, s, g = vc.known_answer_boards()["horizontal_corridor"]
    _, o = _oracle_for(oracle, p)
    f = oracle.v1_frontier(o, vc.lattice(p, s), vc.lattice(p, g))
    assert (f["status"], f["cost"], f["rounds"], f["examined_work"]) == (0, 100, 11, 19)
    assert f["goal_state"] == 10 * 9 + 0 and f["states"] == [9 * k for k in range(10, 0, -1)] + [8]
    f = oracle.v1_frontier(o, vc.lattice(p, s), vc.lattice(p, g), max_rounds=4)
    assert (f["status"], f["rounds"], f["examined_work"]) == (2, 4, 7)


def test_frontier_workspace_hand_worked(oracle):
    """Workspace of the corridor frontier above when it closes the goal: each
    closed (k, E) state relaxed (k+1, E) to 10(k+1) and (k-1, W) to
    10k + 10 + 3 (a turn); (9, W) would come from the goal, which is never
    expanded.  Predecessors are the relaxing states; the start's is none."""
    p, s, g = vc.known_answer_boards()["horizontal_corridor"]
    _, o = _oracle_for(oracle, p)
    w = oracle.v1_frontier_workspace(o, vc.lattice(p, s), vc.lattice(p, g))
    lab, pre = w["labels"], w["predecessors"]
    assert (w["status"], w["goal_state"], w["rounds"]) == (0, 90, 11)
    inf, none = (1 << 64) - 1, 0xFFFFFFFF
    want_lab = {8: 0}
    want_pre = {}
    for k in range(1, 11):
        want_lab[9 * k] = 10 * k
        want_pre[9 * k] = 9 * (k - 1) if k > 1 else 8
    for k in range(0, 9):
        want_lab[9 * k + 4] = 10 * k + 23
        want_pre[9 * k + 4] = 9 * (k + 1)
    for st in range(len(lab)):
        assert int(lab[st]) == want_lab.get(st, inf), st
        assert int(pre[st]) == want_pre.get(st, none), st


@pytest.mark.parametrize("name", sorted(vc.corpus()))
def test_frontier_workspace_is_consistent(oracle, name):
    """On every corpus case: the walk along the workspace predecessors from the
    goal state is the frontier's path, the goal label is the published cost,
    and every finite label is its predecessor's label plus the step cost
    (orthogonal / diagonal step, plus the bend on a heading change)."""
    profile, start, goal = vc.corpus()[name]
    nodes, o = _oracle_for(oracle, profile)
    s, t = vc.lattice(profile, start), vc.lattice(profile, goal)
    w = oracle.v1_frontier_workspace(o, s, t)
    f = oracle.v1_frontier(o, s, t)
    lab, pre = [int(v) for v in w["labels"]], [int(v) for v in w["predecessors"]]
    assert (w["status"], w["rounds"]) == (f["status"], f["rounds"])
    orth, diag, bend = profile.costs
    start_state = next(st for st in range(len(lab)) if lab[st] == 0)
    assert start_state % 9 == 8 and pre[start_state] == 0xFFFFFFFF
    for st, (l, pr) in enumerate(zip(lab, pre)):
        if st == start_state or l == (1 << 64) - 1:
            assert pr == 0xFFFFFFFF
            continue
        h, ph = st % 9, pr % 9
        assert h != 8 and pr != 0xFFFFFFFF
        step = (diag if h % 2 else orth) + (bend if ph != 8 and ph != h else 0)
        assert lab[pr] + step == l, (st, pr)
    if f["status"] == 0:
        cur, walk = f["goal_state"], []
        while True:
            walk.append(cur)
            if cur == start_state:
                break
            cur = pre[cur]
        assert walk == f["states"]
        gold = GOLDEN[name]
        assert lab[f["goal_state"]] == gold.get("cuda_frontier_path_cost", gold.get("base_policy_scalar_cost"))


@pytest.mark.parametrize("name", sorted(vc.corpus()))
def test_frontier_restatement_costs_and_paths(oracle, name):
    """The frontier's costs are the reference's published optimal costs and its
    lowest-predecessor paths equal the exact-label (Dijkstra) oracle's: with a
    consistent heuristic every tight predecessor is closed before the goal."""
    profile, start, goal = vc.corpus()[name]
    _, o = _oracle_for(oracle, profile)
    s, g = vc.lattice(profile, start), vc.lattice(profile, goal)
    f = oracle.v1_frontier(o, s, g)
    cost, states = o.canonical(s, g)
    gold = GOLDEN[name]
    want = gold.get("cuda_frontier_path_cost", gold.get("base_policy_scalar_cost"))
    if name == "disconnected_fields":
        assert f["status"] == 1 and cost is None
    else:
        assert f["status"] == 0 and f["cost"] == cost == want and f["states"] == states
        assert f["rounds"] <= 9 * len(o.nodes)
