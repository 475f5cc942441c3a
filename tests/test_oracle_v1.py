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
    text = (Path(__file__).parent / "golden" / "replay_batch_workspace_owner_v1.txt").read_bytes()
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
