"""GPU parity of the V1 generators on the reference corpus.

Sweep: every query's reverse compact path against the oracle's
lowest-predecessor path, the recomputed path cost against the reference's
published optimal cost, and the round count and examined work against the
reference's published sweep telemetry.  Frontier: the full-workspace readback
(every state's label and predecessor) bit-exact against the frontier
restatement, plus rounds, examined work, goal state and the predecessor walk.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from apgar_amd import v1
from tests import v1_corpus as vc
from tests.test_v1_board import board_for

pytestmark = pytest.mark.gpu
GOLDEN = json.loads((Path(__file__).parent / "golden" / "corpus_golden.json").read_text())


def path_cost(board, states, costs, surcharges=(0, 0, 0), adjustments=None):
    """Cost of a reverse state path under the V1 transition model."""
    adjustments = adjustments or {}
    orth, diag, bend = costs
    total = 0
    seq = list(reversed(states))  # start .. goal
    for p, s in zip(seq[:-1], seq[1:]):
        pn, ph = divmod(p, 9)
        sn, sh = divmod(s, 9)
        assert board.nodes[pn].neighbors[sh] == sn
        w = (diag + surcharges[1]) if sh % 2 else (orth + surcharges[0])
        if ph != 8 and ph != sh:
            w += bend + surcharges[2]
        if sh < 4:
            edge = pn * 8 + sh
        else:
            edge = board.nodes[pn].neighbors[sh] * 8 + sh - 4
        adj = adjustments.get(edge, 0)
        assert adj != v1.BANNED
        total += w + adj
    return total


@pytest.mark.parametrize("name", sorted(vc.corpus()))
def test_corpus_case(apgx, oracle, gpu, name):
    print(f"[case] v1 corpus {name}", flush=True)
    profile, start, goal = vc.corpus()[name]
    nodes, b = board_for(profile)
    s_xy, g_xy = vc.lattice(profile, start), vc.lattice(profile, goal)
    s, g = b.node_index(*s_xy), b.node_index(*g_xy)
    view = v1.View(b)
    results, paths, states, tel, _ = view.execute([v1.make_query(0, s, g, b)])
    r, p = results[0], paths[0]
    gold = GOLDEN[name]
    o = oracle.V1(nodes, profile.costs, profile.heading, profile.tile_w, profile.tile_h)
    ocost, ostates = o.canonical(s_xy, g_xy)
    if name == "disconnected_fields":
        assert r.completion == v1.DISCONNECTED and p.state_count == 0 and ocost is None
    else:
        assert r.completion == v1.REACHED
        got = states[p.state_offset:p.state_offset + p.state_count]
        assert got == ostates
        want = gold.get("cuda_sweep_path_cost", gold.get("base_policy_scalar_cost"))
        assert path_cost(b, got, profile.costs) == want == ocost
        assert r.goal_state == got[0]
    if "cuda_sweep_rounds" in gold:
        assert r.rounds == gold["cuda_sweep_rounds"]
        assert r.examined_work == gold["cuda_sweep_examined_work"]
    assert r.query_id == 1 and r.workspace_owner == 1
    assert r.device_view_fingerprint == b.header.device_view_fingerprint
    view.close()


def test_policy_known_answers(apgx, gpu):
    print("[case] v1 policy known answers", flush=True)
    profile, start, goal = vc.known_answer_boards()["horizontal_detour"]
    _, b = board_for(profile)
    s, g = b.node_index(*vc.lattice(profile, start)), b.node_index(*vc.lattice(profile, goal))
    middle = v1.physical_edge_index(b, 4, 0, 0)
    view = v1.View(b)
    edges = [v1.PolicyEdge(middle, 0, v1.BANNED), v1.PolicyEdge(middle, 0, 200)]
    qs = [v1.make_query(0, s, g, b),
          v1.make_query(1, s, g, b, policy_offset=0, policy_count=1),
          v1.make_query(2, s, g, b, policy_offset=1, policy_count=1)]
    results, paths, states, _, _ = view.execute(qs, edges)
    costs = []
    for i, (r, p) in enumerate(zip(results, paths)):
        assert r.completion == v1.REACHED
        got = states[p.state_offset:p.state_offset + p.state_count]
        adj = {middle: edges[i - 1].adjustment} if i else {}
        costs.append(path_cost(b, got, profile.costs, adjustments=adj))
    assert costs == [100, 126, 126]
    view.close()

    profile, start, goal = vc.known_answer_boards()["horizontal_corridor"]
    _, b = board_for(profile)
    s, g = b.node_index(*vc.lattice(profile, start)), b.node_index(*vc.lattice(profile, goal))
    middle = v1.physical_edge_index(b, 4, 0, 0)
    view = v1.View(b)
    edges = [v1.PolicyEdge(middle, 0, 17), v1.PolicyEdge(middle, 0, v1.BANNED)]
    qs = [v1.make_query(0, s, g, b, policy_count=1, surcharges=(2, 3, 5)),
          v1.make_query(1, g, s, b, policy_count=1, surcharges=(2, 3, 5)),
          v1.make_query(2, s, g, b, policy_offset=1, policy_count=1)]
    results, paths, states, _, _ = view.execute(qs, edges)
    for i in (0, 1):
        p = paths[i]
        got = states[p.state_offset:p.state_offset + p.state_count]
        assert path_cost(b, got, profile.costs, (2, 3, 5), {middle: 17}) == 137
    assert results[2].completion == v1.DISCONNECTED
    view.close()


def test_cheap_diagonal_cost(apgx, oracle, gpu):
    profile, start, goal = vc.known_answer_boards()["cheap_diagonal"]
    nodes, b = board_for(profile)
    s, g = b.node_index(*vc.lattice(profile, start)), b.node_index(*vc.lattice(profile, goal))
    view = v1.View(b)
    results, paths, states, _, _ = view.execute([v1.make_query(0, s, g, b)])
    p = paths[0]
    got = states[p.state_offset:p.state_offset + p.state_count]
    assert path_cost(b, got, profile.costs) == 37
    o = oracle.V1(nodes, profile.costs, profile.heading, profile.tile_w, profile.tile_h)
    assert o.canonical(vc.lattice(profile, start), vc.lattice(profile, goal))[1] == got
    view.close()


def test_batch_of_random_queries(apgx, oracle, gpu):
    print("[case] v1 512-query batch on negative_coordinates", flush=True)
    profile, _, _ = vc.corpus()["negative_coordinates"]
    nodes, b = board_for(profile)
    rng = np.random.default_rng(3)
    n = len(b.nodes)
    pairs = [tuple(rng.choice(n, size=2, replace=False)) for _ in range(512)]
    qs = [v1.make_query(i, int(s), int(g), b) for i, (s, g) in enumerate(pairs)]
    view = v1.View(b)
    results, paths, states, tel, _ = view.execute(qs)
    o = oracle.V1(nodes, profile.costs, profile.heading, profile.tile_w, profile.tile_h)
    for i in range(0, 512, 37):
        s, g = pairs[i]
        sx = (b.nodes[s].lattice_x, b.nodes[s].lattice_y)
        gx = (b.nodes[g].lattice_x, b.nodes[g].lattice_y)
        cost, ost = o.canonical(sx, gx)
        p = paths[i]
        got = states[p.state_offset:p.state_offset + p.state_count]
        assert got == ost and path_cost(b, got, profile.costs) == cost
    assert tel.kernel_milliseconds > 0 and tel.chunk_rounds == 8
    view.close()


def test_cancel_and_budget(apgx, gpu):
    profile, start, goal = vc.corpus()["high_turn_maze"]
    _, b = board_for(profile)
    s, g = b.node_index(*vc.lattice(profile, start)), b.node_index(*vc.lattice(profile, goal))
    view = v1.View(b)
    r, p, _, _, _ = view.execute([v1.make_query(0, s, g, b)], cancel=True)
    assert r[0].completion == v1.CANCELLED and p[0].state_count == 0
    r, p, _, tel, _ = view.execute([v1.make_query(0, s, g, b, maximum_rounds=2)], maximum_rounds=2)
    assert r[0].completion == v1.BUDGET_EXHAUSTED and r[0].rounds == 2
    assert tel.finalization_launch_count == 1
    view.close()


def _frontier_check(b, o, s_xy, g_xy, rb, qi, policy=None, max_rounds=100_000):
    """One frontier query of a full-workspace readback against the
    restatement: rounds, examined work, goal state, every state's label and
    predecessor (bit-exact), ownership stamps, and the predecessor walk."""
    from oracle import py as oracle_py

    f = oracle_py.v1_frontier(o, s_xy, g_xy, policy, max_rounds=max_rounds)
    w = oracle_py.v1_frontier_workspace(o, s_xy, g_xy, policy, max_rounds=max_rounds)
    r, ws = rb.results[qi], rb.workspace
    assert ws.readback_kind == v1.READBACK_FULL_WORKSPACE and rb.states == []
    env = ws.envelopes[qi]
    lo, n = env.workspace_offset, env.workspace_state_count
    assert (env.rounds, env.examined_work) == (r.rounds, r.examined_work)
    assert r.rounds == f["rounds"] and r.examined_work == f["examined_work"]
    assert r.heading_turn_relaxations == 0
    assert np.array_equal(ws.labels[lo:lo + n], w["labels"])
    assert np.array_equal(ws.predecessors[lo:lo + n], w["predecessors"])
    assert (ws.state_owners[lo:lo + n] == r.workspace_owner).all()
    assert (ws.predecessor_owners[lo:lo + n] == r.workspace_owner).all()
    if f["status"] == 2:
        assert r.completion == v1.BUDGET_EXHAUSTED
        return None
    if f["status"] == 1:
        assert r.completion == v1.DISCONNECTED
        return None
    assert r.completion == v1.REACHED and r.goal_state == f["goal_state"]
    got = v1.reconstruct(ws, qi, r)
    assert got == f["states"]
    return got


@pytest.mark.parametrize("name", sorted(vc.corpus()))
def test_frontier_corpus_case(apgx, oracle, gpu, name):
    """Bucketed-frontier generator: rounds, examined work, goal state and path
    bit-exact against the frontier restatement; costs against the published
    optimal costs; results identical to the sweep's."""
    print(f"[case] v1 frontier corpus {name}", flush=True)
    profile, start, goal = vc.corpus()[name]
    nodes, b = board_for(profile)
    s_xy, g_xy = vc.lattice(profile, start), vc.lattice(profile, goal)
    s, g = b.node_index(*s_xy), b.node_index(*g_xy)
    view = v1.View(b)
    q = v1.make_query(0, s, g, b, generator=v1.GENERATOR_FRONTIER)
    rb = view.execute([q], generator=v1.GENERATOR_FRONTIER)
    results, tel = rb.results, rb.telemetry
    o = oracle.V1(nodes, profile.costs, profile.heading, profile.tile_w, profile.tile_h)
    got = _frontier_check(b, o, s_xy, g_xy, rb, 0)
    sres, spaths, sstates, _, _ = view.execute([v1.make_query(0, s, g, b)])
    if got is not None:
        gold = GOLDEN[name]
        want = gold.get("cuda_frontier_path_cost", gold.get("base_policy_scalar_cost"))
        assert path_cost(b, got, profile.costs) == want
        sp = spaths[0]
        assert got == sstates[sp.state_offset:sp.state_offset + sp.state_count]
        assert results[0].goal_state == sres[0].goal_state
    assert tel.chunk_rounds == 32
    view.close()


def test_frontier_batch_with_policies(apgx, oracle, gpu):
    """64 random queries with surcharges, penalties and a ban: every query's
    telemetry and path against the restatement, costs equal to the sweep's."""
    print("[case] v1 frontier batch with policies", flush=True)
    profile, _, _ = vc.corpus()["dense_corridors"]
    nodes, b = board_for(profile)
    rng = np.random.default_rng(8)
    n = len(b.nodes)
    pen_edges = sorted({int(rng.integers(0, n)) * 8 + int(rng.integers(0, 4)) for _ in range(24)})
    pen_edges = [e for e in pen_edges if b.nodes[e // 8].neighbors[e % 8] != 0xFFFFFFFF]
    edges = [v1.PolicyEdge(e, 0, int(rng.integers(1, 60))) for e in pen_edges]
    edges[3].adjustment = v1.BANNED
    qs_f, qs_s, pairs = [], [], []
    for i in range(64):
        s, g = (int(v) for v in rng.choice(n, size=2, replace=False))
        sur = tuple(int(v) for v in rng.integers(0, 5, 3))
        pc = len(edges) if i % 2 else 0
        pairs.append((s, g, sur, pc))
        qs_f.append(v1.make_query(i, s, g, b, policy_count=pc, surcharges=sur, generator=v1.GENERATOR_FRONTIER))
        qs_s.append(v1.make_query(i, s, g, b, policy_count=pc, surcharges=sur))
    view = v1.View(b)
    frb = view.execute(qs_f, edges, generator=v1.GENERATOR_FRONTIER)
    fr = frb.results
    sr, sp, ss, _, _ = view.execute(qs_s, edges)
    o = oracle.V1(nodes, profile.costs, profile.heading, profile.tile_w, profile.tile_h)
    for i, (s, g, sur, pc) in enumerate(pairs):
        s_xy = (b.nodes[s].lattice_x, b.nodes[s].lattice_y)
        g_xy = (b.nodes[g].lattice_x, b.nodes[g].lattice_y)
        bans, pens = [], []
        for e in edges[:pc]:
            node, d = divmod(e.physical_edge_index, 8)
            key = (b.nodes[node].lattice_x, b.nodes[node].lattice_y, d)
            (bans.append(key) if e.adjustment == v1.BANNED else pens.append((key, e.adjustment)))
        pol = oracle.V1.policy(*sur, bans=bans, penalties=pens)
        got = _frontier_check(b, o, s_xy, g_xy, frb, i, pol)
        assert fr[i].completion == sr[i].completion
        if got is not None:
            sw = ss[sp[i].state_offset:sp[i].state_offset + sp[i].state_count]
            adj = {e.physical_edge_index: e.adjustment for e in edges[:pc]}
            assert path_cost(b, got, profile.costs, sur, adj) == path_cost(b, sw, profile.costs, sur, adj)
    view.close()


def test_frontier_budget(apgx, oracle, gpu):
    """A 40-round budget: the partial workspace (labels of open states are
    upper bounds) equals the restatement's after the same 40 rounds."""
    profile, start, goal = vc.corpus()["high_turn_maze"]
    nodes, b = board_for(profile)
    s, g = b.node_index(*vc.lattice(profile, start)), b.node_index(*vc.lattice(profile, goal))
    view = v1.View(b)
    q = v1.make_query(0, s, g, b, maximum_rounds=40, generator=v1.GENERATOR_FRONTIER)
    rb = view.execute([q], maximum_rounds=40, generator=v1.GENERATOR_FRONTIER)
    r, tel = rb.results, rb.telemetry
    assert r[0].completion == v1.BUDGET_EXHAUSTED and r[0].rounds == 40
    assert tel.finalization_launch_count == 1 and tel.dispatched_rounds == 40
    assert tel.kernel_launch_count == 2 + 2 + 1  # init + select, two chunks, finalization
    o = oracle.V1(nodes, profile.costs, profile.heading, profile.tile_w, profile.tile_h)
    _frontier_check(b, o, vc.lattice(profile, start), vc.lattice(profile, goal), rb, 0, max_rounds=40)
    view.close()


@pytest.mark.parametrize("generator", [v1.GENERATOR_SWEEP, v1.GENERATOR_FRONTIER])
def test_host_byte_budget(apgx, gpu, generator):
    """A batch whose worst-case host bytes exceed maximum_host_bytes is refused
    before execution; at exactly the estimate it runs.  Estimate: per query a
    result, a query and (sweep) a compact path header, plus 4 B (sweep, every
    state on a path at worst) or 28 B (frontier workspace) per state."""
    profile, start, goal = vc.corpus()["dense_corridors"]
    _, b = board_for(profile)
    s, g = b.node_index(*vc.lattice(profile, start)), b.node_index(*vc.lattice(profile, goal))
    qs = [v1.make_query(i, s, g, b, generator=generator) for i in range(3)]
    S = len(b.nodes) * 9
    per_state = 28 if generator == v1.GENERATOR_FRONTIER else 4
    est = 3 * (120 + 120 + (0 if generator == v1.GENERATOR_FRONTIER else 32)) + 3 * S * per_state
    view = v1.View(b)
    with pytest.raises(apgx.ApgxError) as e:
        view.execute(qs, generator=generator, maximum_host_bytes=est - 1)
    assert "maximum_host_bytes" in e.value.detail
    rb = view.execute(qs, generator=generator, maximum_host_bytes=est)
    assert all(r.completion == v1.REACHED for r in rb.results)
    view.close()
