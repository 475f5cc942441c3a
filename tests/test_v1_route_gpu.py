"""GPU parity of the single-route path (apgx_v1_route_execute / _readback).

Both generators on every corpus case against the oracle's restatement of the
reference's single-route generators: every state's label and predecessor, the
goal state, rounds, examined work and turn relaxations bit-exact; rounds and
examined work also against the reference's published figures
(tests/golden/corpus_golden.json) and the path cost against its published
optimal cost; every readback passes the host trust boundary.  Then the round
budget, cancellation, the replay decorator, repeated routes on one view, the
deterministic byte accounting and the reference's request errors.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from apgar_amd import v1
from tests import route_cases as rc
from tests import v1_corpus as vc

pytestmark = pytest.mark.gpu
GOLDEN = json.loads((Path(__file__).parent / "golden" / "corpus_golden.json").read_text())
INF = 0xFFFFFFFFFFFFFFFF


def _check_equal(rb, r, b, generator):
    res = rb.result
    assert res.completion == rc.COMPLETION[r["completion"]]
    assert np.array_equal(rb.labels, r["labels"])
    assert np.array_equal(rb.predecessors, r["predecessors"])
    t = res.telemetry
    assert (t.rounds, t.examined_work, t.heading_turn_relaxations) == (r["rounds"], r["examined_work"], r["turns"])
    if r["completion"] == 0:
        assert res.goal_state == r["goal_state"]
    else:
        assert res.goal_state == v1.INVALID
    assert res.header.device_view_fingerprint == b.header.device_view_fingerprint
    assert res.header.generator == generator and res.header.schema_version == 1
    assert t.batch_device_bytes == v1.expected_route_bytes(generator, len(b.nodes))
    assert t.peak_device_bytes == t.persistent_device_bytes + t.batch_device_bytes
    assert t.persistent_device_bytes == b.header.estimated_persistent_device_bytes and t.kernel_milliseconds > 0


@pytest.mark.parametrize("name", sorted(vc.corpus()))
@pytest.mark.parametrize("generator", [0, 1])
def test_corpus_routes(apgx, oracle, gpu, name, generator):
    print(f"[case] v1 single route {name} generator={generator}", flush=True)
    profile, b, o, s_xy, g_xy, s, g = rc.corpus_case(oracle, name)
    view = v1.View(b)
    rb = view.route(s, g, generator=generator, routing_profile_fingerprint=77)
    view.close()
    r = oracle.v1_route(o, s_xy, g_xy, generator)
    _check_equal(rb, r, b, generator)
    gold = GOLDEN[name]
    key = "cuda_frontier" if generator == 0 else "cuda_sweep"
    if f"{key}_rounds" in gold:
        assert rb.result.telemetry.rounds == gold[f"{key}_rounds"]
        assert rb.result.telemetry.examined_work == gold[f"{key}_examined_work"]
    req = v1.route_request_for(s, g, generator=generator, routing_profile_fingerprint=77)
    valid, inv, detail, cost, _ = v1.validate_route(b, req, rb)
    assert valid, (inv, detail)
    if name == "disconnected_fields":
        assert rb.result.completion == v1.DISCONNECTED
    else:
        assert cost == gold.get(f"{key}_path_cost", gold.get("base_policy_scalar_cost"))


def test_corridor_and_budget(apgx, oracle, gpu):
    print("[case] v1 single route corridor, budgets", flush=True)
    b, o = rc.corridor(oracle)
    view = v1.View(b)
    for gen in (0, 1):
        rb = view.route(0, 1, generator=gen)
        _check_equal(rb, oracle.v1_route(o, (0, 0), (1, 0), gen), b, gen)
        rb = view.route(0, 1, generator=gen, maximum_rounds=1)
        assert rb.result.completion == v1.BUDGET_EXHAUSTED and rb.result.telemetry.rounds == 1
        _check_equal(rb, oracle.v1_route(o, (0, 0), (1, 0), gen, max_rounds=1), b, gen)
        req = v1.route_request_for(0, 1, generator=gen, maximum_rounds=1)
        assert v1.validate_route(b, req, rb)[0]
    view.close()
    profile, b, o, s_xy, g_xy, s, g = rc.corpus_case(oracle, "high_turn_maze")
    view = v1.View(b)
    for gen, rounds in ((0, 5), (1, 2)):
        rb = view.route(s, g, generator=gen, maximum_rounds=rounds)
        _check_equal(rb, oracle.v1_route(o, s_xy, g_xy, gen, max_rounds=rounds), b, gen)
    view.close()


def test_cancellation_and_repeat(apgx, oracle, gpu):
    print("[case] v1 single route cancellation and repeated routes", flush=True)
    profile, b, o, s_xy, g_xy, s, g = rc.corpus_case(oracle, "dense_corridors")
    view = v1.View(b)
    for gen in (0, 1):
        rb = view.route(s, g, generator=gen, cancel=True)
        assert rb.result.completion == v1.CANCELLED and rb.result.telemetry.rounds == 0
        req = v1.route_request_for(s, g, generator=gen)
        assert v1.validate_route(b, req, rb)[0]
        # A clear flag is sampled, does not cancel, and routes equal the plain ones.
        flag = (v1.U32 * 1)(0)
        a = view.route(s, g, generator=gen, cancel=flag)
        c = view.route(s, g, generator=gen)
        assert a.result.completion == c.result.completion == v1.REACHED
        assert np.array_equal(a.labels, c.labels) and np.array_equal(a.predecessors, c.predecessors)
    view.close()


def test_goal_predecessor_fault_is_caught(apgx, oracle, gpu):
    print("[case] v1 single route goal-predecessor decorator", flush=True)
    profile, b, o, s_xy, g_xy, s, g = rc.corpus_case(oracle, "high_turn_maze")
    view = v1.View(b)
    rb = view.route(s, g, fault=v1.ROUTE_FAULT_GOAL_PREDECESSOR_SELF_CYCLE)
    view.close()
    assert rb.predecessors[rb.result.goal_state] == rb.result.goal_state
    valid, inv, detail, _, cls = v1.validate_route(b, v1.route_request_for(s, g), rb)
    assert not valid and cls == v1.CHECK_INTERNAL_INVARIANT and inv == "gpu.predecessor.self_reference.v1"


def test_request_errors(apgx, oracle, gpu):
    print("[case] v1 single route request errors", flush=True)
    profile, b, o, s_xy, g_xy, s, g = rc.corpus_case(oracle, "cross_tile_edges")
    view = v1.View(b)
    for kw, code in (({"start": s, "goal": s}, 4), ({"start": s, "goal": len(b.nodes)}, 4),
                     ({"start": s, "goal": g, "maximum_rounds": 0}, 4), ({"start": s, "goal": g, "generator": 5}, 4),
                     ({"start": s, "goal": g, "maximum_device_bytes": b.header.estimated_persistent_device_bytes}, 2)):
        with pytest.raises(apgx.ApgxError) as err:
            view.route(kw.pop("start"), kw.pop("goal"), **kw)
        assert err.value.code == code, err.value.detail
    # The exact byte budget is enough.
    need = b.header.estimated_persistent_device_bytes + v1.expected_route_bytes(0, len(b.nodes))
    assert view.route(s, g, maximum_device_bytes=need).result.completion == v1.REACHED
    view.close()
