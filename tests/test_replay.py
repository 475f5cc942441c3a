"""Invariant replays: candidate batches (apgx_batch_replay_*) and single
routes (apgx_route_replay_*).

CPU: the parser against the reference's own sixteen checked-in artifacts
(tests/golden/reference_replays: checksums, grammar, the generator/fault
association rule, and the policy identity each pins), the canonical grammar's
rejections, and this library's artifacts up to the association gate.  GPU:
every artifact under tests/golden/replays reproduces its expected invariant.
"""
from pathlib import Path

import pytest

from apgar_amd import replay
from tests import v1_corpus as vc

GOLDEN = Path(__file__).parent / "golden"
REFERENCE = sorted((GOLDEN / "reference_replays").glob("*.replay"))
OURS = sorted((GOLDEN / "replays").glob("*.replay"))
ASSOC = replay.ASSOCIATION_ERROR


def resign(payload: bytes) -> bytes:
    """Payload lines + a fresh FNV-1a checksum trailer."""
    return payload + f"checksum_fnv1a64={vc.stable_hash_string(payload)}\n".encode()


def payload_of(text: bytes) -> bytes:
    return text[: text.rstrip(b"\n").rfind(b"\n") + 1]


def test_reference_artifacts_parse_and_pin_the_policy_identity(apgx):
    batch = [p for p in REFERENCE if p.name.startswith("batch_")]
    assert len(batch) == 14
    for p in batch:
        art, err = replay.parse(p.read_bytes())
        if "_wrong_generator_" in p.name:
            assert art is None and err == ASSOC, p.name
            continue
        assert art is not None, (p.name, err)
        # FingerprintCandidateGenerationPolicy restated in the library equals
        # the identity the reference wrote into its own artifact.
        assert replay.policy_identity(art) == art.expected_policy_identity, p.name
        assert art.fixture.decode() == "tests/fixtures/m1_exactness.kicad_pcb"
        assert art.expected_invariant.decode() in p.read_text()


def test_reference_artifacts_field_values(apgx):
    art, _ = replay.parse((GOLDEN / "reference_replays" / "batch_workspace_owner_v1.replay").read_bytes())
    assert (art.generator, art.fault, art.batch_id, art.query_id, art.input_ordinal) == (0, 2, 3102, 9001, 37)
    assert art.fixture_checksum == 15781378684595989985
    assert art.expected_device_view_fingerprint == 2764614999745307367
    assert art.expected_failure == 1 and art.maximum_rounds == 100000
    art, _ = replay.parse((GOLDEN / "reference_replays" / "batch_unauthenticated_cuda_producer_v1.replay").read_bytes())
    assert art.fault == 13 and art.expected_failure == 2


def test_single_route_artifacts_are_another_format(apgx):
    for p in REFERENCE:
        if p.name.startswith("goal_predecessor"):
            art, err = replay.parse(p.read_bytes())
            assert art is None and err == "replay fields are missing or out of canonical order"


def test_reference_fixture_is_out_of_scope(apgx):
    """The reference's fixtures are KiCad boards; this library starts at the
    compiled board, so such an artifact stops at the fixture stage."""
    o = replay.run(GOLDEN / "reference_replays" / "batch_workspace_owner_v1.replay", dry_run=True)
    assert (o.exit_code, o.stage_name) == (2, "fixture")
    assert o.message.decode() == "fixture is missing or its checksum differs"


BASE = (GOLDEN / "reference_replays" / "batch_query_identity_v1.replay").read_bytes()


@pytest.mark.parametrize("mutate, message", [
    (lambda t: t.replace(b"layer=0\n", b"layer=00\n"), "replay contains an invalid integer field"),
    (lambda t: t.replace(b"layer=0\n", b"layer=+0\n"), "replay contains an invalid integer field"),
    (lambda t: t.replace(b"layer=0\n", b"layer=4294967296\n"), "replay contains an invalid integer field"),
    (lambda t: t.replace(b"layer=0\n", b"layer=4294967295\n"), None),
    (lambda t: t.replace(b"batch_id=3106\n", b"batch_id=0\n"), "replay contains an invalid integer field"),
    (lambda t: t.replace(b"maximum_rounds=100000\n", b"maximum_rounds=0\n"), "replay contains an invalid integer field"),
    (lambda t: t.replace(b"layer=0\n", b"layer=\n"), "replay contains an invalid integer field"),
    (lambda t: t.replace(b"generator=bucketed_frontier\n", b"generator=heading_aware_sweep\n"), ASSOC),
    (lambda t: t.replace(b"objective=base_scalar_cost\n", b"objective=length_biased\n"), ASSOC),
    (lambda t: t.replace(b"banned_resources=none\n", b"banned_resources=\n"), ASSOC),
    (lambda t: t.replace(b"expected_failure=internal_invariant\n", b"expected_failure=unsupported\n"), ASSOC),
    (lambda t: t.replace(b"layer=0\nmaximum_rounds", b"maximum_rounds"), "replay fields are missing or out of canonical order"),
    (lambda t: t + b"extra=1\n", "replay payload has unknown fields"),
    (lambda t: t[: t.find(b"expected_invariant=")], "replay payload is truncated"),
])
def test_grammar_rejections(apgx, mutate, message):
    art, err = replay.parse(resign(mutate(payload_of(BASE))))
    if message is None:
        assert art is not None and art.layer == 4294967295
    else:
        assert art is None and err == message


@pytest.mark.parametrize("text, message", [
    (BASE.replace(b"query_id=9001", b"query_id=9002"), "replay payload checksum mismatch"),
    (BASE[:-1], "canonical replay must end every line with LF"),
    (BASE.replace(b"\n", b"\r\n"), "canonical replay must not contain NUL or CR bytes"),
    (BASE.replace(b"fault=", b"fa\x00ult="), "canonical replay must not contain NUL or CR bytes"),
    (BASE.replace(b"fault=", b"fa\xffult="), "canonical replay must be valid UTF-8"),
    (BASE.replace(b"fault=", b"fa\xc0\xafult="), "canonical replay must be valid UTF-8"),
    (payload_of(BASE), "missing final replay checksum"),
    (b"", "canonical replay must end every line with LF"),
])
def test_envelope_rejections(apgx, text, message):
    art, err = replay.parse(text)
    assert art is None and err == message


def test_valid_utf8_text_fields_are_accepted_by_the_envelope(apgx):
    # A multi-byte character in a text field passes the envelope and fails
    # only the semantic contract.
    art, err = replay.parse(resign(payload_of(BASE).replace(b"fault=query_identity", "fault=query_identité".encode())))
    assert art is None and err == ASSOC


def test_committed_artifacts_are_current(apgx):
    """tests/golden/make_replays.py regenerates the committed files byte for byte."""
    from tests.golden import make_replays as mk

    data, header, routing_fp = mk.fixture()
    assert (GOLDEN / "replays" / "high_turn_maze_v1.fixture").read_bytes() == data
    assert mk.policy_identity() == 10775560220227872104  # equals the reference artifacts' policy
    fx_sum = vc.stable_hash_string(data)
    for k, (fault, (generator, invariant)) in enumerate(mk.CONTRACT.items()):
        name = fault if fault.startswith("batch_") else f"batch_{fault}"
        want = mk.artifact("tests/golden/replays/high_turn_maze_v1.fixture", fx_sum, generator, fault, invariant, k,
                           header, routing_fp, mk.policy_identity())
        assert (GOLDEN / "replays" / f"{name}_v1.replay").read_bytes() == want


def test_our_artifacts_pass_the_association_gate(apgx):
    todo, parse_rejected, gate_rejected = replay.default_artifacts()
    assert len(todo) == 13 and len(parse_rejected) == 1 and len(gate_rejected) == 2
    for p in todo:
        o = replay.run(p, dry_run=True)
        assert (o.exit_code, o.stage_name) == (0, "association"), (p.name, o.message)
    assert replay.reject(parse_rejected[0]) == 0
    for p in gate_rejected:
        assert replay.reject_at_gate(p) == 0


# ---------------------------------------------------------- gpu_replay v1 (single route)
ROUTE_BASE = (GOLDEN / "reference_replays" / "goal_predecessor_self_cycle_v1.replay").read_bytes()
ROUTE_ASSOC = "replay format, schema, generator, fault, or expected outcome is unsupported"


def test_reference_route_artifacts_parse(apgx):
    """The reference's two gpu_replay artifacts: every field as it wrote it."""
    art, err = replay.parse_route(ROUTE_BASE)
    assert art is not None, err
    assert art.fixture.decode() == "tests/fixtures/m1_exactness.kicad_pcb"
    assert art.fixture_checksum == 15781378684595989985
    assert (art.generator, art.fault, art.layer, art.maximum_rounds) == (0, 1, 0, 100000)
    assert art.expected_board_content_hash == 1860510029311497781
    assert art.expected_profile_fingerprint == 8851933838860818274
    assert art.expected_routing_profile_fingerprint == 1820138087830637135
    assert art.expected_device_view_fingerprint == 2764614999745307367
    assert art.expected_invariant.decode() == "gpu.predecessor.self_reference.v1"
    wrong, err = replay.parse_route((GOLDEN / "reference_replays" /
                                     "goal_predecessor_self_cycle_wrong_routing_profile_v1.replay").read_bytes())
    assert wrong is not None, err
    assert wrong.expected_routing_profile_fingerprint != art.expected_routing_profile_fingerprint
    # Batch artifacts are another format for this parser.
    batch, err = replay.parse_route((GOLDEN / "reference_replays" / "batch_workspace_owner_v1.replay").read_bytes())
    assert batch is None and err == "replay fields are missing or out of canonical order"


@pytest.mark.parametrize("mutate, message", [
    (lambda t: t.replace(b"generator=bucketed_frontier\n", b"generator=heading_aware_sweep\n"), ROUTE_ASSOC),
    (lambda t: t.replace(b"fault=goal_predecessor_self_cycle\n", b"fault=none\n"), ROUTE_ASSOC),
    (lambda t: t.replace(b"schema_version=1\n", b"schema_version=2\n"), ROUTE_ASSOC),
    (lambda t: t.replace(b"expected_failure=internal_invariant\n", b"expected_failure=unsupported\n"), ROUTE_ASSOC),
    (lambda t: t.replace(b"maximum_rounds=100000\n", b"maximum_rounds=0\n"), "replay contains an invalid integer field"),
    (lambda t: t.replace(b"layer=0\n", b"layer=01\n"), "replay contains an invalid integer field"),
    (lambda t: t.replace(b"layer=0\n", b"layer=4294967296\n"), "replay contains an invalid integer field"),
    (lambda t: t.replace(b"layer=0\nmaximum_rounds", b"maximum_rounds"), "replay fields are missing or out of canonical order"),
    (lambda t: t + b"extra=1\n", "replay payload has unknown fields"),
])
def test_route_grammar_rejections(apgx, mutate, message):
    art, err = replay.parse_route(resign(mutate(payload_of(ROUTE_BASE))))
    assert art is None and err == message


def test_route_checksum_and_fixture(apgx):
    art, err = replay.parse_route(ROUTE_BASE.replace(b"layer=0", b"layer=1"))
    assert art is None and err == "replay payload checksum mismatch"
    o = replay.run(GOLDEN / "reference_replays" / "goal_predecessor_self_cycle_v1.replay", dry_run=True)
    assert (o.exit_code, o.stage_name) == (2, "fixture")
    assert o.message.decode() == "fixture is missing or its checksum differs"


def test_route_association_gate_names_each_mismatch(apgx):
    text = (GOLDEN / "replays" / "goal_predecessor_self_cycle_v1.replay").read_bytes()
    fixture = (GOLDEN / "replays" / "high_turn_maze_v1.fixture").read_bytes()
    o = replay.run_bytes(text, fixture, dry_run=True)
    assert (o.exit_code, o.stage_name) == (0, "association")
    art, _ = replay.parse_route(text)
    for field, name in (("expected_board_content_hash", "board"), ("expected_profile_fingerprint", "compiler_profile"),
                        ("expected_routing_profile_fingerprint", "routing_profile"),
                        ("expected_device_view_fingerprint", "device_view")):
        value = getattr(art, field)
        drifted = payload_of(text).replace(f"{field}={value}\n".encode(), f"{field}={value ^ 1}\n".encode())
        o = replay.run_bytes(resign(drifted), fixture, dry_run=True)
        assert (o.exit_code, o.stage_name) == (2, "association")
        assert o.message.decode() == f"replay semantic association mismatch: {name}"


def test_route_artifacts_are_current(apgx):
    from tests.golden import make_replays as mk

    data, header, routing_fp = mk.fixture()
    want = mk.route_artifact("tests/golden/replays/high_turn_maze_v1.fixture", vc.stable_hash_string(data), header,
                             routing_fp)
    assert (GOLDEN / "replays" / "goal_predecessor_self_cycle_v1.replay").read_bytes() == want


def test_fixture_tampering_is_rejected(apgx):
    text = (GOLDEN / "replays" / "batch_query_identity_v1.replay").read_bytes()
    fixture = bytearray((GOLDEN / "replays" / "high_turn_maze_v1.fixture").read_bytes())
    fixture[-20] ^= 1
    o = replay.run_bytes(text, bytes(fixture), dry_run=True)
    assert (o.exit_code, o.stage_name) == (2, "fixture")
    # Same tampering with the artifact re-pinned to the new fixture bytes: the
    # board changes, so the association gate rejects it before any GPU work.
    art, _ = replay.parse(text)
    repinned = payload_of(text).replace(f"fixture_checksum_fnv1a64={art.fixture_checksum}".encode(),
                                        f"fixture_checksum_fnv1a64={vc.stable_hash_string(bytes(fixture))}".encode())
    o = replay.run_bytes(resign(repinned), bytes(fixture), dry_run=True)
    assert o.exit_code == 2 and o.stage_name in ("fixture", "association"), o.message


@pytest.mark.gpu
def test_every_artifact_reproduces_on_the_gpu(apgx, gpu):
    print("[case] candidate-batch replays", flush=True)
    todo, _, _ = replay.default_artifacts()
    for p in todo:
        text = p.read_bytes()
        art, _ = replay.parse_route(text) if replay.is_route_artifact(text) else replay.parse(text)
        print(f"[case] replay {p.name}", flush=True)
        o = replay.run(p)
        assert o.exit_code == 0 and o.stage_name == "reproduce", (p.name, o.message)
        assert o.invariant == art.expected_invariant
        assert o.message.decode().startswith("reproduced=internal_invariant invariant=")
    assert replay.main([]) == 0
