"""Writes the replay artifacts under tests/golden/replays/.

One fixture -- the compiled high_turn_maze corpus case (profile scalars,
request endpoints, represented nodes with legal masks; see
apgar_amd/replay.py) -- and, per fault contract of
schemas/gpu_batch_replay/v1.md that this library can reproduce, one canonical
artifact naming it, plus two negatives: a checksum-valid artifact that pairs a
compact-path fault with the frontier (rejected by the parser's association
rule) and one whose routing-profile fingerprint differs from the fixture's
(rejected by the association gate before any GPU work).  The single-route
format (gpu_replay v1) gets its one contract, the frontier's goal-predecessor
self cycle, and the same routing-profile negative.

The corpus case has no Board IR, so its board content hash, rule-bucket
identity and routed-net profile (net 1, generation 1, width 200000, clearance
150000, layer 0, the case's heading mask) are declared here; the compiler
profile fingerprint is the case's own (tests/v1_corpus.py) and the device view
fingerprint is computed by apgx_v1_build_board.  Run from the repository root:
    python -m tests.golden.make_replays
"""
from __future__ import annotations

from pathlib import Path

from apgar_amd import replay, v1
from tests import v1_corpus as vc
from tests.v1_faults import CONTRACT

OUT = Path(__file__).resolve().parent / "replays"
CASE = "high_turn_maze"
SEED, ORDINAL = 1311768467463790320, 11


def _hash_string(*parts):
    h = vc.Fnv1a()
    for p in parts:
        h.string(p)
    return h.v


def routing_profile_fingerprint(heading_mask):
    """FingerprintRoutingProfile (src/routing/candidate_policy.cc:162-175)."""
    h = vc.Fnv1a()
    h.string("APGAR-ROUTING-PROFILE-V1")
    h.u64(1)          # net id
    h.u32(1)          # net generation
    h.i64(200000)     # nominal width
    h.i64(150000)     # clearance
    h.u64(1)          # one allowed layer
    h.u32(0)
    h.byte(heading_mask)
    return h.v


def fixture():
    profile, start, goal = vc.corpus()[CASE]
    nodes = vc.compile_obstacle_free(profile)
    prof = v1.CompiledProfile(1, profile.heading, _hash_string("APGX-REPLAY-CORPUS-CASE-V1", CASE),
                              profile.fingerprint(), _hash_string("APGX-REPLAY-RULE-BUCKET-V1", CASE),
                              0, 0, profile.step, profile.tile_w, profile.tile_h, *profile.costs, 0)
    data = replay.fixture_bytes(prof, nodes, vc.lattice(profile, start), vc.lattice(profile, goal),
                                routing_profile_fingerprint=routing_profile_fingerprint(profile.heading))
    board = v1.build_board(nodes, tile_w=profile.tile_w, tile_h=profile.tile_h, costs=profile.costs,
                           heading_mask=profile.heading, step=profile.step, source_hash=prof.source_board_content_hash,
                           profile_fingerprint=prof.compiler_profile_fingerprint,
                           rule_bucket=prof.rule_bucket_identity)
    return data, board.header, routing_profile_fingerprint(profile.heading)


def artifact(fixture_path, fixture_sum, generator, fault, invariant, k, header, routing_fp, policy_identity):
    gen = "bucketed_frontier" if generator == v1.GENERATOR_FRONTIER else "heading_aware_sweep"
    fields = [
        ("format", "apgar_gpu_batch_invariant_replay"), ("schema_version", 1),
        ("device_candidate_batch_schema_version", 1), ("fixture", fixture_path),
        ("fixture_checksum_fnv1a64", fixture_sum), ("generator", gen), ("fault", fault), ("layer", 0),
        ("maximum_rounds", 100000), ("maximum_workspace_states_per_query", 1000000),
        ("maximum_frontier_states_per_query", 1000000), ("maximum_device_bytes", 1000000000),
        ("maximum_host_bytes", 1000000000), ("batch_id", 4101 + k), ("query_id", 7001 + k),
        ("input_ordinal", 20 + k), ("policy_schema_version", 1), ("objective", "base_scalar_cost"),
        ("deterministic_seed", SEED), ("candidate_ordinal", ORDINAL), ("orthogonal_step_surcharge", 0),
        ("diagonal_step_surcharge", 0), ("bend_surcharge", 0), ("banned_resources", "none"),
        ("resource_penalties", "none"), ("expected_policy_identity", policy_identity),
        ("expected_board_content_hash", header.source_board_content_hash),
        ("expected_compiler_profile_fingerprint", header.compiler_profile_fingerprint),
        ("expected_compiler_version", header.compiler_version),
        ("expected_rule_bucket_identity", header.rule_bucket_identity),
        ("expected_routing_profile_fingerprint", routing_fp),
        ("expected_device_view_fingerprint", header.device_view_fingerprint),
        ("expected_failure", "internal_invariant"), ("expected_invariant", invariant),
    ]
    payload = "".join(f"{key}={value}\n" for key, value in fields).encode()
    return payload + f"checksum_fnv1a64={vc.stable_hash_string(payload)}\n".encode()


def route_artifact(fixture_path, fixture_sum, header, routing_fp):
    """A single-route gpu_replay v1 artifact (tools/gpu_replay_parser.cc:13-29):
    the frontier with the goal-predecessor self cycle."""
    fields = [
        ("format", "apgar_gpu_invariant_replay"), ("schema_version", 1), ("fixture", fixture_path),
        ("fixture_checksum_fnv1a64", fixture_sum), ("generator", "bucketed_frontier"),
        ("fault", "goal_predecessor_self_cycle"), ("layer", 0), ("maximum_rounds", 100000),
        ("expected_board_content_hash", header.source_board_content_hash),
        ("expected_profile_fingerprint", header.compiler_profile_fingerprint),
        ("expected_routing_profile_fingerprint", routing_fp),
        ("expected_device_view_fingerprint", header.device_view_fingerprint),
        ("expected_failure", "internal_invariant"), ("expected_invariant", "gpu.predecessor.self_reference.v1"),
    ]
    payload = "".join(f"{key}={value}\n" for key, value in fields).encode()
    return payload + f"checksum_fnv1a64={vc.stable_hash_string(payload)}\n".encode()


def policy_identity():
    h = vc.Fnv1a()
    h.string("APGAR-CANDIDATE-POLICY-V1")
    h.u32(1)
    h.byte(0)
    h.u64(SEED)
    h.u32(ORDINAL)
    for _ in range(5):
        h.u64(0)
    return h.v


def main():
    OUT.mkdir(exist_ok=True)
    data, header, routing_fp = fixture()
    fx_name = f"{CASE}_v1.fixture"
    (OUT / fx_name).write_bytes(data)
    fx_path = f"tests/golden/replays/{fx_name}"
    fx_sum = vc.stable_hash_string(data)
    pid = policy_identity()
    for k, (fault, (generator, invariant)) in enumerate(CONTRACT.items()):
        name = fault if fault.startswith("batch_") else f"batch_{fault}"
        (OUT / f"{name}_v1.replay").write_bytes(
            artifact(fx_path, fx_sum, generator, fault, invariant, k, header, routing_fp, pid))
    (OUT / "batch_compact_path_wrong_generator_v1.replay").write_bytes(
        artifact(fx_path, fx_sum, v1.GENERATOR_FRONTIER, "compact_path_bounds", CONTRACT["compact_path_bounds"][1],
                 len(CONTRACT), header, routing_fp, pid))
    (OUT / "batch_query_identity_wrong_routing_profile_v1.replay").write_bytes(
        artifact(fx_path, fx_sum, v1.GENERATOR_FRONTIER, "query_identity", CONTRACT["query_identity"][1],
                 len(CONTRACT) + 1, header, routing_fp + 1, pid))
    (OUT / "goal_predecessor_self_cycle_v1.replay").write_bytes(route_artifact(fx_path, fx_sum, header, routing_fp))
    (OUT / "goal_predecessor_self_cycle_wrong_routing_profile_v1.replay").write_bytes(
        route_artifact(fx_path, fx_sum, header, routing_fp + 1))


if __name__ == "__main__":
    main()
