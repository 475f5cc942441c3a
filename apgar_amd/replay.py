"""Invariant replays over libapgx.so: candidate batches (gpu_batch_replay v1)
and single routes (gpu_replay v1).

Mirrors tools/gpu_batch_replay.cc and tools/gpu_replay.cc: `python -m apgar_amd.replay` with no
arguments reproduces every default artifact and checks that every rejected
artifact fails at the generator/fault association rule; with arguments it
reproduces the named artifacts.  Exit status 0 all reproduced, 1 a failure was
not reproduced, 2 an artifact, fixture or association error.  Fixture paths in
artifacts are relative to the repository root.
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

from . import check, lib
from . import v1

ROOT = Path(__file__).resolve().parent.parent
REPLAYS = ROOT / "tests" / "golden" / "replays"
ASSOCIATION_ERROR = "replay format, schema, generator, policy, fault, or expected outcome is unsupported"
U32, U64, I64, P = C.c_uint32, C.c_uint64, C.c_int64, C.POINTER
STAGES = {1: "parse", 2: "fixture", 3: "association", 4: "execute", 5: "reproduce"}
DRY_RUN = 1


class BatchReplay(C.Structure):
    _fields_ = [("fixture", C.c_char * 256), ("fixture_checksum", U64), ("generator", U32), ("fault", U32),
                ("layer", U32), ("maximum_rounds", U32), ("maximum_workspace_states_per_query", U64),
                ("maximum_frontier_states_per_query", U64), ("maximum_device_bytes", U64),
                ("maximum_host_bytes", U64), ("batch_id", U64), ("query_id", U64), ("input_ordinal", U32),
                ("policy_schema_version", U32), ("deterministic_seed", U64), ("candidate_ordinal", U32),
                ("expected_compiler_version", U32), ("orthogonal_step_surcharge", U64),
                ("diagonal_step_surcharge", U64), ("bend_surcharge", U64), ("expected_policy_identity", U64),
                ("expected_board_content_hash", U64), ("expected_compiler_profile_fingerprint", U64),
                ("expected_rule_bucket_identity", U64), ("expected_routing_profile_fingerprint", U64),
                ("expected_device_view_fingerprint", U64), ("expected_failure", U32), ("reserved", U32),
                ("expected_invariant", C.c_char * 64)]


class RouteReplay(C.Structure):
    _fields_ = [("fixture", C.c_char * 256), ("fixture_checksum", U64), ("generator", U32), ("fault", U32),
                ("layer", U32), ("maximum_rounds", U32), ("expected_board_content_hash", U64),
                ("expected_profile_fingerprint", U64), ("expected_routing_profile_fingerprint", U64),
                ("expected_device_view_fingerprint", U64), ("expected_invariant", C.c_char * 64)]


class FixtureHeader(C.Structure):
    _fields_ = [("magic", C.c_char * 16), ("schema_version", U32), ("layer", U32), ("start_x", I64),
                ("start_y", I64), ("goal_x", I64), ("goal_y", I64), ("routing_profile_fingerprint", U64),
                ("profile", v1.CompiledProfile), ("node_count", U64)]


class Outcome(C.Structure):
    _fields_ = [("exit_code", C.c_int32), ("stage", U32), ("invariant", C.c_char * 64), ("message", C.c_char * 256)]

    @property
    def stage_name(self):
        return STAGES.get(self.stage, "?")


_SIGS = {
    "apgx_batch_replay_parse": (C.c_int, [C.c_char_p, U64, P(BatchReplay), C.c_char_p, U64]),
    "apgx_batch_replay_policy_identity": (U64, [P(BatchReplay)]),
    "apgx_batch_replay_run": (C.c_int, [C.c_char_p, U64, C.c_char_p, U64, U32, P(Outcome)]),
    "apgx_route_replay_parse": (C.c_int, [C.c_char_p, U64, P(RouteReplay), C.c_char_p, U64]),
    "apgx_route_replay_run": (C.c_int, [C.c_char_p, U64, C.c_char_p, U64, U32, P(Outcome)]),
}
ROUTE_FORMAT = b"format=apgar_gpu_invariant_replay\n"


def _lib():
    h = lib()
    if not getattr(h, "_replay_sigs", False):
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype, fn.argtypes = res, args
        h._replay_sigs = True
    return h


def exported_symbols():
    return list(_SIGS)


def parse(text: bytes):
    """(artifact, "") on success, (None, the reference's error message) otherwise."""
    out, err = BatchReplay(), C.create_string_buffer(256)
    rc = _lib().apgx_batch_replay_parse(text, len(text), C.byref(out), err, len(err))
    return (out, "") if rc == 0 else (None, err.value.decode())


def is_route_artifact(text: bytes) -> bool:
    """Single-route (gpu_replay v1) artifacts open with their own format line."""
    return text.startswith(ROUTE_FORMAT)


def parse_route(text: bytes):
    """Single-route artifact: (artifact, "") or (None, the reference's message)."""
    out, err = RouteReplay(), C.create_string_buffer(256)
    rc = _lib().apgx_route_replay_parse(text, len(text), C.byref(out), err, len(err))
    return (out, "") if rc == 0 else (None, err.value.decode())


def policy_identity(artifact: BatchReplay) -> int:
    return _lib().apgx_batch_replay_policy_identity(C.byref(artifact))


def fixture_bytes(profile: v1.CompiledProfile, nodes, start, goal, *, layer=0, routing_profile_fingerprint=0):
    """A replay fixture: header + compiled nodes [(x, y, layer, mask)]."""
    h = FixtureHeader()
    h.magic = b"APGX-V1-FIXTURE"
    h.schema_version, h.layer = 1, layer
    (h.start_x, h.start_y), (h.goal_x, h.goal_y) = start, goal
    h.routing_profile_fingerprint = routing_profile_fingerprint
    h.profile = profile
    h.node_count = len(nodes)
    arr = (v1.CompiledNode * max(len(nodes), 1))()
    for i, (x, y, lay, m) in enumerate(nodes):
        arr[i].lattice_x, arr[i].lattice_y, arr[i].layer, arr[i].legal_edges = x, y, lay, m
    return bytes(h) + bytes(arr)[: C.sizeof(v1.CompiledNode) * len(nodes)]


def run_bytes(text: bytes, fixture: bytes | None, dry_run=False) -> Outcome:
    """Runs a batch (gpu_batch_replay) or single-route (gpu_replay) artifact."""
    out = Outcome()
    fx = fixture if fixture is not None else None
    runner = _lib().apgx_route_replay_run if is_route_artifact(text) else _lib().apgx_batch_replay_run
    check(runner(text, len(text), fx, len(fx) if fx is not None else 0, DRY_RUN if dry_run else 0, C.byref(out)))
    return out


def run(path, dry_run=False) -> Outcome:
    """Reproduces one artifact; its fixture path is resolved against the
    repository root (a missing fixture is reported, not raised)."""
    text = Path(path).read_bytes()
    art, _ = parse_route(text) if is_route_artifact(text) else parse(text)
    fixture = None
    if art is not None:
        fp = ROOT / art.fixture.decode()
        fixture = fp.read_bytes() if fp.is_file() else None
    return run_bytes(text, fixture, dry_run)


def default_artifacts():
    """(reproduced, parse-rejected, gate-rejected) artifact lists under
    tests/golden/replays: *_wrong_generator_* must fail the parser's
    generator/fault association rule, any other *_wrong_* the association gate
    before GPU work, and every other artifact must reproduce."""
    allp = sorted(REPLAYS.glob("*.replay"))
    parse_rejected = [p for p in allp if "_wrong_generator_" in p.name]
    gate_rejected = [p for p in allp if "_wrong_" in p.name and p not in parse_rejected]
    return [p for p in allp if "_wrong_" not in p.name], parse_rejected, gate_rejected


def reject_at_gate(path) -> int:
    o = run(path, dry_run=True)
    if o.exit_code != 2 or o.stage_name != "association":
        print(f"{path}: association drift was not rejected before execution: {o.message.decode()}", file=sys.stderr)
        return 1
    print(f"rejected=association artifact={path}")
    return 0


def reject(path) -> int:
    art, err = parse(Path(path).read_bytes())
    if art is not None or err != ASSOCIATION_ERROR:
        print(f"{path}: invalid generator/fault association did not reach the semantic parser: {err}",
              file=sys.stderr)
        return 1
    print(f"rejected=generator_fault_association artifact={path}")
    return 0


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    code = 0

    def fold(rc):
        nonlocal code
        code = 2 if rc == 2 else (code or rc)

    if argv:
        todo, rejected, gated = argv, [], []
    else:
        todo, rejected, gated = default_artifacts()
    for p in gated:
        fold(reject_at_gate(p))
    for p in todo:
        o = run(p)
        stream = sys.stdout if o.exit_code == 0 else sys.stderr
        print(o.message.decode() if o.exit_code == 0 else f"{p}: {o.message.decode()}", file=stream)
        fold(o.exit_code)
    for p in rejected:
        fold(reject(p))
    return code


if __name__ == "__main__":
    sys.exit(main())
