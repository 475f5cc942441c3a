"""ctypes mirror of the V1 (reference record) C-ABI of libapgx.so.

Records are byte-identical to include/apgar/gpu/planar_router.h:170-423.  The
helpers here only marshal records; all flattening, validation and search runs
in the library.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import NamedTuple, Optional

import numpy as np

from . import check, lib

INVALID = 0xFFFFFFFF
BANNED = 0xFFFFFFFFFFFFFFFF
GENERATOR_FRONTIER, GENERATOR_SWEEP = 0, 1
REACHED, DISCONNECTED, BUDGET_EXHAUSTED, CANCELLED = 0, 1, 2, 3
READBACK_COMPACT_PATHS, READBACK_FULL_WORKSPACE = 0, 1
# apgx_v1_pending_inject_fault ids, in the reference decorator's order
FAULT_IDS = {"none": 0, "workspace_bounds": 1, "workspace_owner": 2, "query_telemetry": 3, "memory_accounting": 4,
             "batch_telemetry": 5, "query_identity": 6, "false_disconnected": 7, "compact_path_bounds": 8,
             "compact_path_endpoint": 9, "compact_path_cycle": 10, "compact_path_heading": 11,
             "compact_path_edge": 12}
FAULT_NONE = 0
CHECK_OK, CHECK_INTERNAL_INVARIANT, CHECK_VALIDATION_FAILED = 0, 1, 2
U8, U32, U64, I64, P = C.c_uint8, C.c_uint32, C.c_uint64, C.c_int64, C.POINTER


class Header(C.Structure):
    _fields_ = [("schema_version", U32), ("compiler_version", U32),
                ("source_board_content_hash", U64), ("compiler_profile_fingerprint", U64),
                ("rule_bucket_identity", U64), ("device_view_fingerprint", U64),
                ("represented_nodes", U64), ("represented_states", U64),
                ("estimated_persistent_device_bytes", U64), ("lattice_origin_x", I64),
                ("lattice_origin_y", I64), ("lattice_step", I64), ("orthogonal_step", U32),
                ("diagonal_step", U32), ("bend", U32), ("heading_mask", U8), ("reserved", U8 * 7)]


class LayerRange(C.Structure):
    _fields_ = [("layer", U32), ("node_offset", U32), ("node_count", U32), ("pad", U32),
                ("minimum_lattice_x", I64), ("maximum_lattice_x", I64),
                ("minimum_lattice_y", I64), ("maximum_lattice_y", I64)]


class Node(C.Structure):
    _fields_ = [("lattice_x", I64), ("lattice_y", I64), ("layer", U32), ("legal_edges", U8),
                ("reserved", U8 * 3), ("neighbors", U32 * 8)]


class Run(C.Structure):
    _fields_ = [("node_offset", U32), ("node_count", U32), ("direction", U8), ("reserved", U8 * 3)]


class PolicyEdge(C.Structure):
    _fields_ = [("physical_edge_index", U32), ("reserved", U32), ("adjustment", U64)]


class Query(C.Structure):
    _fields_ = [("schema_version", U32), ("input_ordinal", U32), ("batch_id", U64),
                ("query_id", U64), ("workspace_owner", U64), ("policy_identity", U64),
                ("routing_profile_fingerprint", U64), ("start_node", U32), ("goal_node", U32),
                ("policy_offset", U32), ("policy_count", U32), ("workspace_offset", U64),
                ("workspace_state_count", U64), ("frontier_state_capacity", U64),
                ("orthogonal_step_surcharge", U64), ("diagonal_step_surcharge", U64),
                ("bend_surcharge", U64), ("maximum_rounds", U32), ("generator", U8),
                ("reserved", U8 * 3)]


class Result(C.Structure):
    _fields_ = [("schema_version", U32), ("compiler_version", U32), ("input_ordinal", U32),
                ("start_node", U32), ("goal_node", U32), ("goal_state", U32), ("rounds", U32),
                ("generator", U8), ("completion", U8), ("reserved", U8 * 2), ("batch_id", U64),
                ("query_id", U64), ("workspace_owner", U64), ("policy_identity", U64),
                ("routing_profile_fingerprint", U64), ("source_board_content_hash", U64),
                ("compiler_profile_fingerprint", U64), ("rule_bucket_identity", U64),
                ("device_view_fingerprint", U64), ("examined_work", U64),
                ("heading_turn_relaxations", U64)]


class CompactPath(C.Structure):
    _fields_ = [("schema_version", U32), ("state_count", U32), ("query_id", U64),
                ("workspace_owner", U64), ("state_offset", U64)]


class BatchRequest(C.Structure):
    _fields_ = [("schema_version", U32), ("generator", U32), ("batch_id", U64),
                ("maximum_rounds", U32), ("reserved", U32),
                ("maximum_workspace_states_per_query", U64),
                ("maximum_frontier_states_per_query", U64), ("maximum_device_bytes", U64),
                ("maximum_host_bytes", U64)]


class BatchTelemetry(C.Structure):
    _fields_ = [("persistent_device_bytes", U64), ("batch_device_bytes", U64),
                ("workspace_capacity_device_bytes", U64), ("peak_device_bytes", U64),
                ("batch_host_bytes", U64), ("device_to_host_readback_bytes", U64),
                ("kernel_launch_count", U64), ("blocking_status_readback_count", U64),
                ("dispatched_rounds", U32), ("finalization_launch_count", U32),
                ("chunk_rounds", U32), ("reserved", U32), ("kernel_milliseconds", C.c_double)]


class CompiledProfile(C.Structure):
    _fields_ = [("compiler_version", U32), ("heading_mask", U32),
                ("source_board_content_hash", U64), ("compiler_profile_fingerprint", U64),
                ("rule_bucket_identity", U64), ("lattice_origin_x", I64), ("lattice_origin_y", I64),
                ("lattice_step", I64), ("tile_width_nodes", U32), ("tile_height_nodes", U32),
                ("orthogonal_step", U32), ("diagonal_step", U32), ("bend", U32), ("reserved", U32)]


class CompiledNode(C.Structure):
    _fields_ = [("lattice_x", I64), ("lattice_y", I64), ("layer", U32), ("legal_edges", U8),
                ("reserved", U8 * 3)]


class Validation(C.Structure):
    _fields_ = [("valid", U32), ("query_index", U32), ("failure", U32), ("reserved", U32),
                ("invariant", C.c_char * 64), ("detail", C.c_char * 128)]


class RouteRequest(C.Structure):
    _fields_ = [("generator", U32), ("start_node", U32), ("goal_node", U32), ("maximum_rounds", U32),
                ("routing_profile_fingerprint", U64), ("maximum_device_bytes", U64)]


class RouteHeader(C.Structure):
    _fields_ = [("schema_version", U32), ("compiler_version", U32), ("start_node", U32), ("goal_node", U32),
                ("source_board_content_hash", U64), ("compiler_profile_fingerprint", U64),
                ("routing_profile_fingerprint", U64), ("rule_bucket_identity", U64),
                ("device_view_fingerprint", U64), ("generator", U8), ("reserved", U8 * 7)]


class KernelTelemetry(C.Structure):
    _fields_ = [("persistent_device_bytes", U64), ("batch_device_bytes", U64), ("peak_device_bytes", U64),
                ("examined_work", U64), ("heading_turn_relaxations", U64), ("rounds", U32), ("reserved", U32),
                ("kernel_milliseconds", C.c_double)]


class RouteResult(C.Structure):
    _fields_ = [("header", RouteHeader), ("completion", U32), ("goal_state", U32), ("state_count", U64),
                ("telemetry", KernelTelemetry)]


class RouteReadback(NamedTuple):
    result: RouteResult
    labels: np.ndarray
    predecessors: np.ndarray


ROUTE_FAULT_NONE, ROUTE_FAULT_GOAL_PREDECESSOR_SELF_CYCLE = 0, 1


class GeneratorShape(C.Structure):
    _fields_ = [("chunk_rounds", U32), ("fixed_launches", U32), ("fixed_examined_work_per_round", U32),
                ("examined_work_per_node_per_round", U32), ("heading_turn_relaxations_per_examined_work", U32),
                ("readback_kind", U32), ("launches_once_per_chunk", U32), ("reserved", U32)]


class QueryEnvelope(C.Structure):
    _fields_ = [("workspace_offset", U64), ("workspace_state_count", U64), ("examined_work", U64),
                ("heading_turn_relaxations", U64), ("rounds", U32), ("reserved", U32)]


class WorkspaceView(C.Structure):
    _fields_ = [("readback_kind", U32), ("reserved", U32), ("queries", P(QueryEnvelope)),
                ("labels", P(U64)), ("label_count", U64), ("predecessors", P(U32)), ("predecessor_count", U64),
                ("state_owners", P(U64)), ("state_owner_count", U64), ("predecessor_owners", P(U64)),
                ("predecessor_owner_count", U64)]


@dataclass
class Workspace:
    """The workspace half of a readback: readback kind, per-query envelopes
    and, for the frontier, the flat query-major arrays (numpy)."""
    readback_kind: int
    envelopes: list
    labels: np.ndarray
    predecessors: np.ndarray
    state_owners: np.ndarray
    predecessor_owners: np.ndarray

    def copy(self):
        return Workspace(self.readback_kind, [QueryEnvelope(e.workspace_offset, e.workspace_state_count,
                                                            e.examined_work, e.heading_turn_relaxations, e.rounds, 0)
                                              for e in self.envelopes],
                         self.labels.copy(), self.predecessors.copy(), self.state_owners.copy(),
                         self.predecessor_owners.copy())


class Readback(NamedTuple):
    results: list
    paths: list
    states: list
    telemetry: "BatchTelemetry"
    workspace: Optional[Workspace]


_SIGS = {
    "apgx_v1_build_board": (C.c_int, [P(CompiledProfile), P(CompiledNode), U64, P(C.c_void_p)]),
    "apgx_v1_board_sizes": (C.c_int, [C.c_void_p, P(U64), P(U64), P(U64), P(U64)]),
    "apgx_v1_board_copy": (C.c_int, [C.c_void_p, P(Header), P(LayerRange), P(Node), P(Run), P(U32)]),
    "apgx_v1_board_release": (None, [C.c_void_p]),
    "apgx_v1_view_fingerprint": (U64, [P(Header), P(LayerRange), U64, P(Node), U64, P(Run), U64,
                                       P(U32), U64]),
    "apgx_v1_view_upload": (C.c_int, [P(Header), P(LayerRange), U64, P(Node), U64, P(Run), U64,
                                      P(U32), U64, P(C.c_void_p)]),
    "apgx_v1_view_release": (None, [C.c_void_p]),
    "apgx_v1_batch_execute": (C.c_int, [C.c_void_p, P(BatchRequest), P(Query), U32, P(PolicyEdge),
                                        U64, P(U32), P(C.c_void_p)]),
    "apgx_v1_batch_readback": (C.c_int, [C.c_void_p, P(Result), P(CompactPath), P(U32), U64,
                                         P(U64), P(BatchTelemetry)]),
    "apgx_v1_batch_readback_workspace": (C.c_int, [C.c_void_p, P(U32), P(QueryEnvelope), P(U64), P(U32),
                                                   P(U64), P(U64), U64, P(U64)]),
    "apgx_v1_pending_inject_fault": (C.c_int, [C.c_void_p, U32]),
    "apgx_v1_generator_shape": (C.c_int, [U32, P(GeneratorShape)]),
    "apgx_v1_expected_batch_bytes": (C.c_int, [U32, U64, U64, U64, U64, P(U64), P(U64), P(U64)]),
    "apgx_v1_pending_release": (None, [C.c_void_p]),
    "apgx_v1_validate_batch": (C.c_int, [P(Header), P(Node), U64, P(BatchRequest), P(Query), U32, P(PolicyEdge),
                                         U64, P(Result), P(CompactPath), P(U32), U64, P(WorkspaceView),
                                         P(BatchTelemetry), P(U64), P(Validation)]),
    "apgx_v1_validate_batch_queries": (C.c_int, [P(Header), P(Node), U64, P(BatchRequest), P(Query), U32,
                                                 P(PolicyEdge), U64, P(Result), P(CompactPath), P(U32), U64,
                                                 P(WorkspaceView), P(BatchTelemetry), P(U64), P(Validation),
                                                 P(Validation)]),
    "apgx_v1_route_execute": (C.c_int, [C.c_void_p, P(RouteRequest), P(U32), P(C.c_void_p)]),
    "apgx_v1_route_readback": (C.c_int, [C.c_void_p, P(RouteResult), P(U64), P(U32), U64]),
    "apgx_v1_route_pending_inject_fault": (C.c_int, [C.c_void_p, U32]),
    "apgx_v1_route_pending_release": (None, [C.c_void_p]),
    "apgx_v1_expected_route_bytes": (C.c_int, [U32, U64, P(U64)]),
    "apgx_v1_validate_route": (C.c_int, [P(Header), P(Node), U64, P(RouteRequest), P(RouteResult), P(U64), P(U32),
                                         P(U64), P(Validation)]),
}


def _lib():
    h = lib()
    if not getattr(h, "_v1_sigs", False):
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        h._v1_sigs = True
    return h


def exported_symbols():
    return list(_SIGS)


@dataclass
class Board:
    header: Header
    layers: C.Array
    nodes: C.Array
    runs: C.Array
    run_nodes: C.Array

    @property
    def counts(self):
        return len(self.layers), len(self.nodes), len(self.runs), len(self.run_nodes)

    def node_index(self, x, y, layer=0):
        for i, n in enumerate(self.nodes):
            if n.lattice_x == x and n.lattice_y == y and n.layer == layer:
                return i
        raise KeyError((x, y, layer))

    def fingerprint(self) -> int:
        nl, nn, nr, nrn = self.counts
        return _lib().apgx_v1_view_fingerprint(C.byref(self.header), self.layers, nl, self.nodes,
                                               nn, self.runs, nr, self.run_nodes, nrn)


def build_board(nodes, *, tile_w, tile_h, costs, heading_mask, step=10, origin=(0, 0),
                compiler_version=1, source_hash=0, profile_fingerprint=0, rule_bucket=0) -> Board:
    """Flattens compiled nodes [(x, y, layer, mask)] through apgx_v1_build_board."""
    prof = CompiledProfile(compiler_version, heading_mask, source_hash, profile_fingerprint,
                           rule_bucket, origin[0], origin[1], step, tile_w, tile_h, costs[0],
                           costs[1], costs[2], 0)
    arr = (CompiledNode * max(len(nodes), 1))()
    for i, (x, y, l, m) in enumerate(nodes):
        arr[i].lattice_x, arr[i].lattice_y, arr[i].layer, arr[i].legal_edges = x, y, l, m
    h = _lib()
    handle = C.c_void_p()
    check(h.apgx_v1_build_board(C.byref(prof), arr, len(nodes), C.byref(handle)))
    try:
        nl, nn, nr, nrn = U64(), U64(), U64(), U64()
        check(h.apgx_v1_board_sizes(handle, C.byref(nl), C.byref(nn), C.byref(nr), C.byref(nrn)))
        header = Header()
        layers = (LayerRange * nl.value)()
        nodes_out = (Node * nn.value)()
        runs = (Run * nr.value)()
        run_nodes = (U32 * nrn.value)()
        check(h.apgx_v1_board_copy(handle, C.byref(header), layers, nodes_out, runs, run_nodes))
    finally:
        h.apgx_v1_board_release(handle)
    return Board(header, layers, nodes_out, runs, run_nodes)


class View:
    def __init__(self, board: Board):
        nl, nn, nr, nrn = board.counts
        self.board = board
        self.handle = C.c_void_p()
        check(_lib().apgx_v1_view_upload(C.byref(board.header), board.layers, nl, board.nodes, nn,
                                         board.runs, nr, board.run_nodes, nrn, C.byref(self.handle)))

    def close(self):
        if self.handle:
            _lib().apgx_v1_view_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def execute(self, queries, policy_edges=(), *, batch_id=1, maximum_rounds=100_000,
                generator=GENERATOR_SWEEP, maximum_device_bytes=(1 << 63), cancel=None, fault=FAULT_NONE,
                maximum_host_bytes=(1 << 63)):
        """Runs a batch and returns its Readback: results, compact paths and
        packed states (sweep; empty for the frontier), telemetry, and the
        workspace (envelopes for both; flat arrays for the frontier).
        `fault` (replay/test only) corrupts query 0 of the host copy with the
        library's readback decorator before it is read back."""
        S = len(self.board.nodes) * 9
        req = BatchRequest(1, generator, batch_id, maximum_rounds, 0, S, S, maximum_device_bytes,
                           maximum_host_bytes)
        qarr = (Query * max(len(queries), 1))(*queries)
        parr = (PolicyEdge * max(len(policy_edges), 1))(*policy_edges)
        pending = C.c_void_p()
        flag = (U32 * 1)(1 if cancel else 0)
        check(_lib().apgx_v1_batch_execute(self.handle, C.byref(req), qarr, len(queries), parr,
                                           len(policy_edges), flag if cancel is not None else None,
                                           C.byref(pending)))
        try:
            check(_lib().apgx_v1_pending_inject_fault(pending, fault))
            used = U64()
            results = (Result * max(len(queries), 1))()
            paths = (CompactPath * max(len(queries), 1))()
            check(_lib().apgx_v1_batch_readback(pending, results, paths, None, 0, C.byref(used), None))
            states = (U32 * max(used.value, 1))()
            tel = BatchTelemetry()
            check(_lib().apgx_v1_batch_readback(pending, results, paths, states, used.value,
                                                C.byref(used), C.byref(tel)))
            ws = _read_workspace(pending, len(queries))
        finally:
            _lib().apgx_v1_pending_release(pending)
        n = len(queries)
        return Readback(list(results)[:n], list(paths)[:n], list(states)[: used.value], tel, ws)


    def route(self, start, goal, *, generator=GENERATOR_FRONTIER, maximum_rounds=100_000,
              routing_profile_fingerprint=0, maximum_device_bytes=(1 << 64) - 1, cancel=None,
              fault=ROUTE_FAULT_NONE):
        """Executes one single route (apgx_v1_route_execute) and reads it back:
        RouteReadback(result, labels, predecessors).  `cancel` (a one-element
        ctypes uint32 array, or True for an already set flag) is the caller's
        cancellation word; `fault` (replay/test only) applies the readback
        decorator."""
        req = RouteRequest(generator, start, goal, maximum_rounds, routing_profile_fingerprint,
                           maximum_device_bytes)
        flag = None
        if cancel is True:
            flag = (U32 * 1)(1)
        elif cancel is not None:
            flag = cancel
        pending = C.c_void_p()
        check(_lib().apgx_v1_route_execute(self.handle, C.byref(req), flag, C.byref(pending)))
        try:
            check(_lib().apgx_v1_route_pending_inject_fault(pending, fault))
            S = len(self.board.nodes) * 9
            labels = np.zeros(S, np.uint64)
            preds = np.zeros(S, np.uint32)
            res = RouteResult()
            check(_lib().apgx_v1_route_readback(pending, C.byref(res), labels.ctypes.data_as(P(U64)),
                                                preds.ctypes.data_as(P(U32)), S))
        finally:
            _lib().apgx_v1_route_pending_release(pending)
        return RouteReadback(res, labels, preds)


def route_request_for(start, goal, *, generator=GENERATOR_FRONTIER, maximum_rounds=100_000,
                      routing_profile_fingerprint=0, maximum_device_bytes=(1 << 64) - 1):
    return RouteRequest(generator, start, goal, maximum_rounds, routing_profile_fingerprint, maximum_device_bytes)


def validate_route(board: Board, request: RouteRequest, readback: RouteReadback):
    """Host trust-boundary check of a single route (apgx_v1_validate_route):
    (valid, invariant id, detail, path cost, failure class)."""
    labels = np.ascontiguousarray(readback.labels, dtype=np.uint64)
    preds = np.ascontiguousarray(readback.predecessors, dtype=np.uint32)
    cost = U64()
    out = Validation()
    check(_lib().apgx_v1_validate_route(C.byref(board.header), board.nodes, len(board.nodes), C.byref(request),
                                        C.byref(readback.result), labels.ctypes.data_as(P(U64)),
                                        preds.ctypes.data_as(P(U32)), C.byref(cost), C.byref(out)))
    return bool(out.valid), out.invariant.decode(), out.detail.decode(), cost.value, out.failure


def expected_route_bytes(generator, node_count):
    out = U64()
    check(_lib().apgx_v1_expected_route_bytes(generator, node_count, C.byref(out)))
    return out.value


def _read_workspace(pending, n):
    kind, total = U32(), U64()
    env = (QueryEnvelope * max(n, 1))()
    check(_lib().apgx_v1_batch_readback_workspace(pending, C.byref(kind), env, None, None, None, None, 0,
                                                  C.byref(total)))
    t = total.value
    lab = np.zeros(t, np.uint64)
    pre = np.zeros(t, np.uint32)
    sown = np.zeros(t, np.uint64)
    pown = np.zeros(t, np.uint64)
    if t:
        check(_lib().apgx_v1_batch_readback_workspace(
            pending, None, None, lab.ctypes.data_as(P(U64)), pre.ctypes.data_as(P(U32)),
            sown.ctypes.data_as(P(U64)), pown.ctypes.data_as(P(U64)), t, C.byref(total)))
    return Workspace(kind.value, list(env)[:n], lab, pre, sown, pown)


def reconstruct(workspace: Workspace, qi, result):
    """Reverse state path (goal .. start) of a reached frontier query, walked
    along the readback's predecessors from its goal state (no validation)."""
    off = workspace.envelopes[qi].workspace_offset
    out, cur = [], result.goal_state
    start = result.start_node * 9 + 8
    while cur != INVALID and len(out) <= workspace.envelopes[qi].workspace_state_count:
        out.append(cur)
        if cur == start:
            return out
        cur = int(workspace.predecessors[off + cur])
    raise ValueError("predecessor walk does not reach the start")


def make_query(i, start, goal, board: Board, *, batch_id=1, policy_offset=0, policy_count=0,
               surcharges=(0, 0, 0), maximum_rounds=100_000, generator=GENERATOR_SWEEP):
    S = len(board.nodes) * 9
    q = Query()
    q.schema_version = 1
    q.input_ordinal = i
    q.batch_id = batch_id
    q.query_id = i + 1
    q.workspace_owner = i + 1
    q.start_node, q.goal_node = start, goal
    q.policy_offset, q.policy_count = policy_offset, policy_count
    q.workspace_offset = i * S
    q.workspace_state_count = S
    q.orthogonal_step_surcharge, q.diagonal_step_surcharge, q.bend_surcharge = surcharges
    q.maximum_rounds = maximum_rounds
    q.generator = generator
    q.frontier_state_capacity = S if generator == GENERATOR_FRONTIER else 0
    return q


def physical_edge_index(board: Board, x, y, direction, layer=0):
    """Canonical physical edge id node*8 + dir (< 4) of a canonical (x, y, dir) resource."""
    return board.node_index(x, y, layer) * 8 + direction


def request_for(board: Board, *, batch_id=1, maximum_rounds=100_000, generator=GENERATOR_SWEEP,
                maximum_device_bytes=(1 << 63)):
    """The BatchRequest View.execute sends for these arguments."""
    S = len(board.nodes) * 9
    return BatchRequest(1, generator, batch_id, maximum_rounds, 0, S, S, maximum_device_bytes, 1 << 63)


def _validation_args(board: Board, request, queries, policy_edges, results, paths, states, telemetry, workspace):
    """The leading arguments of apgx_v1_validate_batch[_queries] and the
    ctypes objects they point into (kept alive by the caller)."""
    n = len(queries)
    S = len(board.nodes) * 9
    if workspace is None:
        workspace = Workspace(READBACK_COMPACT_PATHS,
                              [QueryEnvelope(i * S, S, r.examined_work, r.heading_turn_relaxations, r.rounds, 0)
                               for i, r in enumerate(results)],
                              *(np.zeros(0, d) for d in (np.uint64, np.uint32, np.uint64, np.uint64)))
    env = (QueryEnvelope * max(n, 1))(*workspace.envelopes)
    arrays = [np.ascontiguousarray(a) for a in (workspace.labels, workspace.predecessors, workspace.state_owners,
                                                  workspace.predecessor_owners)]
    wv = WorkspaceView(workspace.readback_kind, 0, env,
                       arrays[0].ctypes.data_as(P(U64)), len(arrays[0]), arrays[1].ctypes.data_as(P(U32)),
                       len(arrays[1]), arrays[2].ctypes.data_as(P(U64)), len(arrays[2]),
                       arrays[3].ctypes.data_as(P(U64)), len(arrays[3]))
    qarr = (Query * max(n, 1))(*queries)
    parr = (PolicyEdge * max(len(policy_edges), 1))(*policy_edges)
    rarr = (Result * max(n, 1))(*results)
    carr = (CompactPath * max(n, 1))(*paths)
    sarr = (U32 * max(len(states), 1))(*states)
    args = (C.byref(board.header), board.nodes, len(board.nodes), C.byref(request), qarr, n, parr,
            len(policy_edges), rarr, carr, sarr, len(states), C.byref(wv), C.byref(telemetry))
    return args, (env, arrays, wv, qarr, parr, rarr, carr, sarr)


def validate(board: Board, request, queries, policy_edges, results, paths, states, telemetry, workspace=None):
    """Host trust-boundary check of a readback (apgx_v1_validate_batch).
    `workspace` None means a compact readback with canonical envelopes.
    Returns (valid, query_index, invariant id, detail, path costs, failure class)."""
    n = len(queries)
    args, _keep = _validation_args(board, request, queries, policy_edges, results, paths, states, telemetry,
                                   workspace)
    costs = (U64 * max(n, 1))()
    out = Validation()
    check(_lib().apgx_v1_validate_batch(*args, costs, C.byref(out)))
    return (bool(out.valid), out.query_index, out.invariant.decode(), out.detail.decode(), list(costs)[:n],
            out.failure)


def validate_queries(board: Board, request, queries, policy_edges, results, paths, states, telemetry,
                     workspace=None):
    """Per-query verdicts of a readback (apgx_v1_validate_batch_queries).
    Returns (batch verdict, [query verdicts], path costs); a verdict is a tuple
    (valid, query_index, invariant id, detail, failure class)."""
    n = len(queries)
    args, _keep = _validation_args(board, request, queries, policy_edges, results, paths, states, telemetry,
                                   workspace)
    costs = (U64 * max(n, 1))()
    verdicts = (Validation * max(n, 1))()
    batch = Validation()
    check(_lib().apgx_v1_validate_batch_queries(*args, costs, verdicts, C.byref(batch)))
    as_tuple = lambda v: (bool(v.valid), v.query_index, v.invariant.decode(), v.detail.decode(), v.failure)
    return as_tuple(batch), [as_tuple(verdicts[i]) for i in range(n)], list(costs)[:n]


def generator_shape(generator) -> GeneratorShape:
    out = GeneratorShape()
    check(_lib().apgx_v1_generator_shape(generator, C.byref(out)))
    return out


def expected_batch_bytes(generator, query_count, policy_edge_count, node_count, compact_path_states=0):
    """(device, host, readback) bytes of a batch under this backend's accounting."""
    d, h, r = U64(), U64(), U64()
    check(_lib().apgx_v1_expected_batch_bytes(generator, query_count, policy_edge_count, node_count,
                                              compact_path_states, C.byref(d), C.byref(h), C.byref(r)))
    return d.value, h.value, r.value
