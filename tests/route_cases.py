"""Shared helpers of the single-route tests: corpus boards, oracle-made
readbacks in the library's record form, and the hand-worked corridor."""
from __future__ import annotations

import numpy as np

from apgar_amd import v1
from tests import v1_corpus as vc
from tests.test_v1_board import board_for

COMPLETION = {0: v1.REACHED, 1: v1.DISCONNECTED, 2: v1.BUDGET_EXHAUSTED}


def corpus_case(oracle, name):
    profile, start, goal = vc.corpus()[name]
    nodes, b = board_for(profile)
    o = oracle.V1(nodes, profile.costs, profile.heading, profile.tile_w, profile.tile_h)
    s_xy, g_xy = vc.lattice(profile, start), vc.lattice(profile, goal)
    return profile, b, o, s_xy, g_xy, b.node_index(*s_xy), b.node_index(*g_xy)


def corridor(oracle):
    """Two nodes east-west, orth 10, diag 14, bend 3 (hand-worked in
    tests/test_oracle_v1_route.py)."""
    nodes = [(0, 0, 0, 1 << 0), (1, 0, 0, 1 << 4)]
    b = v1.build_board(nodes, tile_w=4, tile_h=4, costs=(10, 14, 3), heading_mask=1)
    o = oracle.V1(nodes, (10, 14, 3), 1, 4, 4)
    return b, o


def oracle_readback(oracle, b, o, s_xy, g_xy, generator, max_rounds=100_000, routing_fp=0):
    """The oracle's single route as a readback the host validator takes."""
    r = oracle.v1_route(o, s_xy, g_xy, generator, max_rounds=max_rounds)
    s, g = b.node_index(*s_xy), b.node_index(*g_xy)
    res = v1.RouteResult()
    h = res.header
    h.schema_version = 1
    h.compiler_version = b.header.compiler_version
    h.start_node, h.goal_node = s, g
    h.source_board_content_hash = b.header.source_board_content_hash
    h.compiler_profile_fingerprint = b.header.compiler_profile_fingerprint
    h.routing_profile_fingerprint = routing_fp
    h.rule_bucket_identity = b.header.rule_bucket_identity
    h.device_view_fingerprint = b.header.device_view_fingerprint
    h.generator = generator
    res.completion = COMPLETION[r["completion"]]
    res.goal_state = r["goal_state"] if r["completion"] == 0 else v1.INVALID
    res.state_count = 9 * len(b.nodes)
    t = res.telemetry
    t.persistent_device_bytes = b.header.estimated_persistent_device_bytes
    t.batch_device_bytes = v1.expected_route_bytes(generator, len(b.nodes))
    t.peak_device_bytes = t.persistent_device_bytes + t.batch_device_bytes
    t.examined_work = r["examined_work"]
    t.heading_turn_relaxations = r["turns"]
    t.rounds = r["rounds"]
    req = v1.route_request_for(s, g, generator=generator, maximum_rounds=max_rounds, routing_profile_fingerprint=routing_fp)
    return req, v1.RouteReadback(res, r["labels"].copy(), r["predecessors"].copy()), r


def clone(rb: v1.RouteReadback) -> v1.RouteReadback:
    res = v1.RouteResult()
    import ctypes

    ctypes.memmove(ctypes.byref(res), ctypes.byref(rb.result), ctypes.sizeof(res))
    return v1.RouteReadback(res, np.array(rb.labels, copy=True), np.array(rb.predecessors, copy=True))
