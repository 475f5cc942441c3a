"""V1 device-view flattening (host C++ in libapgx.so; CPU only).

Pinned by the reference's published persistent device bytes per corpus case
(which fix node, layer, run and run-node counts), by an independent ordering in
the oracle, and by structural invariants of the runs.
"""
import json
from pathlib import Path

import pytest

from apgar_amd import v1
from tests import v1_corpus as vc

GOLDEN = json.loads((Path(__file__).parent / "golden" / "corpus_golden.json").read_text())


def board_for(profile):
    nodes = vc.compile_obstacle_free(profile)
    return nodes, v1.build_board(nodes, tile_w=profile.tile_w, tile_h=profile.tile_h,
                                 costs=profile.costs, heading_mask=profile.heading,
                                 step=profile.step, profile_fingerprint=profile.fingerprint())


@pytest.mark.parametrize("name", sorted(vc.corpus()))
def test_persistent_bytes_match_published(apgx, name):
    profile, _, _ = vc.corpus()[name]
    _, b = board_for(profile)
    nl, nn, nr, nrn = b.counts
    assert b.header.estimated_persistent_device_bytes == vc.persistent_bytes(nn, nr, nrn, nl)
    assert b.header.estimated_persistent_device_bytes == int(GOLDEN[name]["descriptor"]["persistent_device_bytes"])
    assert b.header.represented_states == 9 * nn
    assert b.fingerprint() == b.header.device_view_fingerprint


@pytest.mark.parametrize("name", sorted(vc.corpus()))
def test_flatten_order_matches_oracle_and_runs_partition_edges(apgx, oracle, name):
    profile, _, _ = vc.corpus()[name]
    nodes, b = board_for(profile)
    o = oracle.V1(nodes, profile.costs, profile.heading, profile.tile_w, profile.tile_h)
    order = o.flatten_order(profile.tile_w, profile.tile_h)
    got = [(n.lattice_x, n.lattice_y) for n in b.nodes]
    assert got == [(nodes[i][0], nodes[i][1]) for i in order]
    # every legal directed edge appears in exactly one maximal run
    seen = set()
    for r in b.runs:
        seq = [b.run_nodes[r.node_offset + k] for k in range(r.node_count)]
        for s, t in zip(seq[:-1], seq[1:]):
            assert b.nodes[s].neighbors[r.direction] == t
            assert (s, r.direction) not in seen
            seen.add((s, r.direction))
    legal = {(i, d) for i, n in enumerate(b.nodes) for d in range(8) if n.legal_edges >> d & 1}
    assert seen == legal


def test_fingerprint_changes_with_content(apgx):
    profile, _, _ = vc.corpus()["dense_corridors"]
    _, b = board_for(profile)
    fp = b.fingerprint()
    b.nodes[0].lattice_x += 1
    assert b.fingerprint() != fp


def test_duplicate_nodes_rejected(apgx):
    import apgar_amd

    with pytest.raises(apgar_amd.ApgxError) as e:
        v1.build_board([(0, 0, 0, 0), (0, 0, 0, 0)], tile_w=4, tile_h=4, costs=(1, 1, 1),
                       heading_mask=7)
    assert e.value.code == 4
