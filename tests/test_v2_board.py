"""DeviceDenseBoardV2 wire records (apgx_v2_board_*, apgx_lattice_create_v2).

CPU: the fingerprint against an independent restatement of its field order,
the derived fields against numpy, the wire round trip, every rejection, and
the net-batch record layouts the schema documents.  GPU: a board uploaded
from its sealed form routes exactly as the same plane uploaded directly."""
import ctypes as C

import numpy as np
import pytest

from tests import v1_corpus as vc


def _board(apgx, cost_bytes=2, seed=3):
    cost = apgx.generate_lattice(64, 48, 3, 1, 8, blocked_ppm=100000, seed=seed)
    return apgx.seal_board(cost, 20, cost_bytes=cost_bytes, compiler_version=1, source_board_content_hash=5,
                           compiler_profile_fingerprint=6, rule_bucket_identity=7, origin=(100, -7), step=25)


def restated_fingerprint(h, plane):
    """Field order of DeviceDenseBoardV2's fingerprint, restated from the schema."""
    f = vc.Fnv1a()
    f.string("APGAR-DEVICE-DENSE-BOARD-V2")
    f.u32(h.schema_version); f.u32(h.compiler_version)
    f.u64(h.source_board_content_hash); f.u64(h.compiler_profile_fingerprint); f.u64(h.rule_bucket_identity)
    f.u64(h.estimated_persistent_device_bytes)
    f.i64(h.lattice_origin_x); f.i64(h.lattice_origin_y); f.i64(h.lattice_step)
    for v in (h.width, h.height, h.layers, h.cost_bytes, h.via_cost, h.min_cell_cost, h.max_cell_cost):
        f.u32(v)
    f.u64(h.free_cells)
    for b in np.ascontiguousarray(plane).tobytes():
        f.byte(b)
    return f.v


@pytest.mark.parametrize("cost_bytes", [2, 4])
def test_seal_derives_fields_and_fingerprint(apgx, cost_bytes):
    h, plane = _board(apgx, cost_bytes)
    free = plane[plane > 0]
    assert (h.free_cells, h.min_cell_cost, h.max_cell_cost) == (free.size, free.min(), free.max())
    assert h.estimated_persistent_device_bytes == 128 + plane.nbytes
    assert h.device_view_fingerprint == restated_fingerprint(h, plane)
    assert apgx.lib().apgx_v2_board_fingerprint(C.byref(h), plane.ctypes.data) == h.device_view_fingerprint
    apgx.validate_board(h, plane)


def test_wire_round_trip(apgx):
    h, plane = _board(apgx)
    data = apgx.board_bytes(h, plane)
    assert len(data) == 128 + plane.nbytes and C.sizeof(apgx.V2BoardHeader) == 128
    h2, p2 = apgx.parse_board_bytes(data)
    assert bytes(h2) == bytes(h) and np.array_equal(p2, plane)
    apgx.validate_board(h2, p2)


def _mutations(apgx):
    def field(name, value):
        def f(h, p):
            setattr(h, name, value)
            return h, p
        return f

    def plane_byte(h, p):
        # One free cost changed within (min, max): only the fingerprint can tell.
        p = p.copy()
        flat = p.reshape(-1)
        i = int(np.flatnonzero((flat >= 3) & (flat <= 6))[0])
        flat[i] ^= 1
        return h, p

    def over_max(h, p):
        p = p.astype(np.uint32)
        p.reshape(-1)[5] = 1048576  # APGX_MAX_CELL_COST + 1
        h.cost_bytes = 4
        return h, p

    def short_plane(h, p):
        return h, p.reshape(-1)[:-1].copy()

    def reserved(h, p):
        h.reserved[3] = 1
        return h, p

    return {
        "plane byte": (plane_byte, "device view fingerprint mismatch"),
        "source hash": (field("source_board_content_hash", 6), "device view fingerprint mismatch"),
        "origin": (field("lattice_origin_x", 101), "device view fingerprint mismatch"),
        "free cells": (field("free_cells", 1), "cost summary disagrees"),
        "max cost": (field("max_cell_cost", 9), "cost summary disagrees"),
        "persistent": (field("estimated_persistent_device_bytes", 1), "persistent byte formula"),
        "schema": (field("schema_version", 1), "schema version is not supported"),
        "cost bytes": (field("cost_bytes", 3), "cost_bytes must be 2 or 4"),
        "step": (field("lattice_step", 0), "lattice step"),
        "layers": (field("layers", 17), "dimensions out of range"),
        "reserved": (reserved, "reserved bytes"),
        "plane size": (short_plane, "plane size disagrees"),
        "cost above maximum": (over_max, "exceeds APGX_MAX_CELL_COST"),
    }


@pytest.mark.parametrize("name", ["plane byte", "source hash", "origin", "free cells", "max cost", "persistent",
                                  "schema", "cost bytes", "step", "layers", "reserved", "plane size",
                                  "cost above maximum"])
def test_rejections(apgx, name):
    h, plane = _board(apgx)
    mutate, message = _mutations(apgx)[name]
    h2 = apgx.V2BoardHeader.from_buffer_copy(bytes(h))
    h2, p2 = mutate(h2, plane)
    with pytest.raises(apgx.ApgxError) as e:
        apgx.validate_board(h2, p2)
    assert message in e.value.detail
    # The upload path validates before touching any device.
    handle = C.c_void_p()
    p2 = np.ascontiguousarray(p2)
    rc = apgx.lib().apgx_lattice_create_v2(C.byref(h2), p2.ctypes.data, p2.nbytes, C.byref(handle))
    assert rc != 0 and not handle.value


def test_net_batch_record_layouts(apgx):
    """The offsets schemas/device_net_batch/v2.md documents."""
    assert C.sizeof(apgx.NetQuery) == 32
    assert (apgx.NetQuery.pin_offset.offset, apgx.NetQuery.pin_count.offset, apgx.NetQuery.window.offset) == (8, 12, 16)
    assert C.sizeof(apgx.Window) == 16
    assert C.sizeof(apgx.NetResult) == 32
    assert [getattr(apgx.NetResult, f).offset for f in ("status", "cost", "connection_count", "connection_offset",
                                                         "rounds")] == [8, 12, 16, 20, 24]
    assert C.sizeof(apgx.Connection) == 24
    assert [getattr(apgx.Connection, f).offset for f in ("pin_index", "distance", "cell_offset", "cell_count")] == \
        [0, 4, 8, 16]


@pytest.mark.gpu
def test_sealed_upload_routes_like_direct_upload(apgx, gpu):
    print("[case] v2 sealed board upload", flush=True)
    h, plane = _board(apgx, seed=9)
    nets, pins = apgx.generate_nets(plane, 24, 2, 5, 8, 32, seed=4)
    a = apgx.Lattice(plane, via_cost=20)
    b = apgx.Lattice.from_sealed(h, plane)
    ra, _ = a.route(nets, pins)
    rb, _ = b.route(nets, pins)
    assert [(r.status, r.cost) for r in ra] == [(r.status, r.cost) for r in rb]
    for x, y in zip(ra, rb):
        assert len(x.connections) == len(y.connections)
        for (pa, da, ca), (pb, db, cb) in zip(x.connections, y.connections):
            assert (pa, da) == (pb, db) and np.array_equal(ca, cb)
    a.close()
    b.close()
