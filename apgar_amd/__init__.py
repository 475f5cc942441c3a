"""apgar_amd — MI355X-native batched lattice search (ctypes view of libapgx.so).

The compute path is the HIP library ``libapgx.so`` built in-tree by
``apgar_amd.build``.  This module only mirrors its C-ABI records with ctypes and
offers thin Python conveniences over them; it never computes a route itself.
Loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libapgx.so"

APGX_OK = 0
ERROR_NAMES = {1: "unsupported", 2: "resource_exhausted", 3: "backend_failure", 4: "internal_invariant"}
UNREACHABLE = 0xFFFFFFFF
NET_REACHED, NET_DISCONNECTED, NET_BUDGET_EXHAUSTED, NET_INVALID = 0, 1, 2, 3
NEGOTIATION_FEASIBLE, NEGOTIATION_STALLED, NEGOTIATION_PASS_LIMIT = 0, 1, 2
LATTICE_SCHEMA_VERSION = 2


class ApgxError(RuntimeError):
    def __init__(self, code: int, detail: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {detail}")
        self.code = code
        self.detail = detail


class DeviceInfo(C.Structure):
    _fields_ = [("backend", C.c_char * 32), ("device_name", C.c_char * 256),
                ("gcn_arch_name", C.c_char * 64), ("compute_units", C.c_uint32),
                ("runtime_version", C.c_uint32), ("driver_version", C.c_uint32),
                ("reserved", C.c_uint32), ("global_memory_bytes", C.c_uint64)]


class LatticeDesc(C.Structure):
    _fields_ = [("schema_version", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("layers", C.c_uint32), ("cost_bytes", C.c_uint32), ("via_cost", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


class Window(C.Structure):
    _fields_ = [("x0", C.c_uint32), ("y0", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32)]


class NetQuery(C.Structure):
    _fields_ = [("net_id", C.c_uint64), ("pin_offset", C.c_uint32), ("pin_count", C.c_uint32),
                ("window", Window)]


class NetResult(C.Structure):
    _fields_ = [("net_id", C.c_uint64), ("status", C.c_uint32), ("cost", C.c_uint32),
                ("connection_count", C.c_uint32), ("connection_offset", C.c_uint32),
                ("rounds", C.c_uint32), ("reserved", C.c_uint32)]


class Connection(C.Structure):
    _fields_ = [("pin_index", C.c_uint32), ("distance", C.c_uint32), ("cell_offset", C.c_uint64),
                ("cell_count", C.c_uint32), ("reserved", C.c_uint32)]


class RouteOptions(C.Structure):
    _fields_ = [("max_rounds", C.c_uint32), ("tile_w", C.c_uint32), ("tile_h", C.c_uint32),
                ("rounds_per_sync", C.c_uint32), ("max_device_bytes", C.c_uint64),
                ("use_graph", C.c_uint32), ("profile_rounds", C.c_uint32),
                ("frontier_delta", C.c_uint32), ("max_sweeps", C.c_uint32),
                ("no_goal_bound", C.c_uint32), ("sweep_schedule", C.c_uint32), ("landmarks", C.c_uint32),
                ("landmark_weight", C.c_uint32), ("reserved", C.c_uint32)]


class RouteStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("wall_ms", C.c_double), ("tile_visits", C.c_uint64),
                ("cell_visits", C.c_uint64), ("algorithmic_bytes", C.c_uint64),
                ("relaxations", C.c_uint64), ("inner_sweeps", C.c_uint64),
                ("workspace_bytes", C.c_uint64), ("rounds", C.c_uint32), ("phases", C.c_uint32),
                ("tile_w", C.c_uint32), ("tile_h", C.c_uint32), ("host_syncs", C.c_uint32),
                ("graph_launches", C.c_uint32), ("round_kernel_ms", C.c_double),
                ("tiles_touched", C.c_uint64), ("layer_sweeps", C.c_uint64), ("full_sweeps", C.c_uint64),
                ("changing_full_sweeps", C.c_uint64), ("free_cell_visits", C.c_uint64)]


class TuneCandidate(C.Structure):
    _fields_ = [("tile_w", C.c_uint32), ("tile_h", C.c_uint32), ("frontier_delta", C.c_uint32),
                ("max_sweeps", C.c_uint32), ("device_ms", C.c_double)]


class TuneReport(C.Structure):
    _fields_ = [("sample_nets", C.c_uint32), ("candidate_count", C.c_uint32), ("best", C.c_uint32),
                ("reserved", C.c_uint32), ("candidates", TuneCandidate * 24)]


NEGOTIATION_PASS_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint32, C.c_double)


class NegotiationPolicy(C.Structure):
    _fields_ = [("maximum_passes", C.c_uint32), ("batch_nets", C.c_uint32),
                ("initial_present_cost", C.c_uint64), ("present_cost_increment", C.c_uint64),
                ("historical_cost_increment", C.c_uint64),
                ("maximum_congestion_cost_value", C.c_uint64), ("on_pass", NEGOTIATION_PASS_FN),
                ("on_pass_context", C.c_void_p), ("group_nets", C.c_uint32), ("reserved", C.c_uint32)]


class NegotiationSummary(C.Structure):
    _fields_ = [("termination", C.c_uint32), ("completed_passes", C.c_uint32),
                ("routed_nets", C.c_uint32), ("batches", C.c_uint32),
                ("total_overuse_units", C.c_uint64), ("tile_visits", C.c_uint64),
                ("search_device_ms", C.c_double), ("wall_ms", C.c_double),
                ("search_rounds", C.c_uint64), ("group_routed_nets", C.c_uint64),
                ("redone_nets", C.c_uint64)]


class BoardGenParams(C.Structure):
    _fields_ = [(k, C.c_uint32) for k in (
        "width", "height", "layers", "cost_min", "cost_max", "keepout_ppm", "keepout_min_side",
        "keepout_max_side", "nets", "min_pins", "max_pins", "bga_fields", "bga_pads", "bga_pitch",
        "pad_pin_ppm", "scatter_radius", "window_margin", "pin_clearance")] + [("seed", C.c_uint64)]


class LatticeGenParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("layers", C.c_uint32),
                ("cost_min", C.c_uint32), ("cost_max", C.c_uint32), ("blocked_ppm", C.c_uint32),
                ("keepout_ppm", C.c_uint32), ("keepout_min_side", C.c_uint32),
                ("keepout_max_side", C.c_uint32), ("reserved", C.c_uint32), ("seed", C.c_uint64)]


class PricePolicy(C.Structure):
    _fields_ = [("initial_present_factor", C.c_uint64), ("present_factor_increment", C.c_uint64),
                ("historical_price_increment", C.c_uint64)]


class PriceLimits(C.Structure):
    _fields_ = [("maximum_price_entries", C.c_uint64), ("maximum_hot_resources", C.c_uint64),
                ("maximum_price_value", C.c_uint64), ("maximum_aggregate_price", C.c_uint64)]


class PriceEpoch(C.Structure):
    _fields_ = [("outcome", C.c_uint32), ("reserved", C.c_uint32), ("outcome_cell", C.c_uint64),
                ("epoch_index", C.c_uint64), ("present_factor", C.c_uint64), ("hot_resources", C.c_uint64),
                ("price_entries", C.c_uint64), ("aggregate_price", C.c_uint64)]


class PriorConfig(C.Structure):
    _fields_ = [("lattice_layers", C.c_uint32), ("grid_h", C.c_uint32), ("grid_w", C.c_uint32),
                ("reserved", C.c_uint32), ("seed", C.c_uint64)]


class PriorTrainOptions(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32), ("learning_rate", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float),
                ("augment", C.c_uint32), ("use_graph", C.c_uint32), ("seed", C.c_uint64)]


class PriorTrainReport(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("first_epoch_loss", C.c_float), ("last_epoch_loss", C.c_float),
                ("device_ms", C.c_double), ("wall_ms", C.c_double), ("graph", C.c_uint32),
                ("reserved", C.c_uint32)]


class PriorDecode(C.Structure):
    _fields_ = [("tile", C.c_uint32), ("max_value", C.c_uint32), ("unit", C.c_float), ("gain", C.c_float)]


class PriceAssociations(C.Structure):
    _fields_ = [("board_content_hash", C.c_uint64), ("compiler_profile_fingerprint", C.c_uint64),
                ("geometry_compiler_version", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("layers", C.c_uint32)]


class PriceSnapshot(C.Structure):
    _fields_ = [("associations", PriceAssociations), ("capacity_units", C.c_uint32), ("reserved", C.c_uint32),
                ("policy_identity", C.c_uint64), ("prior_snapshot_identity", C.c_uint64),
                ("epoch_index", C.c_uint64), ("present_factor", C.c_uint64), ("price_count", C.c_uint64),
                ("snapshot_identity", C.c_uint64)]


class PriceCheck(C.Structure):
    _fields_ = [("valid", C.c_uint32), ("reserved", C.c_uint32), ("entry_index", C.c_uint64),
                ("invariant", C.c_char * 64)]


# apgx_price_entry as a numpy record
PRICE_ENTRY = np.dtype([("cell", "<u8"), ("present", "<u8"), ("history", "<u8"), ("total", "<u8")])


class V2BoardHeader(C.Structure):
    """DeviceDenseBoardV2 header (128 bytes), schemas/device_dense_board/v2.md."""
    _fields_ = [("schema_version", C.c_uint32), ("compiler_version", C.c_uint32),
                ("source_board_content_hash", C.c_uint64), ("compiler_profile_fingerprint", C.c_uint64),
                ("rule_bucket_identity", C.c_uint64), ("device_view_fingerprint", C.c_uint64),
                ("estimated_persistent_device_bytes", C.c_uint64), ("lattice_origin_x", C.c_int64),
                ("lattice_origin_y", C.c_int64), ("lattice_step", C.c_int64), ("width", C.c_uint32),
                ("height", C.c_uint32), ("layers", C.c_uint32), ("cost_bytes", C.c_uint32),
                ("via_cost", C.c_uint32), ("min_cell_cost", C.c_uint32), ("max_cell_cost", C.c_uint32),
                ("reserved0", C.c_uint32), ("free_cells", C.c_uint64), ("reserved", C.c_uint8 * 16)]


_P = C.POINTER
_SIGNATURES = {
    "apgx_device_info_query": (C.c_int, [_P(DeviceInfo)]),
    "apgx_last_error": (C.c_char_p, []),
    "apgx_abi_version": (C.c_uint32, []),
    "apgx_device_synchronize": (C.c_int, []),
    "apgx_lattice_create": (C.c_int, [_P(LatticeDesc), C.c_void_p, _P(C.c_void_p)]),
    "apgx_lattice_create_from_device": (C.c_int, [_P(LatticeDesc), C.c_void_p, _P(C.c_void_p)]),
    "apgx_v2_board_fingerprint": (C.c_uint64, [_P(V2BoardHeader), C.c_void_p]),
    "apgx_device_class": (C.c_int, [C.c_char_p, C.c_uint64]),
    "apgx_price_policy_identity": (C.c_uint64, [_P(PricePolicy)]),
    "apgx_price_snapshot_identity": (C.c_uint64, [_P(PriceSnapshot), C.c_void_p]),
    "apgx_price_snapshot_capture": (C.c_int, [_P(PriceAssociations), C.c_uint32, _P(PricePolicy), C.c_uint64,
                                              C.c_uint64, _P(C.c_uint32), _P(C.c_uint32), C.c_void_p, C.c_uint64,
                                              _P(PriceSnapshot)]),
    "apgx_price_snapshot_validate": (C.c_int, [_P(PriceSnapshot), C.c_void_p, _P(PriceAssociations), C.c_uint32,
                                               _P(PricePolicy), _P(PriceLimits), C.c_uint64, C.c_uint64,
                                               _P(PriceCheck)]),
    "apgx_price_snapshot_restore": (C.c_int, [_P(PriceSnapshot), C.c_void_p, _P(C.c_uint32), _P(C.c_uint32)]),
    "apgx_v2_board_seal": (C.c_int, [_P(V2BoardHeader), C.c_void_p, C.c_uint64]),
    "apgx_v2_board_validate": (C.c_int, [_P(V2BoardHeader), C.c_void_p, C.c_uint64]),
    "apgx_lattice_create_v2": (C.c_int, [_P(V2BoardHeader), C.c_void_p, C.c_uint64, _P(C.c_void_p)]),
    "apgx_lattice_release": (None, [C.c_void_p]),
    "apgx_lattice_device_costs": (C.c_void_p, [C.c_void_p]),
    "apgx_lattice_build_landmarks": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double)]),
    "apgx_lattice_landmark_count": (C.c_uint32, [C.c_void_p]),
    "apgx_lattice_landmark_field": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_uint32), C.c_uint64]),
    "apgx_route_nets": (C.c_int, [C.c_void_p, _P(NetQuery), C.c_uint32, _P(C.c_uint32), C.c_uint64,
                                  _P(RouteOptions), _P(C.c_void_p)]),
    "apgx_search_sizes": (C.c_int, [C.c_void_p, _P(C.c_uint64), _P(C.c_uint64)]),
    "apgx_search_readback": (C.c_int, [C.c_void_p, _P(NetResult), _P(Connection), C.c_uint64,
                                       _P(C.c_uint32), C.c_uint64]),
    "apgx_search_stats": (C.c_int, [C.c_void_p, _P(RouteStats)]),
    "apgx_search_release": (None, [C.c_void_p]),
    "apgx_generate_board": (C.c_int, [_P(BoardGenParams), _P(C.c_uint32), _P(NetQuery), _P(C.c_uint32),
                                      _P(C.c_uint64)]),
    "apgx_negotiation_policy_defaults": (None, [_P(NegotiationPolicy)]),
    "apgx_negotiate": (C.c_int, [C.c_void_p, _P(C.c_uint8), _P(NetQuery), C.c_uint32, _P(C.c_uint32),
                                 C.c_uint64, _P(NegotiationPolicy), _P(C.c_uint32), _P(RouteOptions),
                                 _P(C.c_void_p)]),
    "apgx_negotiation_get_summary": (C.c_int, [C.c_void_p, _P(NegotiationSummary), _P(C.c_uint64),
                                               C.c_uint32]),
    "apgx_negotiation_planes": (C.c_int, [C.c_void_p, _P(C.c_uint32), _P(C.c_uint32)]),
    "apgx_negotiation_routes": (C.c_int, [C.c_void_p, _P(C.c_uint32), _P(C.c_uint32), C.c_uint64,
                                          _P(C.c_uint64)]),
    "apgx_negotiation_release": (None, [C.c_void_p]),
    "apgx_autotune": (C.c_int, [C.c_void_p, _P(NetQuery), C.c_uint32, _P(C.c_uint32), C.c_uint64,
                                C.c_uint32, _P(RouteOptions), _P(RouteOptions), _P(TuneReport)]),
    "apgx_price_defaults": (None, [_P(PricePolicy), _P(PriceLimits)]),
    "apgx_price_update": (C.c_int, [C.c_uint64, _P(PricePolicy), _P(PriceLimits), C.c_uint64, _P(C.c_uint32),
                                    _P(C.c_uint8), _P(C.c_uint32), _P(C.c_uint32), _P(C.c_uint32),
                                    _P(PriceEpoch)]),
    "apgx_price_update_no_alloc": (C.c_int, [C.c_uint64, _P(PricePolicy), _P(PriceLimits), C.c_uint64, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _P(PriceEpoch)]),
    "apgx_prior_create": (C.c_int, [_P(PriorConfig), _P(C.c_void_p)]),
    "apgx_prior_release": (None, [C.c_void_p]),
    "apgx_prior_param_count": (C.c_uint64, [C.c_void_p]),
    "apgx_prior_get_params": (C.c_int, [C.c_void_p, _P(C.c_float), C.c_uint64]),
    "apgx_prior_set_params": (C.c_int, [C.c_void_p, _P(C.c_float), C.c_uint64]),
    "apgx_prior_forward": (C.c_int, [C.c_void_p, _P(C.c_float), C.c_uint32, _P(C.c_float)]),
    "apgx_prior_gradients": (C.c_int, [C.c_void_p, _P(C.c_float), _P(C.c_float), C.c_uint32,
                                       _P(C.c_float), _P(C.c_float)]),
    "apgx_prior_train_defaults": (None, [_P(PriorTrainOptions)]),
    "apgx_prior_train": (C.c_int, [C.c_void_p, _P(C.c_float), _P(C.c_float), C.c_uint32,
                                   _P(PriorTrainOptions), _P(PriorTrainReport)]),
    "apgx_prior_features": (C.c_int, [C.c_void_p, _P(C.c_uint32), C.c_uint64, C.c_uint32, _P(C.c_float)]),
    "apgx_prior_pool_history": (C.c_int, [_P(C.c_uint32), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                          C.c_float, _P(C.c_float)]),
    "apgx_prior_seed_history": (C.c_int, [C.c_void_p, C.c_void_p, _P(C.c_uint32), C.c_uint64,
                                          _P(PriorDecode), _P(C.c_uint32)]),
    "apgx_generate_lattice": (C.c_int, [_P(LatticeGenParams), _P(C.c_uint32)]),
    "apgx_generate_pairs": (C.c_int, [_P(C.c_uint32), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint64, _P(C.c_uint32)]),
    "apgx_generate_nets": (C.c_int, [_P(C.c_uint32), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                     _P(NetQuery), _P(C.c_uint32), _P(C.c_uint64)]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads libapgx.so (raises if it has not been built: there is no fallback)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m apgar_amd.build`")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def check(code: int) -> None:
    if code != APGX_OK:
        raise ApgxError(code, lib().apgx_last_error().decode(errors="replace"))


def u32p(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_P(C.c_uint32))


def device_class() -> str:
    """Candidate device class of the current device, e.g. "hip-gfx950"."""
    buf = C.create_string_buffer(64)
    check(lib().apgx_device_class(buf, len(buf)))
    return buf.value.decode()


def synchronize() -> None:
    check(lib().apgx_device_synchronize())


def device_info() -> DeviceInfo:
    info = DeviceInfo()
    check(lib().apgx_device_info_query(C.byref(info)))
    return info


# ----------------------------------------------------------------- generators
def generate_lattice(width, height, layers, cost_min, cost_max, blocked_ppm=0, keepout_ppm=0,
                     keepout_min_side=0, keepout_max_side=0, seed=0) -> np.ndarray:
    """Deterministic cost plane (uint32, 0 = blocked), shape (layers, height, width)."""
    p = LatticeGenParams(width, height, layers, cost_min, cost_max, blocked_ppm, keepout_ppm,
                         keepout_min_side, keepout_max_side, 0, seed)
    out = np.empty(width * height * layers, dtype=np.uint32)
    check(lib().apgx_generate_lattice(C.byref(p), u32p(out)))
    return out.reshape(layers, height, width)


def generate_pairs(cost: np.ndarray, count: int, seed: int) -> np.ndarray:
    L, H, W = cost.shape
    flat = np.ascontiguousarray(cost.reshape(-1), dtype=np.uint32)
    out = np.empty(2 * count, dtype=np.uint32)
    check(lib().apgx_generate_pairs(u32p(flat), W, H, L, count, seed, u32p(out)))
    return out.reshape(count, 2)


def generate_nets(cost: np.ndarray, count: int, min_pins: int, max_pins: int, min_side: int,
                  max_side: int, seed: int):
    L, H, W = cost.shape
    flat = np.ascontiguousarray(cost.reshape(-1), dtype=np.uint32)
    nets = (NetQuery * count)()
    pins = np.empty(count * max_pins, dtype=np.uint32)
    total = C.c_uint64(0)
    check(lib().apgx_generate_nets(u32p(flat), W, H, L, count, min_pins, max_pins, min_side,
                                   max_side, seed, nets, u32p(pins), C.byref(total)))
    return nets, pins[: total.value].copy()


def generate_board(**params):
    """Allocator board (apgx_generate_board): returns (cost (L, H, W) uint32, nets, pins)."""
    p = BoardGenParams(**params)
    cost = np.empty(p.width * p.height * p.layers, dtype=np.uint32)
    nets = (NetQuery * p.nets)()
    pins = np.empty(p.nets * p.max_pins, dtype=np.uint32)
    total = C.c_uint64(0)
    check(lib().apgx_generate_board(C.byref(p), u32p(cost), nets, u32p(pins), C.byref(total)))
    return cost.reshape(p.layers, p.height, p.width), nets, pins[:total.value].copy()


# ------------------------------------------------------------ sealed V2 boards
def seal_board(cost: np.ndarray, via_cost: int = 0, *, cost_bytes: int | None = None, compiler_version: int = 1,
               source_board_content_hash: int = 0, compiler_profile_fingerprint: int = 0,
               rule_bucket_identity: int = 0, origin=(0, 0), step: int = 1):
    """Seals a (layers, height, width) cost plane as DeviceDenseBoardV2:
    returns (header, plane) with the derived fields and fingerprint filled."""
    L, H, W = cost.shape
    if cost_bytes is None:
        cost_bytes = 2 if int(cost.max(initial=0)) <= 0xFFFF else 4
    plane = np.ascontiguousarray(cost, dtype=np.uint16 if cost_bytes == 2 else np.uint32)
    h = V2BoardHeader()
    h.schema_version, h.compiler_version = LATTICE_SCHEMA_VERSION, compiler_version
    h.source_board_content_hash = source_board_content_hash
    h.compiler_profile_fingerprint = compiler_profile_fingerprint
    h.rule_bucket_identity = rule_bucket_identity
    h.lattice_origin_x, h.lattice_origin_y = origin
    h.lattice_step = step
    h.width, h.height, h.layers, h.cost_bytes, h.via_cost = W, H, L, cost_bytes, via_cost
    check(lib().apgx_v2_board_seal(C.byref(h), plane.ctypes.data, plane.nbytes))
    return h, plane


def validate_board(header: V2BoardHeader, plane: np.ndarray) -> None:
    """Raises ApgxError unless the sealed board is intact."""
    plane = np.ascontiguousarray(plane)
    check(lib().apgx_v2_board_validate(C.byref(header), plane.ctypes.data, plane.nbytes))


def board_bytes(header: V2BoardHeader, plane: np.ndarray) -> bytes:
    """Wire form: the 128-byte header followed by the plane."""
    return bytes(header) + np.ascontiguousarray(plane).tobytes()


def parse_board_bytes(data: bytes):
    """(header, plane) of a wire-form board; the plane's dtype and shape come
    from the header (nothing is validated here)."""
    h = V2BoardHeader.from_buffer_copy(data[:C.sizeof(V2BoardHeader)])
    dt = np.uint16 if h.cost_bytes == 2 else np.uint32
    plane = np.frombuffer(data, dtype=dt, offset=C.sizeof(V2BoardHeader))
    if plane.size == h.layers * h.height * h.width:
        plane = plane.reshape(h.layers, h.height, h.width)
    return h, plane


# --------------------------------------------------------------------- search
@dataclass
class NetRoute:
    net_id: int
    status: int
    cost: int
    rounds: int
    connections: list  # (pin_index, distance, cells ndarray)


class Lattice:
    """Owns one uploaded cost plane (uint16 or uint32) on the current device."""

    def __init__(self, cost: np.ndarray, via_cost: int = 0, cost_bytes: int | None = None):
        L, H, W = cost.shape
        if cost_bytes is None:
            cost_bytes = 2 if int(cost.max(initial=0)) <= 0xFFFF else 4
        plane = np.ascontiguousarray(cost, dtype=np.uint16 if cost_bytes == 2 else np.uint32)
        self.shape = (L, H, W)
        self.via_cost = via_cost
        self.cost_bytes = cost_bytes
        self.desc = LatticeDesc(LATTICE_SCHEMA_VERSION, W, H, L, cost_bytes, via_cost, (C.c_uint32 * 2)())
        handle = C.c_void_p()
        check(lib().apgx_lattice_create(C.byref(self.desc), plane.ctypes.data, C.byref(handle)))
        self.handle = handle

    @classmethod
    def from_sealed(cls, header: V2BoardHeader, plane: np.ndarray) -> "Lattice":
        """Uploads a sealed DeviceDenseBoardV2 (validated first)."""
        self = cls.__new__(cls)
        plane = np.ascontiguousarray(plane)
        self.shape = (header.layers, header.height, header.width)
        self.via_cost = header.via_cost
        self.cost_bytes = header.cost_bytes
        self.desc = LatticeDesc(LATTICE_SCHEMA_VERSION, header.width, header.height, header.layers,
                                header.cost_bytes, header.via_cost, (C.c_uint32 * 2)())
        handle = C.c_void_p()
        self.handle = None
        check(lib().apgx_lattice_create_v2(C.byref(header), plane.ctypes.data, plane.nbytes, C.byref(handle)))
        self.handle = handle
        return self

    def close(self):
        if self.handle:
            lib().apgx_lattice_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def route(self, nets, pins: np.ndarray, *, max_rounds=0, tile=(0, 0), rounds_per_sync=0,
              max_device_bytes=0, use_graph=0, profile_rounds=0, frontier_delta=0, max_sweeps=0,
              goal_bound=True, sweep_schedule=0, landmarks=False, landmark_weight=0):
        """Routes nets (a ctypes NetQuery array) and returns (routes, stats).
        landmarks=True adds the ALT bound of the fields build_landmarks made."""
        pins = np.ascontiguousarray(pins, dtype=np.uint32)
        opts = RouteOptions(max_rounds, tile[0], tile[1], rounds_per_sync, max_device_bytes,
                            use_graph, profile_rounds, frontier_delta, max_sweeps,
                            0 if goal_bound else 1, sweep_schedule, 1 if landmarks else 0, landmark_weight)
        search = C.c_void_p()
        check(lib().apgx_route_nets(self.handle, nets, len(nets), u32p(pins), pins.size,
                                    C.byref(opts), C.byref(search)))
        try:
            nconn, ncell = C.c_uint64(0), C.c_uint64(0)
            check(lib().apgx_search_sizes(search, C.byref(nconn), C.byref(ncell)))
            results = (NetResult * max(len(nets), 1))()
            conns = (Connection * max(nconn.value, 1))()
            cells = np.empty(max(ncell.value, 1), dtype=np.uint32)
            check(lib().apgx_search_readback(search, results, conns, nconn.value, u32p(cells),
                                             ncell.value))
            stats = RouteStats()
            check(lib().apgx_search_stats(search, C.byref(stats)))
        finally:
            lib().apgx_search_release(search)
        routes = []
        for i in range(len(nets)):
            r = results[i]
            cs = []
            for c in range(r.connection_count):
                k = conns[r.connection_offset + c]
                cs.append((k.pin_index, k.distance,
                           cells[k.cell_offset:k.cell_offset + k.cell_count].copy()))
            routes.append(NetRoute(r.net_id, r.status, r.cost, r.rounds, cs))
        return routes, stats

    def build_landmarks(self, count: int = 8) -> float:
        """Builds `count` landmark distance fields (apgx_lattice_build_landmarks);
        returns the build's device milliseconds."""
        ms = C.c_double(0.0)
        check(lib().apgx_lattice_build_landmarks(self.handle, count, C.byref(ms)))
        return ms.value

    @property
    def landmark_count(self) -> int:
        return int(lib().apgx_lattice_landmark_count(self.handle))

    def landmark_field(self, index: int) -> np.ndarray:
        """d(landmark index, cell) as an (L, H, W) uint32 array (0x1FFFFFFF:
        blocked or unreachable)."""
        L, H, W = self.shape
        out = np.empty(L * H * W, dtype=np.uint32)
        check(lib().apgx_lattice_landmark_field(self.handle, index, u32p(out), out.size))
        return out.reshape(L, H, W)

    def autotune(self, nets, pins: np.ndarray, sample_nets=256, **base):
        """Picks tile shape and schedule on a strided sample of the nets.

        Returns (options, report): options is a dict of route() keyword
        arguments, report a list of (tile_w, tile_h, frontier_delta,
        max_sweeps, device_ms) per candidate with the chosen index."""
        pins = np.ascontiguousarray(pins, dtype=np.uint32)
        b = RouteOptions()
        for k, v in base.items():
            if k == "tile":
                b.tile_w, b.tile_h = v
            elif k == "goal_bound":
                b.no_goal_bound = 0 if v else 1
            elif k == "landmarks":
                b.landmarks = 1 if v else 0
            else:
                setattr(b, k, v)
        out, rep = RouteOptions(), TuneReport()
        check(lib().apgx_autotune(self.handle, nets, len(nets), u32p(pins), pins.size, sample_nets,
                                  C.byref(b), C.byref(out), C.byref(rep)))
        opts = dict(base)
        opts.update(tile=(out.tile_w, out.tile_h), frontier_delta=out.frontier_delta,
                    max_sweeps=out.max_sweeps)
        cands = [(c.tile_w, c.tile_h, c.frontier_delta, c.max_sweeps, c.device_ms)
                 for c in rep.candidates[:rep.candidate_count]]
        return opts, {"sample_nets": rep.sample_nets, "best": rep.best, "candidates": cands}

    def negotiate(self, nets, pins: np.ndarray, *, capacity=None, initial_history=None,
                  search_options=None, **policy):
        """Negotiated-congestion allocator (apgx_negotiate).  policy keywords:
        maximum_passes, batch_nets, initial_present_cost, present_cost_increment,
        historical_cost_increment, maximum_congestion_cost_value, group_nets, and on_pass
        (a callable (pass, overuse, routed nets, wall ms) run after every pass).  Returns a dict
        like the oracle's: termination, passes, routed, overuse, overuse_per_pass,
        history / usage planes (L, H, W) and per-net tree cells (None: unrouted)."""
        pins = np.ascontiguousarray(pins, dtype=np.uint32)
        pol = NegotiationPolicy()
        lib().apgx_negotiation_policy_defaults(C.byref(pol))
        on_pass = policy.pop("on_pass", None)
        for k, v in policy.items():
            setattr(pol, k, v)
        if on_pass is not None:
            # Kept alive for the call: on_pass(pass, overuse, routed, wall_ms).
            pol.on_pass = NEGOTIATION_PASS_FN(lambda _ctx, p, o, r, ms: on_pass(p, o, r, ms))
        L, H, W = self.shape
        n = L * H * W
        cap = None if capacity is None else np.ascontiguousarray(capacity, dtype=np.uint8).reshape(-1)
        init = (None if initial_history is None
                else np.ascontiguousarray(initial_history, dtype=np.uint32).reshape(-1))
        opts = RouteOptions()
        for k, v in (search_options or {}).items():
            if k == "tile":
                opts.tile_w, opts.tile_h = v
            else:
                setattr(opts, k, v)
        handle = C.c_void_p()
        check(lib().apgx_negotiate(self.handle,
                                   cap.ctypes.data_as(_P(C.c_uint8)) if cap is not None else None,
                                   nets, len(nets), u32p(pins), pins.size, C.byref(pol),
                                   u32p(init) if init is not None else None, C.byref(opts),
                                   C.byref(handle)))
        try:
            summ = NegotiationSummary()
            per_pass = np.zeros(max(pol.maximum_passes, 1), dtype=np.uint64)
            check(lib().apgx_negotiation_get_summary(handle, C.byref(summ),
                                                     per_pass.ctypes.data_as(_P(C.c_uint64)),
                                                     per_pass.size))
            hist = np.empty(n, dtype=np.uint32)
            usage = np.empty(n, dtype=np.uint32)
            check(lib().apgx_negotiation_planes(handle, u32p(hist), u32p(usage)))
            total = C.c_uint64(0)
            rlen = np.zeros(max(len(nets), 1), dtype=np.uint32)
            check(lib().apgx_negotiation_routes(handle, u32p(rlen), None, 0, C.byref(total)))
            cells = np.empty(max(total.value, 1), dtype=np.uint32)
            check(lib().apgx_negotiation_routes(handle, u32p(rlen), u32p(cells), cells.size,
                                                C.byref(total)))
        finally:
            lib().apgx_negotiation_release(handle)
        routes, off = [], 0
        for i in range(len(nets)):
            k = int(rlen[i])
            routes.append(cells[off:off + k].copy() if k else None)
            off += k
        return {"termination": summ.termination, "passes": summ.completed_passes,
                "routed": summ.routed_nets, "overuse": int(summ.total_overuse_units),
                "overuse_per_pass": [int(v) for v in per_pass[:summ.completed_passes]],
                "batches": summ.batches, "tile_visits": summ.tile_visits,
                "search_device_ms": summ.search_device_ms, "wall_ms": summ.wall_ms,
                "search_rounds": summ.search_rounds, "group_routed_nets": summ.group_routed_nets,
                "redone_nets": summ.redone_nets,
                "history": hist.reshape(L, H, W), "occupancy": usage.reshape(L, H, W),
                "routes": routes}

    def route_pairs(self, pairs: np.ndarray, window=None, **kw):
        """Single-pair queries (pairs: (n, 2) source/target cells)."""
        n = len(pairs)
        nets = (NetQuery * n)()
        for i in range(n):
            nets[i].net_id = i + 1
            nets[i].pin_offset = 2 * i
            nets[i].pin_count = 2
            if window is not None:
                nets[i].window = Window(*window)
        return self.route(nets, np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1), **kw)


def nets_from_lists(pin_lists, windows=None):
    """Builds a NetQuery array and flat pin array from python lists."""
    nets = (NetQuery * len(pin_lists))()
    flat = []
    for i, pl in enumerate(pin_lists):
        nets[i].net_id = i + 1
        nets[i].pin_offset = len(flat)
        nets[i].pin_count = len(pl)
        if windows is not None and windows[i] is not None:
            nets[i].window = Window(*windows[i])
        flat.extend(int(p) for p in pl)
    return nets, np.asarray(flat, dtype=np.uint32)


def _f32p(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_P(C.c_float))


class Prior:
    """Learned history-cost prior (apgx_prior_*): a conv encoder-decoder over
    per-tile board maps, trained and run by the library's bf16 MFMA kernels."""

    TILE = 64
    UNIT = 10.0

    def __init__(self, lattice_layers: int, grid_h: int, grid_w: int, seed: int = 1):
        self.layers, self.grid = lattice_layers, (grid_h, grid_w)
        cfg = PriorConfig(lattice_layers=lattice_layers, grid_h=grid_h, grid_w=grid_w, seed=seed)
        h = C.c_void_p()
        check(lib().apgx_prior_create(C.byref(cfg), C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().apgx_prior_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def param_count(self) -> int:
        return int(lib().apgx_prior_param_count(self.handle))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.param_count, dtype=np.float32)
        check(lib().apgx_prior_get_params(self.handle, _f32p(out), out.size))
        return out

    def set_params(self, params: np.ndarray):
        params = np.ascontiguousarray(params, dtype=np.float32)
        check(lib().apgx_prior_set_params(self.handle, _f32p(params), params.size))

    def forward(self, inputs: np.ndarray) -> np.ndarray:
        """inputs [n][2L][gh][gw] -> outputs [n][L][gh][gw]."""
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        n = x.shape[0]
        out = np.empty((n, self.layers) + self.grid, dtype=np.float32)
        check(lib().apgx_prior_forward(self.handle, _f32p(x), n, _f32p(out)))
        return out

    def gradients(self, inputs: np.ndarray, targets: np.ndarray):
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        g = np.empty(self.param_count, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        check(lib().apgx_prior_gradients(self.handle, _f32p(x), _f32p(t), x.shape[0], _f32p(loss), _f32p(g)))
        return float(loss[0]), g

    def train(self, inputs: np.ndarray, targets: np.ndarray, **options) -> dict:
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        o = PriorTrainOptions()
        lib().apgx_prior_train_defaults(C.byref(o))
        for k, v in options.items():
            setattr(o, k, v)
        rep = PriorTrainReport()
        check(lib().apgx_prior_train(self.handle, _f32p(x), _f32p(t), x.shape[0], C.byref(o), C.byref(rep)))
        return {k: getattr(rep, k) for k, _ in PriorTrainReport._fields_ if k != "reserved"}

    def seed_history(self, lattice: "Lattice", pins: np.ndarray, gain: float = 1.0,
                     max_value: int = 10_000_000) -> np.ndarray:
        pins = np.ascontiguousarray(pins, dtype=np.uint32)
        L, H, W = lattice.shape
        out = np.empty((L, H, W), dtype=np.uint32)
        dec = PriorDecode(tile=self.TILE, max_value=max_value, unit=self.UNIT, gain=gain)
        check(lib().apgx_prior_seed_history(self.handle, lattice.handle, u32p(pins), pins.size, C.byref(dec),
                                            u32p(out)))
        return out


def prior_features(lattice: "Lattice", pins: np.ndarray, tile: int = 64) -> np.ndarray:
    """Input planes [2L][gh][gw] of one board (pin density, blocked fraction per tile)."""
    pins = np.ascontiguousarray(pins, dtype=np.uint32)
    L, H, W = lattice.shape
    out = np.empty((2 * L, -(-H // tile), -(-W // tile)), dtype=np.float32)
    check(lib().apgx_prior_features(lattice.handle, u32p(pins), pins.size, tile, _f32p(out)))
    return out


def prior_targets(history: np.ndarray, tile: int = 64, unit: float = 10.0) -> np.ndarray:
    """Training targets [L][gh][gw] = log1p(tile mean history / unit)."""
    h = np.ascontiguousarray(history, dtype=np.uint32)
    L, H, W = h.shape
    out = np.empty((L, -(-H // tile), -(-W // tile)), dtype=np.float32)
    check(lib().apgx_prior_pool_history(u32p(h), W, H, L, tile, unit, _f32p(out)))
    return out


def price_update(epoch: int, usage: np.ndarray, history: np.ndarray, capacity=None, **settings):
    """Candidate-allocator price epoch (apgx_price_update) on host planes.
    settings: apgx_price_policy / apgx_price_limits fields.  Returns (epoch
    summary dict, present, history, total); planes are unchanged copies of
    the input history (present / total zero) when the epoch fails."""
    pol, lim = PricePolicy(), PriceLimits()
    lib().apgx_price_defaults(C.byref(pol), C.byref(lim))
    for k, v in settings.items():
        setattr(pol if hasattr(pol, k) else lim, k, v)
    u = np.ascontiguousarray(usage, dtype=np.uint32).reshape(-1)
    h = np.ascontiguousarray(history, dtype=np.uint32).reshape(-1).copy()
    cap = None if capacity is None else np.ascontiguousarray(capacity, dtype=np.uint8).reshape(-1)
    present = np.zeros_like(h)
    total = np.zeros_like(h)
    out = PriceEpoch()
    check(lib().apgx_price_update(epoch, C.byref(pol), C.byref(lim), u.size, u32p(u),
                                  cap.ctypes.data_as(_P(C.c_uint8)) if cap is not None else None,
                                  u32p(h), u32p(present), u32p(total), C.byref(out)))
    summary = {k: getattr(out, k) for k, _ in PriceEpoch._fields_ if k != "reserved"}
    return summary, present, h, total


def price_update_no_alloc(epoch: int, cells: int, usage, history, capacity=None, present=None, total=None,
                          **settings):
    """The no-allocation price epoch (apgx_price_update_no_alloc) on DEVICE
    planes given as integer device addresses (history updated in place).
    Returns the epoch summary dict."""
    pol, lim = _price_policy(**settings)
    out = PriceEpoch()
    check(lib().apgx_price_update_no_alloc(epoch, C.byref(pol), C.byref(lim), cells, usage, capacity, history,
                                           present, total, C.byref(out)))
    return {k: getattr(out, k) for k, _ in PriceEpoch._fields_ if k != "reserved"}


def _price_policy(**settings):
    pol, lim = PricePolicy(), PriceLimits()
    lib().apgx_price_defaults(C.byref(pol), C.byref(lim))
    for k, v in settings.items():
        setattr(pol if hasattr(pol, k) else lim, k, v)
    return pol, lim


def price_policy_identity(**policy) -> int:
    return lib().apgx_price_policy_identity(C.byref(_price_policy(**policy)[0]))


def price_snapshot_capture(associations: PriceAssociations, capacity_units: int, epoch: int, prior_identity: int,
                           present: np.ndarray, history: np.ndarray, **policy):
    """Captures a price snapshot of host u32 planes; returns (snapshot,
    entries as a PRICE_ENTRY array)."""
    pol, _ = _price_policy(**policy)
    p = np.ascontiguousarray(present, dtype=np.uint32).reshape(-1)
    h = np.ascontiguousarray(history, dtype=np.uint32).reshape(-1)
    snap = PriceSnapshot()
    check(lib().apgx_price_snapshot_capture(C.byref(associations), capacity_units, C.byref(pol), epoch,
                                            prior_identity, u32p(p), u32p(h), None, 0, C.byref(snap)))
    entries = np.zeros(max(snap.price_count, 1), PRICE_ENTRY)
    check(lib().apgx_price_snapshot_capture(C.byref(associations), capacity_units, C.byref(pol), epoch,
                                            prior_identity, u32p(p), u32p(h), entries.ctypes.data, entries.size,
                                            C.byref(snap)))
    return snap, entries[: snap.price_count]


def price_snapshot_validate(snapshot: PriceSnapshot, entries: np.ndarray, expected: PriceAssociations,
                            capacity_units: int, *, maximum_epoch_index=1024, maximum_overuse_per_epoch=1 << 32,
                            **settings):
    """ValidatePriorSnapshot on a snapshot; returns (valid, invariant id, entry index)."""
    pol, lim = _price_policy(**settings)
    e = np.ascontiguousarray(entries, dtype=PRICE_ENTRY)
    out = PriceCheck()
    check(lib().apgx_price_snapshot_validate(C.byref(snapshot), e.ctypes.data if e.size else None, C.byref(expected),
                                             capacity_units, C.byref(pol), C.byref(lim), maximum_epoch_index,
                                             maximum_overuse_per_epoch, C.byref(out)))
    return bool(out.valid), out.invariant.decode(), out.entry_index


def price_snapshot_identity(snapshot: PriceSnapshot, entries: np.ndarray) -> int:
    e = np.ascontiguousarray(entries, dtype=PRICE_ENTRY)
    return lib().apgx_price_snapshot_identity(C.byref(snapshot), e.ctypes.data if e.size else None)


def price_snapshot_restore(snapshot: PriceSnapshot, entries: np.ndarray):
    """Dense (present, history) u32 planes of a snapshot."""
    a = snapshot.associations
    cells = a.width * a.height * a.layers
    present = np.zeros(cells, np.uint32)
    history = np.zeros(cells, np.uint32)
    e = np.ascontiguousarray(entries, dtype=PRICE_ENTRY)
    check(lib().apgx_price_snapshot_restore(C.byref(snapshot), e.ctypes.data if e.size else None, u32p(present),
                                            u32p(history)))
    return present, history
