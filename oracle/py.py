"""ctypes loader for the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Imported only by tests/, the smoke check and the benchmark's CPU-baseline leg,
where it is the checker (or the timed CPU baseline); the product never imports
it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
UNREACHABLE = 0xFFFFFFFF


class OrcLattice(C.Structure):
    _fields_ = [("W", C.c_uint32), ("H", C.c_uint32), ("L", C.c_uint32),
                ("cost", C.POINTER(C.c_uint32)), ("via", C.c_uint32), ("wx0", C.c_uint32),
                ("wy0", C.c_uint32), ("ww", C.c_uint32), ("wh", C.c_uint32),
                ("via_pays_cell", C.c_uint32)]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        h.orc_dense_dijkstra.restype = C.c_int
        h.orc_dense_dijkstra.argtypes = [P(OrcLattice), P(C.c_uint32), C.c_uint32, P(C.c_uint32)]
        h.orc_dense_canonical_path.restype = C.c_int64
        h.orc_dense_canonical_path.argtypes = [P(OrcLattice), P(C.c_uint32), C.c_uint32,
                                               P(C.c_uint32), C.c_uint64]
        h.orc_dense_astar.restype = C.c_uint32
        h.orc_dense_astar.argtypes = [P(OrcLattice), C.c_uint32, C.c_uint32, P(C.c_uint32),
                                      C.c_uint64, P(C.c_uint64), P(C.c_uint64)]
        h.orc_dense_route_net.restype = C.c_int
        h.orc_dense_route_net.argtypes = [P(OrcLattice), P(C.c_uint32), C.c_uint32,
                                          P(C.c_uint32), P(C.c_uint32), P(C.c_uint32),
                                          P(C.c_uint32), C.c_uint64, P(C.c_uint64)]
        _lib = h
    return _lib


def _u32(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Dense:
    """Dense lattice oracle over a (L, H, W) uint32 cost array (0 = blocked)."""

    def __init__(self, cost: np.ndarray, via: int, window=None, via_pays_cell=False):
        self.cost = np.ascontiguousarray(cost, dtype=np.uint32)
        L, H, W = self.cost.shape
        self.shape = (L, H, W)
        wx0, wy0, ww, wh = window if window is not None else (0, 0, 0, 0)
        self.g = OrcLattice(W, H, L, _u32(self.cost), via, wx0, wy0, ww, wh, 1 if via_pays_cell else 0)

    @property
    def n(self):
        L, H, W = self.shape
        return L * H * W

    def dijkstra(self, sources) -> np.ndarray:
        src = np.ascontiguousarray(np.atleast_1d(sources), dtype=np.uint32)
        dist = np.empty(self.n, dtype=np.uint32)
        rc = lib().orc_dense_dijkstra(C.byref(self.g), _u32(src), src.size, _u32(dist))
        if rc != 0:
            raise RuntimeError("oracle dijkstra rejected its input")
        return dist

    def canonical_path(self, dist: np.ndarray, target: int) -> np.ndarray:
        out = np.empty(self.n + 1, dtype=np.uint32)
        k = lib().orc_dense_canonical_path(C.byref(self.g), _u32(dist), target, _u32(out), out.size)
        if k < 0:
            raise RuntimeError("oracle canonical path failed")
        return out[:k].copy()

    def astar(self, source: int, target: int, want_path=False):
        cap = self.n + 1 if want_path else 0
        path = np.empty(max(cap, 1), dtype=np.uint32)
        plen, expd = C.c_uint64(0), C.c_uint64(0)
        d = lib().orc_dense_astar(C.byref(self.g), source, target,
                                  _u32(path) if want_path else None, cap, C.byref(plen), C.byref(expd))
        return d, (path[:plen.value].copy() if want_path else None), expd.value

    def route_net(self, pins):
        pins = np.ascontiguousarray(pins, dtype=np.uint32)
        k = pins.size
        cp, cd, cl = (np.zeros(k, dtype=np.uint32) for _ in range(3))
        cells = np.empty(self.n * 2 + 16, dtype=np.uint32)
        used = C.c_uint64(0)
        rc = lib().orc_dense_route_net(C.byref(self.g), _u32(pins), k, _u32(cp), _u32(cd), _u32(cl),
                                       _u32(cells), cells.size, C.byref(used))
        if rc < 0:
            raise RuntimeError("oracle route_net failed")
        conns = []
        off = 0
        n_done = k - 1 if rc == 0 else 0
        if rc == 0:
            for i in range(n_done):
                conns.append((int(cp[i]), int(cd[i]), cells[off:off + cl[i]].copy()))
                off += int(cl[i])
        return rc, conns


# ------------------------------------------------------ negotiated allocator
class NegoPolicy(C.Structure):
    _fields_ = [("maximum_passes", C.c_uint32), ("batch_nets", C.c_uint32),
                ("initial_present_cost", C.c_uint64), ("present_cost_increment", C.c_uint64),
                ("historical_cost_increment", C.c_uint64),
                ("maximum_congestion_cost_value", C.c_uint64), ("maximum_cell_cost", C.c_uint64)]


class NegoNet(C.Structure):
    _fields_ = [("pin_offset", C.c_uint32), ("pin_count", C.c_uint32), ("wx0", C.c_uint32),
                ("wy0", C.c_uint32), ("ww", C.c_uint32), ("wh", C.c_uint32)]


class NegoResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("termination", C.c_uint32), ("completed_passes", C.c_uint32),
                ("routed_nets", C.c_uint32), ("total_overuse_units", C.c_uint64)]


def negotiate(cost, via, nets, pins, *, capacity=None, initial_history=None, maximum_passes=8,
              batch_nets=1, initial_present_cost=0, present_cost_increment=100,
              historical_cost_increment=10, maximum_congestion_cost_value=10_000_000,
              maximum_cell_cost=(1 << 20) - 1):
    """Runs the allocator oracle.  nets: list of (pin_offset, pin_count, window or None).
    Returns a dict with status, termination, passes, overuse per pass, history,
    occupancy and per-net tree cell arrays (None for unrouted nets)."""
    h = lib()
    if not hasattr(h, "_nego_ready"):
        P = C.POINTER
        h.orc_negotiate.restype = C.c_int
        h.orc_negotiate.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_uint32), C.c_uint32,
                                    P(C.c_uint8), P(NegoNet), C.c_uint32, P(C.c_uint32),
                                    P(NegoPolicy), P(C.c_uint32), P(C.c_uint32), P(C.c_uint32),
                                    P(C.c_uint64), P(C.c_uint32), P(C.c_uint32), C.c_uint64,
                                    P(NegoResult)]
        h._nego_ready = True
    cost = np.ascontiguousarray(cost, dtype=np.uint32)
    L, H, W = cost.shape
    pins = np.ascontiguousarray(pins, dtype=np.uint32)
    arr = (NegoNet * max(len(nets), 1))()
    for i, (off, cnt, win) in enumerate(nets):
        wx0, wy0, ww, wh = win if win is not None else (0, 0, 0, 0)
        arr[i] = NegoNet(off, cnt, wx0, wy0, ww, wh)
    pol = NegoPolicy(maximum_passes, batch_nets, initial_present_cost, present_cost_increment,
                     historical_cost_increment, maximum_congestion_cost_value, maximum_cell_cost)
    n = cost.size
    cap = None if capacity is None else np.ascontiguousarray(capacity, dtype=np.uint8).reshape(-1)
    init = None if initial_history is None else np.ascontiguousarray(initial_history, dtype=np.uint32).reshape(-1)
    hist = np.zeros(n, dtype=np.uint32)
    occ = np.zeros(n, dtype=np.uint32)
    per_pass = np.zeros(max(maximum_passes, 1), dtype=np.uint64)
    rlen = np.zeros(max(len(nets), 1), dtype=np.uint32)
    rcap = n * 4 + 64
    rcells = np.empty(rcap, dtype=np.uint32)
    res = NegoResult()
    rc = h.orc_negotiate(W, H, L, _u32(cost), via,
                         cap.ctypes.data_as(C.POINTER(C.c_uint8)) if cap is not None else None,
                         arr, len(nets), _u32(pins), C.byref(pol),
                         _u32(init) if init is not None else None, _u32(hist), _u32(occ),
                         per_pass.ctypes.data_as(C.POINTER(C.c_uint64)), _u32(rlen), _u32(rcells),
                         rcap, C.byref(res))
    if rc != 0:
        raise RuntimeError("oracle negotiate failed")
    routes, off = [], 0
    for i in range(len(nets)):
        k = int(rlen[i])
        routes.append(rcells[off:off + k].copy() if k else None)
        off += k
    return {"status": res.status, "termination": res.termination, "passes": res.completed_passes,
            "routed": res.routed_nets, "overuse": int(res.total_overuse_units),
            "overuse_per_pass": [int(v) for v in per_pass[:res.completed_passes]],
            "history": hist.reshape(L, H, W), "occupancy": occ.reshape(L, H, W), "routes": routes}


def replay_exact_pass(cost, via, nets, pins, incumbents, occupancy, history, pass_index, *, first=None,
                      budget_s=None, capacity=1, initial_present_cost=0, present_cost_increment=100,
                      maximum_congestion_cost_value=10_000_000, maximum_cell_cost=(1 << 20) - 1):
    """The first nets of one exact-order pass of the allocator oracle, started
    from a given state: the incumbent trees (None: unrouted), usage and history
    planes at the start of pass `pass_index`.  Per net, in net order
    (src/allocator/cpu_sequential_negotiated_routing.cc:1024-1254, as
    orc_negotiate with batches of one): rip up its incumbent, snapshot the cell
    costs of its window, base + history + pf * max(0, usage + 1 - capacity)
    with pf = initial + pass_index * increment (:306-326, :446-480), route it
    (canonical multi-pin growth from pin 0, vias paying the entered cell), and
    install the new tree, or restore the incumbent when the net is
    disconnected.  Stops after `first` nets, or once `budget_s` seconds have
    passed.  nets: (pin_offset, pin_count, window) per net; capacity: one value
    for every cell or a plane of the lattice's shape.  Returns the trees of
    the nets routed, in order, and the seconds spent."""
    import time

    base = np.ascontiguousarray(cost, dtype=np.uint32)
    L, H, W = base.shape
    occ = np.array(occupancy, dtype=np.int64).reshape(L, H, W)
    hist = np.asarray(history, dtype=np.int64).reshape(L, H, W)
    pf = initial_present_cost + pass_index * present_cost_increment
    cap = np.broadcast_to(np.asarray(capacity, dtype=np.int64).reshape(
        (L, H, W) if np.ndim(capacity) else ()), (L, H, W))
    snap = base.copy()
    trees = []
    t0 = time.perf_counter()
    for i, (off, cnt, win) in enumerate(nets):
        if first is not None and i >= first:
            break
        if budget_s is not None and i > 0 and time.perf_counter() - t0 >= budget_s:
            break
        inc = incumbents[i]
        if inc is not None:
            np.subtract.at(occ.reshape(-1), np.asarray(inc, dtype=np.int64), 1)
        x0, y0, w, h = win if win is not None else (0, 0, W, H)
        sl = (slice(None), slice(y0, y0 + h), slice(x0, x0 + w))
        b = base[sl].astype(np.int64)
        cong = hist[sl] + pf * np.maximum(occ[sl] + 1 - cap[sl], 0)
        if (np.where(b > 0, cong, 0) > maximum_congestion_cost_value).any() or \
                (np.where(b > 0, b + cong, 0) > maximum_cell_cost).any():
            raise ValueError(f"net {i}: congestion cost above the value bound")
        snap[sl] = np.where(b > 0, b + cong, 0).astype(np.uint32)
        rc, conns = Dense(snap, via, window=win, via_pays_cell=True).route_net(pins[off:off + cnt])
        if rc == 0:
            tree = np.concatenate([np.asarray(pins[off:off + 1], dtype=np.uint32)] +
                                  [c[2][:-1] for c in conns]).astype(np.uint32)
        else:
            tree = inc
        if tree is not None:
            np.add.at(occ.reshape(-1), np.asarray(tree, dtype=np.int64), 1)
        trees.append(tree)
    return trees, time.perf_counter() - t0


# ------------------------------------------------------------------ V1 model
class V1Node(C.Structure):
    _fields_ = [("x", C.c_int64), ("y", C.c_int64), ("layer", C.c_uint32), ("mask", C.c_uint8),
                ("pad", C.c_uint8 * 3)]


class V1Profile(C.Structure):
    _fields_ = [("orth", C.c_uint32), ("diag", C.c_uint32), ("bend", C.c_uint32),
                ("heading_mask", C.c_uint32), ("tile_w", C.c_uint32), ("tile_h", C.c_uint32)]


class V1Policy(C.Structure):
    _fields_ = [("orth_s", C.c_uint64), ("diag_s", C.c_uint64), ("bend_s", C.c_uint64),
                ("ban_keys", C.POINTER(C.c_int64)), ("n_bans", C.c_uint32), ("n_pens", C.c_uint32),
                ("pen_keys", C.POINTER(C.c_int64)), ("pen_vals", C.POINTER(C.c_uint64))]


def _v1_setup():
    h = lib()
    if getattr(h, "_v1_ready", False):
        return h
    P = C.POINTER
    h.orc_v1_astar.restype = C.c_int
    h.orc_v1_astar.argtypes = [P(V1Node), C.c_uint64, P(V1Profile), P(V1Policy), C.c_uint32,
                               C.c_int64, C.c_int64, C.c_int64, C.c_int64, P(C.c_uint64),
                               P(C.c_int64), C.c_uint64, P(C.c_uint64), P(C.c_uint64)]
    h.orc_v1_flatten_order.restype = C.c_int
    h.orc_v1_flatten_order.argtypes = [P(V1Node), C.c_uint64, C.c_uint32, C.c_uint32, P(C.c_uint32)]
    h.orc_v1_canonical.restype = C.c_int
    h.orc_v1_canonical.argtypes = [P(V1Node), C.c_uint64, P(V1Profile), P(V1Policy), C.c_uint32,
                                   C.c_int64, C.c_int64, C.c_int64, C.c_int64, P(C.c_uint64),
                                   P(C.c_uint32), C.c_uint64, P(C.c_uint64)]
    h.orc_v1_frontier.restype = C.c_int
    h.orc_v1_frontier.argtypes = [P(V1Node), C.c_uint64, P(V1Profile), P(V1Policy), C.c_uint32,
                                  C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_uint32, P(C.c_uint64),
                                  P(C.c_uint32), P(C.c_uint32), P(C.c_uint64), P(C.c_uint32), C.c_uint64,
                                  P(C.c_uint64)]
    h.orc_v1_frontier_workspace.restype = C.c_int
    h.orc_v1_frontier_workspace.argtypes = [P(V1Node), C.c_uint64, P(V1Profile), P(V1Policy), C.c_uint32,
                                            C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_uint32,
                                            P(C.c_uint32), P(C.c_uint32), P(C.c_uint64), P(C.c_uint32)]
    h.orc_v1_route.restype = C.c_int
    h.orc_v1_route.argtypes = [P(V1Node), C.c_uint64, P(V1Profile), C.c_uint32, C.c_uint32, C.c_int64, C.c_int64,
                               C.c_int64, C.c_int64, C.c_uint32, P(C.c_uint32), P(C.c_uint32), P(C.c_uint64),
                               P(C.c_uint64), P(C.c_uint32), P(C.c_uint64), P(C.c_uint32)]
    h._v1_ready = True
    return h


class V1:
    """Planar heading-state oracle over compiled nodes [(x, y, layer, mask)]."""

    def __init__(self, nodes, costs, heading_mask, tile_w, tile_h):
        self.n = len(nodes)
        self.nodes = (V1Node * max(self.n, 1))()
        for i, (x, y, l, m) in enumerate(nodes):
            self.nodes[i].x, self.nodes[i].y, self.nodes[i].layer, self.nodes[i].mask = x, y, l, m
        self.profile = V1Profile(costs[0], costs[1], costs[2], heading_mask, tile_w, tile_h)

    @staticmethod
    def policy(orth_s=0, diag_s=0, bend_s=0, bans=(), penalties=()):
        bk = (C.c_int64 * max(3 * len(bans), 1))(*[v for b in bans for v in b])
        pk = (C.c_int64 * max(3 * len(penalties), 1))(*[v for (k, _) in penalties for v in k])
        pv = (C.c_uint64 * max(len(penalties), 1))(*[v for (_, v) in penalties])
        pol = V1Policy(orth_s, diag_s, bend_s, bk, len(bans), len(penalties), pk, pv)
        pol._keep = (bk, pk, pv)
        return pol

    def astar(self, start, goal, policy=None, layer=0):
        h = _v1_setup()
        cost, plen, exp = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        path = (C.c_int64 * (2 * (9 * self.n + 2)))()
        rc = h.orc_v1_astar(self.nodes, self.n, C.byref(self.profile),
                            C.byref(policy) if policy is not None else None, layer, start[0],
                            start[1], goal[0], goal[1], C.byref(cost), path, 9 * self.n + 2,
                            C.byref(plen), C.byref(exp))
        if rc < 0:
            raise RuntimeError("oracle v1 astar error")
        pts = [(path[2 * i], path[2 * i + 1]) for i in range(plen.value)][::-1]
        return (None if rc == 1 else cost.value), pts, exp.value

    def flatten_order(self, tile_w, tile_h):
        h = _v1_setup()
        order = (C.c_uint32 * max(self.n, 1))()
        if h.orc_v1_flatten_order(self.nodes, self.n, tile_w, tile_h, order):
            raise RuntimeError("flatten order failed")
        return list(order)[: self.n]

    def canonical(self, start, goal, policy=None, layer=0):
        h = _v1_setup()
        cost, plen = C.c_uint64(0), C.c_uint64(0)
        states = (C.c_uint32 * (9 * self.n + 2))()
        rc = h.orc_v1_canonical(self.nodes, self.n, C.byref(self.profile),
                                C.byref(policy) if policy is not None else None, layer, start[0],
                                start[1], goal[0], goal[1], C.byref(cost), states, 9 * self.n + 2,
                                C.byref(plen))
        if rc < 0:
            raise RuntimeError("oracle v1 canonical error")
        return (None if rc == 1 else cost.value), list(states)[: plen.value]


def v1_frontier(o: "V1", start, goal, policy=None, layer=0, max_rounds=100_000):
    """Batch frontier restatement: dict(status 0 reached / 1 disconnected / 2
    budget, cost, goal_state, rounds, examined_work, states) in flattened
    numbering."""
    h = _v1_setup()
    cost, plen, examined = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
    goal_state, rounds = C.c_uint32(0), C.c_uint32(0)
    states = (C.c_uint32 * (9 * o.n + 2))()
    rc = h.orc_v1_frontier(o.nodes, o.n, C.byref(o.profile), C.byref(policy) if policy is not None else None,
                           layer, start[0], start[1], goal[0], goal[1], max_rounds, C.byref(cost),
                           C.byref(goal_state), C.byref(rounds), C.byref(examined), states, 9 * o.n + 2,
                           C.byref(plen))
    if rc < 0:
        raise RuntimeError("oracle v1 frontier error")
    return {"status": rc, "cost": cost.value if rc == 0 else None, "goal_state": goal_state.value,
            "rounds": rounds.value, "examined_work": examined.value,
            "states": list(states)[: plen.value] if rc == 0 else []}


def v1_frontier_workspace(o: "V1", start, goal, policy=None, layer=0, max_rounds=100_000):
    """The frontier restatement's workspace when it stops: dict(status, goal_state,
    rounds, labels, predecessors) with per-state arrays in flattened numbering
    (predecessor 0xFFFFFFFF = none)."""
    import numpy as np
    h = _v1_setup()
    goal_state, rounds = C.c_uint32(0), C.c_uint32(0)
    labels = np.zeros(9 * o.n, np.uint64)
    preds = np.zeros(9 * o.n, np.uint32)
    rc = h.orc_v1_frontier_workspace(o.nodes, o.n, C.byref(o.profile),
                                     C.byref(policy) if policy is not None else None, layer, start[0], start[1],
                                     goal[0], goal[1], max_rounds, C.byref(goal_state), C.byref(rounds),
                                     labels.ctypes.data_as(C.POINTER(C.c_uint64)),
                                     preds.ctypes.data_as(C.POINTER(C.c_uint32)))
    if rc < 0:
        raise RuntimeError("oracle v1 frontier workspace error")
    return {"status": rc, "goal_state": goal_state.value, "rounds": rounds.value, "labels": labels,
            "predecessors": preds}


def v1_route(o: "V1", start, goal, generator, layer=0, max_rounds=100_000):
    """Single-route generator restatement (generator 0 frontier, 1 sweep):
    dict(completion 0 reached / 1 disconnected / 2 budget, rounds,
    examined_work, turns, goal_state, labels, predecessors), flattened
    numbering."""
    import numpy as np
    h = _v1_setup()
    comp, rounds, gs = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
    examined, turns = C.c_uint64(0), C.c_uint64(0)
    labels = np.zeros(9 * o.n, np.uint64)
    preds = np.zeros(9 * o.n, np.uint32)
    rc = h.orc_v1_route(o.nodes, o.n, C.byref(o.profile), generator, layer, start[0], start[1], goal[0], goal[1],
                        max_rounds, C.byref(comp), C.byref(rounds), C.byref(examined), C.byref(turns), C.byref(gs),
                        labels.ctypes.data_as(C.POINTER(C.c_uint64)), preds.ctypes.data_as(C.POINTER(C.c_uint32)))
    if rc < 0:
        raise RuntimeError("oracle v1 route error")
    return {"completion": comp.value, "rounds": rounds.value, "examined_work": examined.value,
            "turns": turns.value, "goal_state": gs.value, "labels": labels, "predecessors": preds}
