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
                ("wy0", C.c_uint32), ("ww", C.c_uint32), ("wh", C.c_uint32)]


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

    def __init__(self, cost: np.ndarray, via: int, window=None):
        self.cost = np.ascontiguousarray(cost, dtype=np.uint32)
        L, H, W = self.cost.shape
        self.shape = (L, H, W)
        wx0, wy0, ww, wh = window if window is not None else (0, 0, 0, 0)
        self.g = OrcLattice(W, H, L, _u32(self.cost), via, wx0, wy0, ww, wh)

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
