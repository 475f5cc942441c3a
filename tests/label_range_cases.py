"""Lattices that put shortest-path distances at and beyond the 29-bit label
range of the dense search (labels hold distances below 2^29 - 1).

Shared by the CPU test, which pins every premise with the oracle's 64-bit
Dijkstra (which distances lie above or below the range), and the GPU test,
which checks that the search returns the exact result below the range and
fails with APGX_E_UNSUPPORTED, never a false "disconnected", above it.
"""
from __future__ import annotations

import numpy as np

LABEL_LIMIT = (1 << 29) - 1   # distances of this size or more do not fit a label
MAX_COST = (1 << 20) - 1      # APGX_MAX_CELL_COST
REG_MAX_COST = 4095           # largest cost the register-resident tiles take


def corridor(length: int, cost: int = MAX_COST, wall_at: int | None = None) -> np.ndarray:
    """1 x 1 x length corridor of uniform cost; an optional blocked cell."""
    c = np.full((1, 1, length), cost, dtype=np.uint32)
    if wall_at is not None:
        c[0, 0, wall_at] = 0
    return c


def serpentine(side: int = 640, cost: int = REG_MAX_COST):
    """1 x side x side maze of uniform cost whose only path snakes row by row:
    odd rows are walls with one gap, alternately at the east and west end.
    Returns (cost plane, source cell, cell -> cell index helper)."""
    c = np.full((1, side, side), cost, dtype=np.uint32)
    for y in range(1, side, 2):
        c[0, y, :] = 0
        gap = side - 1 if (y // 2) % 2 == 0 else 0
        c[0, y, gap] = cost
    return c


def comb(teeth: int, length: int = 500, tooth_cost: int = MAX_COST):
    """1 x (length+1) x (2*teeth - 1) comb: a cost-1 spine along row 0 and
    `teeth` vertical teeth of tooth_cost at even columns (odd columns blocked
    below the spine).  Pins: pin 0 at the spine's first cell, then each
    tooth's far end.  Each connection costs length * tooth_cost plus spine steps."""
    W, H = 2 * teeth - 1, length + 1
    c = np.zeros((1, H, W), dtype=np.uint32)
    c[0, 0, :] = 1
    c[0, 1:, 0::2] = tooth_cost
    pins = [0] + [length * W + 2 * k for k in range(teeth)]
    return c, np.asarray(pins, dtype=np.uint32)
