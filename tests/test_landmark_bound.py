"""The ALT goal bound's mathematics, checked by brute force on the CPU.

GoalBound (apgar_amd/csrc/lattice.hip) keys a neighbour push with the bound
over the target tile's entry edge line: per landmark l, d(l, p) minus the
largest finite d(l, .) over that edge line of the 64 x 64 column, and the
Manhattan distance from the line times the smallest cell cost.  The settle
rule needs that to be at most the true distance from every entry cell to the
pin, and the per-cell bound to be consistent.  Here the landmark fields come
from the oracle's Dijkstra, the true distances to the pin from a reverse
Dijkstra over the same lattice (scipy), on a lattice with disconnected
regions, so that leaving unreachable cells out of the maxima is exercised."""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import dijkstra

from tests.test_oracle_dense import random_lattice

UNR = 0xFFFFFFFF
COL = 64


def _graph(cost, via):
    """Directed edge u -> v weighs cost[v] in a layer, via between layers."""
    L, H, W = cost.shape
    flat = cost.reshape(-1).astype(np.int64)
    idx = np.arange(flat.size).reshape(L, H, W)
    rows, cols, wts = [], [], []
    pairs = [(idx[:, :, :-1], idx[:, :, 1:]), (idx[:, :-1, :], idx[:, 1:, :])]
    for a, b in pairs:
        for u, v in ((a.ravel(), b.ravel()), (b.ravel(), a.ravel())):
            m = (flat[u] > 0) & (flat[v] > 0)
            rows.append(u[m]); cols.append(v[m]); wts.append(flat[v[m]])
    if L > 1 and via:
        for u, v in ((idx[:-1].ravel(), idx[1:].ravel()), (idx[1:].ravel(), idx[:-1].ravel())):
            m = (flat[u] > 0) & (flat[v] > 0)
            rows.append(u[m]); cols.append(v[m]); wts.append(np.full(m.sum(), via))
    return csr_matrix((np.concatenate(wts).astype(float), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(flat.size, flat.size))


@pytest.mark.parametrize("seed", range(3))
def test_entry_edge_bound_is_admissible_and_consistent(oracle, seed):
    rng = np.random.default_rng(700 + seed)
    L, H, W, via = 2, 150, 140, 5
    cost = random_lattice(rng, L, H, W, blocked=0.3, cmax=8)
    # Two single free cells walled in on every layer: a landmark that reaches
    # nothing else and a pin that no landmark reaches.
    for cx_, cy_ in ((10, 12), (100, 120)):
        cost[:, cy_ - 1:cy_ + 2, cx_ - 1:cx_ + 2] = 0
        cost[0, cy_, cx_] = 3
    flat = cost.reshape(-1)
    free = np.flatnonzero(flat)
    o = oracle.Dense(cost, via)
    marks = np.concatenate([[12 * W + 10], rng.choice(free, 5, replace=False)])
    fields = np.stack([o.dijkstra([int(m)]).astype(np.float64) for m in marks])
    fields[fields == UNR] = np.inf
    G = _graph(cost, via)
    min_cost = int(flat[free].min())
    x = np.arange(flat.size) % W
    y = (np.arange(flat.size) // W) % H
    for p in np.concatenate([[120 * W + 100], rng.choice(free, 4, replace=False)]):
        to_p = dijkstra(G.T.tocsr(), indices=[int(p)])[0]  # d(c, p) for every c
        dlp = fields[:, p]
        # Per-cell bound: consistent along every edge.
        h = np.zeros(flat.size)
        for k in range(len(marks)):
            if np.isfinite(dlp[k]):
                with np.errstate(invalid="ignore"):
                    h = np.maximum(h, np.where(np.isfinite(fields[k]), dlp[k] - fields[k], 0.0))
        h = np.maximum(h, (np.abs(x - p % W) + np.abs(y - (p // W) % H)) * min_cost)
        coo = G.tocoo()
        reach = np.isfinite(to_p)
        e = reach[coo.row] & reach[coo.col]
        assert np.all(h[coo.row[e]] <= coo.data[e] + h[coo.col[e]] + 1e-9)
        assert np.all(h[reach] <= to_p[reach] + 1e-9)
        # Tile bound over each entry edge line of each 64 x 64 column.
        for cy in range(0, H, COL):
            for cx in range(0, W, COL):
                x1, y1 = min(cx + COL, W) - 1, min(cy + COL, H) - 1
                for side in range(4):
                    ex0, ex1 = (x1, x1) if side == 1 else ((cx, cx) if side == 0 else (cx, x1))
                    ey0, ey1 = (y1, y1) if side == 3 else ((cy, cy) if side == 2 else (cy, y1))
                    line = (x >= ex0) & (x <= ex1) & (y >= ey0) & (y <= ey1) & (flat > 0)
                    if not line.any():
                        continue
                    px, py = p % W, (p // W) % H
                    dx = max(ex0 - px, 0, px - ex1)
                    dy = max(ey0 - py, 0, py - ey1)
                    b = (dx + dy) * min_cost
                    for k in range(len(marks)):
                        if not np.isfinite(dlp[k]):
                            continue
                        vals = fields[k][line]
                        far = vals[np.isfinite(vals)].max(initial=0.0)
                        if dlp[k] > far:
                            b = max(b, dlp[k] - far)
                    assert b <= to_p[line].min() + 1e-9, (seed, p, cx, cy, side)
