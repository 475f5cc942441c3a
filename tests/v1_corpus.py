"""Reference planar corpus and known-answer boards, restated as test fixtures.

Restates the case descriptions of src/benchmark/planar_corpus.cc:196-377 and the
obstacle-free boards of tests/routing/cpu_astar_test.cc, plus the two compiler
steps those obstacle-free boards need:
  * represented nodes = union of active-region lattice points
    (src/geometry_compiler/compiled_board.cc:395-435);
  * a forward edge (E, NE, N, NW) allowed by the heading mask between two
    represented nodes is legal in both directions (:442-490; with no obstacles
    the exact movement oracle accepts every such segment).
The compiler-profile fingerprint (FingerprintCompilerProfile, :268-295) is the
byte-stable FNV-1a of include/apgar/board_ir/stable_hash.h:12-53.  Everything
here is pinned by tests/golden/corpus_golden.json (the reference's published
node, edge, byte and fingerprint values).
"""
from __future__ import annotations

from dataclasses import dataclass, field

H, V, D = 1, 2, 4
M1 = H | V | D
DX = [1, 1, 0, -1, -1, -1, 0, 1]
DY = [0, 1, 1, 1, 0, -1, -1, -1]


class Fnv1a:
    OFFSET = 14695981039346656037
    PRIME = 1099511628211
    MASK = (1 << 64) - 1

    def __init__(self):
        self.v = self.OFFSET

    def byte(self, b):
        self.v = ((self.v ^ (b & 0xFF)) * self.PRIME) & self.MASK

    def u32(self, x):
        for i in range(4):
            self.byte((x >> (8 * i)) & 0xFF)

    def u64(self, x):
        x &= self.MASK
        for i in range(8):
            self.byte((x >> (8 * i)) & 0xFF)

    def i64(self, x):
        self.u64(x & self.MASK)

    def string(self, s: str):
        self.u64(len(s.encode()))
        for c in s.encode():
            self.byte(c)


def stable_hash_string(data: bytes) -> int:
    h = Fnv1a()
    for c in data:
        h.byte(c)
    return h.v


@dataclass
class Profile:
    roi: tuple          # (minx, miny, maxx, maxy) exact coordinates
    step: int
    tile_w: int
    tile_h: int
    costs: tuple        # (orthogonal, diagonal, bend)
    heading: int = M1
    regions: list = field(default_factory=list)   # [(layer, (minx, miny, maxx, maxy))]
    origin: tuple = (0, 0)

    def fingerprint(self) -> int:
        regs = sorted(set((l, b[0], b[1], b[2], b[3]) for l, b in self.regions))
        h = Fnv1a()
        h.string("APGAR-COMPILER-PROFILE-V1")
        h.u32(1)
        h.i64(self.origin[0]); h.i64(self.origin[1]); h.i64(self.step)
        h.u32(self.tile_w); h.u32(self.tile_h)
        for v in self.roi:
            h.i64(v)
        h.u64(len(regs))
        for l, a, b, c, d in regs:
            h.u32(l); h.i64(a); h.i64(b); h.i64(c); h.i64(d)
        h.byte(self.heading)
        for c in self.costs:
            h.u32(c)
        return h.v

    def add_point(self, x, y):
        e = (self.origin[0] + x * self.step, self.origin[1] + y * self.step)
        self.regions.append((0, (e[0], e[1], e[0], e[1])))

    def add_line(self, a, b):
        dx = (b[0] > a[0]) - (b[0] < a[0])
        dy = (b[1] > a[1]) - (b[1] < a[1])
        n = max(abs(b[0] - a[0]), abs(b[1] - a[1]))
        assert abs(b[0] - a[0]) in (0, n) and abs(b[1] - a[1]) in (0, n)
        for i in range(n + 1):
            self.add_point(a[0] + dx * i, a[1] + dy * i)

    def add_polyline(self, pts):
        for a, b in zip(pts[:-1], pts[1:]):
            self.add_line(a, b)


def _ceil_div(a, b):
    return -((-a) // b)


def compile_obstacle_free(p: Profile):
    """Represented nodes with legal masks: list of (x, y, layer, mask)."""
    nodes = {}
    for layer, (x0, y0, x1, y1) in p.regions:
        lx0 = _ceil_div(x0 - p.origin[0], p.step)
        lx1 = (x1 - p.origin[0]) // p.step
        ly0 = _ceil_div(y0 - p.origin[1], p.step)
        ly1 = (y1 - p.origin[1]) // p.step
        for y in range(ly0, ly1 + 1):
            for x in range(lx0, lx1 + 1):
                nodes.setdefault((layer, x, y), 0)
    for (layer, x, y) in list(nodes):
        for d in range(4):
            hm = H if DY[d] == 0 else (V if DX[d] == 0 else D)
            if not (p.heading & hm):
                continue
            nb = (layer, x + DX[d], y + DY[d])
            if nb in nodes:
                nodes[(layer, x, y)] |= 1 << d
                nodes[nb] |= 1 << ((d + 4) % 8)
    return [(x, y, l, m) for (l, x, y), m in nodes.items()]


def base(roi, step, tw, th, costs, heading=M1):
    return Profile(roi=roi, step=step, tile_w=tw, tile_h=th, costs=costs, heading=heading)


def corpus():
    """name -> (profile, start exact point, goal exact point)."""
    out = {}
    p = base((0, -25, 100, 25), 5, 8, 8, (7, 11, 13))
    p.regions = [(0, p.roi)]
    out["dense_corridors"] = (p, (0, 0), (100, 0))
    p = base((0, -40, 100, 40), 10, 4, 3, (19, 3, 29))
    p.add_polyline([(0, 0), (2, 2), (4, 2), (6, 0), (8, -2), (10, 0)])
    out["sparse_regions"] = (p, (0, 0), (100, 0))
    p = base((0, -20, 100, 30), 10, 5, 2, (17, 23, 31))
    for pt in [(0, 0), (1, 1), (2, 1), (3, 2), (4, 2), (5, 1), (6, 1), (7, 0), (8, -1), (9, -1), (10, 0)]:
        p.add_point(*pt)
    out["fragmented_runs"] = (p, (0, 0), (100, 0))
    p = base((0, -40, 100, 40), 10, 3, 3, (10, 14, 101), H | V)
    p.add_polyline([(0, 0), (2, 0), (2, 4), (4, 4), (4, -4), (6, -4), (6, 4), (8, 4), (8, 0), (10, 0)])
    out["high_turn_maze"] = (p, (0, 0), (100, 0))
    p = base((0, -10, 100, 10), 10, 3, 2, (5, 8, 2))
    p.regions = [(0, p.roi)]
    out["cross_tile_edges"] = (p, (0, 0), (100, 0))
    p = base((-120, -70, 20, -30), 10, 4, 3, (37, 41, 43))
    p.regions = [(0, p.roi)]
    out["negative_coordinates"] = (p, (-100, -50), (0, -50))
    p = base((0, 0, 100, 0), 10, 4, 2, (10, 14, 3), H | V)
    p.add_line((0, 0), (4, 0))
    p.add_line((6, 0), (10, 0))
    out["disconnected_fields"] = (p, (0, 0), (100, 0))
    p = base((0, -20, 100, 20), 10, 4, 3, (10, 14, 7))
    p.add_polyline([(0, 0), (2, 2), (8, 2), (10, 0)])
    p.add_polyline([(0, 0), (2, -2), (8, -2), (10, 0)])
    out["symmetric_dual_corridor"] = (p, (0, 0), (100, 0))
    p = base((0, -20, 120, 20), 10, 5, 3, (11, 16, 9))
    for cy in (-2, 0, 2):
        p.add_polyline([(0, 0), (2, 0), (4, cy), (8, cy), (10, 0), (12, 0)])
    out["multi_channel_bottleneck"] = (p, (0, 0), (120, 0))
    p = base((0, -40, 120, 40), 10, 4, 4, (13, 18, 5))
    p.add_line((0, 0), (2, 0))
    p.add_line((10, 0), (12, 0))
    p.add_line((2, -4), (2, 4))
    p.add_line((10, -4), (10, 4))
    for cy in (-4, -2, 0, 2, 4):
        p.add_line((2, cy), (10, cy))
    out["policy_alternatives"] = (p, (0, 0), (120, 0))
    return out


def default_test_profile(heading=M1):
    """tests/support/compiler_builder.h:11-42 restricted to layer 0."""
    p = base((-20, -30, 120, 30), 10, 4, 4, (10, 14, 3), heading)
    p.regions = [(0, (-20, -30, 120, 30))]
    return p


def known_answer_boards():
    """tests/routing/cpu_astar_test.cc obstacle-free known answers."""
    out = {}
    p = default_test_profile()
    p.roi = (0, -30, 40, 20)
    p.regions = []
    for pt in [(0, 0), (10, 0), (10, 10), (20, -20), (20, -10), (20, 20), (30, -30), (30, -10),
               (30, 10), (40, 0)]:
        p.regions.append((0, (pt[0], pt[1], pt[0], pt[1])))
    p.costs = (10, 1, 10)
    out["cheap_diagonal"] = (p, (0, 0), (30, -30))
    p = default_test_profile(H | V)
    out["horizontal_detour"] = (p, (0, 0), (100, 0))
    p = default_test_profile(H)
    p.roi = (0, 0, 100, 0)
    p.regions = [(0, p.roi)]
    out["horizontal_corridor"] = (p, (0, 0), (100, 0))
    return out


def lattice(profile: Profile, point):
    return ((point[0] - profile.origin[0]) // profile.step, (point[1] - profile.origin[1]) // profile.step)


def persistent_bytes(nodes, runs, run_nodes, layers=1):
    """DeviceCompiledBoardV1 persistent bytes (src/gpu/device_board.cc:67-79)."""
    return 112 + 48 * layers + 56 * nodes + 12 * runs + 4 * run_nodes
