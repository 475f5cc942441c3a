"""Workload definitions shared by tests, the smoke check and the benchmark.

The five lattice workloads of this project (sizes, cost models, seeds):
  single_layer_256   1 x 256 x 256, uint32 cost U[1,16], 10% i.i.d. blocked, 64 pairs, seed 1
  four_layer_1024    4 x 1024 x 1024, uint16 cost U[1,8], 15% rectangular keepouts
                     (sides U[4,64]), vias cost 20, 1024 pairs, seed 2
  eight_layer_2048   8 x 2048 x 2048, same model, 4096 nets of 2-8 pins, windows
                     log-uniform 32..2048, seed 3
  six_layer_1536     allocator boards: 6 x 1536 x 1536, uint16 cost U[1,8], 10%
                     keepouts, five BGA fields of 32 x 32 pads at 4-cell pitch
                     (128 x 128-cell footprints, pads on layer 0), 2000 nets of
                     2-6 pins: pin 0 on a pad, further pins on a pad with
                     probability 0.3, else scattered within 384 cells of the
                     net's field; windows = pin bounding box + 16; pin clearance 1
                     (no pin walled in by keepouts or another pin); capacity 1
                     per cell
  six_layer_512      the same generator on 6 x 512 x 512 with one field and 200
                     nets (learned-prior study boards; seeds 0-199 training,
                     200-249 held out)
  six_layer_512_dense the same with 300 nets: about 570 used pads in the field,
                     near six_layer_1536's 760, so runs need tens of passes
"""
from __future__ import annotations

SINGLE_LAYER_256 = dict(width=256, height=256, layers=1, cost_min=1, cost_max=16,
                        blocked_ppm=100_000, seed=1, via=0, queries=64, cost_bytes=4)
FOUR_LAYER_1024 = dict(width=1024, height=1024, layers=4, cost_min=1, cost_max=8,
                       keepout_ppm=150_000, keepout_min_side=4, keepout_max_side=64, seed=2,
                       via=20, queries=1024, cost_bytes=2)
EIGHT_LAYER_2048 = dict(width=2048, height=2048, layers=8, cost_min=1, cost_max=8,
                        keepout_ppm=150_000, keepout_min_side=4, keepout_max_side=64, seed=3,
                        via=20, nets=4096, min_pins=2, max_pins=8, min_side=32, max_side=2048,
                        cost_bytes=2)


SIX_LAYER_1536 = dict(width=1536, height=1536, layers=6, cost_min=1, cost_max=8,
                      keepout_ppm=100_000, keepout_min_side=4, keepout_max_side=64, nets=2000,
                      min_pins=2, max_pins=6, bga_fields=5, bga_pads=32, bga_pitch=4,
                      pad_pin_ppm=300_000, scatter_radius=384, window_margin=16, pin_clearance=1,
                      via=20)

# One field of the six_layer_1536 family on a 512 x 512 board at about the same
# net density (200 nets per 512^2 against 2000 per 1536^2): the scale at which
# the learned prior's 200 training and 50 held-out allocator runs fit in minutes.
SIX_LAYER_512 = dict(SIX_LAYER_1536, width=512, height=512, nets=200, bga_fields=1, scatter_radius=192)

# The same field under more nets: pad use per field closer to six_layer_1536's,
# where the plain allocator needs tens of passes rather than three or four.
SIX_LAYER_512_DENSE = dict(SIX_LAYER_512, nets=300)


def make_board(apgx, spec, seed, **overrides):
    """Allocator board of the six_layer_1536 family (or a scaled variant)."""
    p = {k: v for k, v in spec.items() if k != "via"}
    p.update(overrides)
    return apgx.generate_board(seed=seed, **p)


def make_lattice(apgx, spec):
    return apgx.generate_lattice(spec["width"], spec["height"], spec["layers"], spec["cost_min"],
                                 spec["cost_max"], spec.get("blocked_ppm", 0),
                                 spec.get("keepout_ppm", 0), spec.get("keepout_min_side", 0),
                                 spec.get("keepout_max_side", 0), spec["seed"])
