#!/usr/bin/env python3
"""Benchmark of the batched lattice search on one MI355X.

One step routes the four_layer_1024 workload (4 x 1024 x 1024 lattice, uint16
cell cost U[1,8], 15% rectangular keepouts, vias cost 20, 1024 single-pair
queries) through the C-ABI.  The cost plane is resident in HBM before timing;
`value` is queries per second over the device envelope of each step (hipEvent
from the seed kernel to the last phase kernel).  The host-inclusive rate
(query upload, result readback) is reported beside it as wall_queries_per_s.

Prints one JSON line.  --gpus above 1 is rejected: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import apgar_amd as apgx  # noqa: E402
from tests.lattice_cases import FOUR_LAYER_1024, make_lattice  # noqa: E402

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = ROOT / "profiles" / "round_kernel_four_layer_1024_traffic.txt"


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--queries", type=int, default=FOUR_LAYER_1024["queries"])
    p.add_argument("--cpu-sample", type=int, default=32, help="queries timed on the CPU oracle")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--tile", type=str, default="0x0")
    p.add_argument("--delta", type=int, default=0, help="tile scheduling bucket width (0: default)")
    p.add_argument("--max-sweeps", type=int, default=0)
    p.add_argument("--no-goal-bound", action="store_true", help="order tiles by distance only")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only run a device-to-device copy of MIB MiB (HBM counter scale reference)")
    return p.parse_args()


def read_traffic():
    """HBM bytes per round-kernel launch from the committed rocprofv3 PMC summary."""
    if not TRAFFIC_FILE.exists():
        return None
    for line in TRAFFIC_FILE.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(cost, pairs, sample, threads):
    """CPU A* oracle (restated reference algorithm) on a bounded sample."""
    from oracle import py as oracle_py

    o = oracle_py.Dense(cost, FOUR_LAYER_1024["via"])
    idx = list(range(min(sample, len(pairs))))

    def one(i):
        s, t = (int(v) for v in pairs[i])
        return o.astar(s, t)[0]

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as pool:
        list(pool.map(one, idx))
    dt = time.perf_counter() - t0
    return {"value": len(idx) / dt, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"first {len(idx)} of the {len(pairs)} four_layer_1024 queries, "
                      f"one A* per thread, {threads} threads, {dt:.1f} s"}


def copy_calibration(mib):
    """A plain copy kernel of known size: reads and writes `mib` MiB once, so
    rocprofv3's FETCH_SIZE / WRITE_SIZE for it give the counters' scale."""
    import torch

    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n, "launches": 3}))


def main():
    args = parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); queries are "
              "independent and would shard by query list, which this build does not launch.",
              file=sys.stderr)
        sys.exit(2)
    spec = FOUR_LAYER_1024
    tile = tuple(int(v) for v in args.tile.split("x"))
    cost = make_lattice(apgx, spec)
    pairs = apgx.generate_pairs(cost, args.queries, 202)
    lat = apgx.Lattice(cost, via_cost=spec["via"], cost_bytes=2)
    print(f"[bench] lattice {cost.shape} uploaded; {args.queries} queries", file=sys.stderr, flush=True)

    def step(profile=False):
        routes, stats = lat.route_pairs(pairs, tile=tile, profile_rounds=1 if profile else 0,
                                        frontier_delta=args.delta, max_sweeps=args.max_sweeps,
                                        goal_bound=not args.no_goal_bound)
        return routes, stats

    for i in range(args.warmup):
        _, st = step()
        print(f"[bench] warmup {i}: device {st.device_ms:.2f} ms rounds {st.rounds}",
              file=sys.stderr, flush=True)
    apgx.synchronize()
    t0 = time.perf_counter()
    dev_ms, all_stats = [], []
    reached = 0
    for i in range(args.steps):
        routes, st = step()
        dev_ms.append(st.device_ms)
        all_stats.append(st)
        reached = sum(r.status == apgx.NET_REACHED for r in routes)
        print(f"[bench] step {i}: device {st.device_ms:.2f} ms wall {st.wall_ms:.2f} ms "
              f"rounds {st.rounds} tiles {st.tile_visits}", file=sys.stderr, flush=True)
    apgx.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    ms = float(np.mean(dev_ms))

    # Roofline of the dominant kernel (the round kernel), timed per launch.
    _, prof = step(profile=True)
    launches = max(prof.rounds, 1)
    bytes_per_launch = prof.algorithmic_bytes / launches
    avg_launch_s = prof.round_kernel_ms / 1e3 / launches
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    traffic = read_traffic()
    st = all_stats[-1]
    line = {
        "metric": "queries/s & % HBM roofline",
        "value": args.queries / (ms / 1e3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "four_layer_1024",
            "lattice": "4x1024x1024, uint16 cost U[1,8], 15% keepouts (sides 4-64), vias 20",
            "queries": args.queries,
            "tile": f"{st.tile_w}x{st.tile_h}",
            "parallelism": "single GPU",
        },
        "wall_ms_per_step": wall * 1e3,
        "wall_queries_per_s": args.queries / wall,
        "reached": reached,
        "rounds": st.rounds,
        "tile_visits": st.tile_visits,
        "inner_sweeps_per_visit": st.inner_sweeps / max(st.tile_visits, 1),
        "relaxations_per_s": st.relaxations / (ms / 1e3),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": (f"RegRoundKernel<{cost.shape[0]}>" if (st.tile_w, st.tile_h) == (64, 64)
                       else f"RoundKernel<{st.tile_w},{st.tile_h}>"),
            "launches": launches,
            "avg_launch_us": avg_launch_s * 1e6,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(cost, pairs, args.cpu_sample, threads)
    print(json.dumps(line))
    lat.close()


if __name__ == "__main__":
    main()
