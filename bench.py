#!/usr/bin/env python3
"""Benchmark of the batched lattice search on one MI355X.

Default workload, four_layer_1024: a 4 x 1024 x 1024 lattice with uint16 cell
costs U[1,8], 15% rectangular keepouts and vias costing 20, searched by 1024
single-pair queries through the C-ABI.  The cost plane is resident in HBM
before timing.  `value` is queries per second over the device envelope of each
step (hipEvent from the seed kernel to the final path walk).  The rate that
includes the host side (query upload, result readback) is reported beside it
as wall_queries_per_s.

By default the searches add the ALT bound of 16 landmark distance fields
(--landmarks 0 turns it off).  The fields are built once per lattice, like the
upload, before the warmup; their build time and the rate with that build
charged to a single step are reported in "landmarks".

eight_layer_2048: an 8 x 2048 x 2048 lattice with the same model, 4096 nets
of 2-8 pins and log-uniform windows of side 32-2048.  Tile shape and schedule
come from the autotuner on a 256-net sample (reported, not timed), and round
chunks replay as captured hipGraphs.  Its rate is nets/s.  The default run
times it after four_layer_1024 and reports it as the line's "eight_layer_2048"
object (its own roofline and CPU baseline; --no-sub skips it);
--workload eight_layer_2048 makes it the line's main value.

The CPU baseline is the oracle's restatement of the reference CPU A* on the
host's cores (at most 16 threads, the CPU share of one GPU) and on one thread,
each on a bounded sample; the CPU model and core counts are reported with it.

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
from tests.lattice_cases import EIGHT_LAYER_2048, FOUR_LAYER_1024, make_lattice  # noqa: E402

HBM_PEAK_GBS = 8000.0


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", choices=["four_layer_1024", "eight_layer_2048", "learned_prior", "allocator"],
                   default="four_layer_1024")
    p.add_argument("--queries", type=int, default=0, help="queries / nets (0: the workload's count)")
    p.add_argument("--cpu-sample", type=int, default=0, help="queries timed on the CPU oracle (0: auto)")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--tile", type=str, default="0x0")
    p.add_argument("--delta", type=int, default=0, help="tile scheduling bucket width (0: default)")
    p.add_argument("--max-sweeps", type=int, default=0)
    p.add_argument("--no-goal-bound", action="store_true", help="order tiles by distance only")
    p.add_argument("--landmarks", type=int, default=16,
                   help="ALT landmark fields added to the goal bound (0: off); built once per lattice "
                        "before the warmup, their build time reported beside the rate")
    p.add_argument("--landmark-weight", type=int, default=0, help="ordering weight of the bound in 1/16 (0: default)")
    p.add_argument("--sweep-schedule", type=int, default=0,
                   help="register tiles: 1 full sweeps, 2 alternate in-layer-only and full (0: default)")
    p.add_argument("--graph", type=int, default=-1, help="1/0: replay round chunks as hipGraphs "
                   "(default: on for eight_layer_2048)")
    p.add_argument("--no-tune", action="store_true", help="eight_layer_2048: skip the autotuner")
    p.add_argument("--no-sub", action="store_true",
                   help="four_layer_1024: skip the eight_layer_2048 object timed in the same run")
    p.add_argument("--sub-steps", type=int, default=3, help="timed steps of the eight_layer_2048 object")
    p.add_argument("--sub-warmup", type=int, default=1, help="warmup steps of the eight_layer_2048 object")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only run a device-to-device copy of MIB MiB (HBM counter scale reference)")
    p.add_argument("--train-seeds", type=int, default=200, help="learned_prior: training boards")
    p.add_argument("--train-seed-start", type=int, default=0,
                   help="learned_prior --collect-only: first training seed collected (to collect in parts)")
    p.add_argument("--heldout-seeds", type=int, default=50, help="learned_prior: held-out boards")
    p.add_argument("--dataset", type=str, default="", help="learned_prior: .npz of collected training "
                   "data (loaded if present, else collected and written)")
    p.add_argument("--collect-only", action="store_true", help="learned_prior: stop after collection")
    p.add_argument("--epochs", type=int, default=300, help="learned_prior: training epochs")
    p.add_argument("--gains", type=str, default="1,4,16", help="learned_prior: decode gains to validate")
    p.add_argument("--gain", type=float, default=0.0, help="learned_prior: use this decode gain (skips validation)")
    p.add_argument("--upper-bound", type=int, default=0, metavar="N",
                   help="learned_prior: only seed N held-out boards with their own plain-run history "
                        "(per cell and per tile), the best any history prior could do")
    p.add_argument("--full-size-boards", type=int, default=0,
                   help="learned_prior: also compare on this many six_layer_1536 boards")
    p.add_argument("--prior-family", choices=["six_layer_512", "six_layer_512_dense"], default="six_layer_512",
                   help="learned_prior: board family of the study (training, validation, held-out)")
    p.add_argument("--board-seed", type=int, default=200, help="allocator: six_layer_1536 board seed")
    p.add_argument("--passes", type=int, default=40, help="allocator: pass limit")
    p.add_argument("--group-nets", type=int, default=0,
                   help="allocator: speculative group size of the exact order (0: window-disjoint levels)")
    p.add_argument("--prior-gain", type=float, default=0.0,
                   help="allocator: also run the board seeded by the learned prior, trained on --dataset "
                        "(six_layer_512 plain runs), at this decode gain")
    return p.parse_args()


def run_config(workload, opts, landmarks):
    """The search configuration a profile summary must match to be quoted."""
    return {"workload": workload, "landmarks": int(landmarks),
            "landmark_weight": int(opts.get("landmark_weight", 0)),
            "tile": "%dx%d" % tuple(opts.get("tile", (0, 0))),
            "frontier_delta": int(opts.get("frontier_delta", 0)),
            "max_sweeps": int(opts.get("max_sweeps", 0)),
            "sweep_schedule": int(opts.get("sweep_schedule", 0)),
            "goal_bound": bool(opts.get("goal_bound", True))}


def read_profile(kind, config):
    """Figures of the committed rocprofv3 summary of `kind` ("traffic" or
    "sq_counters") for the round kernel under exactly this configuration: the
    summary's `config:` line must equal it.  Returns (fields, source) or
    (None, reason)."""
    f = ROOT / "profiles" / f"round_kernel_{config['workload']}_{kind}.txt"
    if not f.exists():
        return None, f"no profiles/{f.name}"
    fields, cfg = {}, None
    for line in f.read_text().splitlines():
        key, sep, val = line.partition(":")
        if not sep:
            continue
        if key == "config":
            cfg = json.loads(val)
        elif key in ("traffic_bytes_per_launch", "traffic_bytes_per_search", "valu_busy", "lds_busy"):
            fields[key] = float(val)
    if cfg != config:
        return None, f"profiles/{f.name} was measured under another configuration ({cfg})"
    return fields, f"profiles/{f.name}"


def host_cpu():
    """CPU model and the cores this process may use (its affinity set)."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"model": model, "visible_cpus": os.cpu_count(), "usable_cpus": usable}


HEURISTIC = ("oracle A*: key (f, h, g, y, x, layer, seq), heuristic (|dx|+|dy|) x smallest free cell cost "
             "+ |dl| x via (the reference's geometric heuristic restated); the GPU search also uses "
             "landmark (ALT) bounds, so the two do different amounts of work")


def _timed_sample(fn, items, threads, budget_s):
    """Runs fn over items on `threads` threads; with budget_s, stops starting
    new items once that much time has passed (a bounded sample)."""
    t0 = time.perf_counter()
    done = 0
    if threads == 1:
        for it in items:
            fn(it)
            done += 1
            if budget_s and time.perf_counter() - t0 >= budget_s:
                break
    else:
        with ThreadPoolExecutor(max_workers=threads) as pool:
            list(pool.map(fn, items))
        done = len(items)
    return done, time.perf_counter() - t0


def cpu_pairs_baseline(cost, via, pairs, sample, threads):
    """CPU A* oracle (restated reference algorithm) on bounded samples of the
    pairs: one A* per thread on `threads` threads, and one thread alone."""
    from oracle import py as oracle_py

    o = oracle_py.Dense(cost, via)

    def one(i):
        s, t = (int(v) for v in pairs[i])
        return o.astar(s, t)[0]

    n, dt = _timed_sample(one, list(range(min(sample, len(pairs)))), threads, 0)
    n1, dt1 = _timed_sample(one, list(range(len(pairs))), 1, 15.0)
    return {"value": n / dt, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"first {n} of the {len(pairs)} four_layer_1024 queries, one A* per thread, "
                      f"{threads} threads, {dt:.1f} s",
            "single_thread": {"value": n1 / dt1, "cores": 1,
                              "sample": f"first {n1} queries on one thread, {dt1:.1f} s"},
            "heuristic": HEURISTIC, "host": host_cpu()}


def cpu_exact_pass_baseline(lat, cost, via, nets, pins, budget_s):
    """The allocator oracle's exact-order pass on one thread (the order is
    sequential, so more threads do not help it): pass 1 replayed from the
    GPU's state after pass 0 for as many nets as fit in budget_s seconds, the
    trees compared with the GPU's pass 1, the time extrapolated to the pass."""
    from oracle import py as oracle_py

    a = lat.negotiate(nets, pins, maximum_passes=1, batch_nets=0, group_nets=0)
    b = lat.negotiate(nets, pins, maximum_passes=2, batch_nets=0, group_nets=0)
    hist = 10 * np.maximum(a["occupancy"].astype(np.int64) - 1, 0)
    onets = [(n.pin_offset, n.pin_count, (n.window.x0, n.window.y0, n.window.width, n.window.height)) for n in nets]
    trees, secs = oracle_py.replay_exact_pass(cost, via, onets, pins, a["routes"], a["occupancy"], hist, 1,
                                              budget_s=budget_s)
    k = len(trees)
    same = all((t is None) == (g is None) and (t is None or np.array_equal(t, g)) for t, g in zip(trees, b["routes"]))
    return {"value": len(nets) * secs / k * 1e3, "unit": "ms per exact-order pass (extrapolated)", "cores": 1,
            "kind": "port",
            "sample": f"the first {k} of the {len(nets)} nets of pass 1, replayed on one thread from the GPU's "
                      f"state after pass 0, {secs:.1f} s ({k / secs:.2f} nets/s)",
            "sample_trees_equal_gpu": bool(same), "host": host_cpu()}


def cpu_nets_baseline(cost, via, nets, pins, sample, threads):
    """CPU oracle multi-pin routing (repeated Dijkstra from the tree) on an evenly
    strided sample of nets, one net per thread, and on one thread alone."""
    from oracle import py as oracle_py

    idx = [i * len(nets) // sample for i in range(sample)]

    def one(i):
        n = nets[i]
        win = (n.window.x0, n.window.y0, n.window.width, n.window.height)
        return oracle_py.Dense(cost, via, win).route_net(pins[n.pin_offset:n.pin_offset + n.pin_count])

    n, dt = _timed_sample(one, idx, threads, 0)
    n1, dt1 = _timed_sample(one, idx, 1, 15.0)
    return {"value": n / dt, "unit": "nets/s", "cores": threads, "kind": "port",
            "sample": f"{n} of the {len(nets)} eight_layer_2048 nets (every "
                      f"{len(nets) // sample}th), one net per thread, {threads} threads, {dt:.1f} s",
            "single_thread": {"value": n1 / dt1, "cores": 1,
                              "sample": f"the first {n1} of those nets on one thread, {dt1:.1f} s"},
            "heuristic": "oracle multi-source Dijkstra from the growing tree (no goal bound)",
            "host": host_cpu()}


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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def learned_prior(args):
    """Learned history-cost prior on six_layer_512 (or six_layer_512_dense)
    boards: plain allocator runs on
    training seeds 0.. give the targets, the network trains on the GPU, the
    decode gain is chosen on the last tenth of the training seeds, and held-out
    seeds 200.. compare time to zero overflow, learned seed vs zero history."""
    from apgar_amd import prior_study as ps
    from tests.lattice_cases import SIX_LAYER_1536, SIX_LAYER_512, SIX_LAYER_512_DENSE

    spec = SIX_LAYER_512_DENSE if args.prior_family == "six_layer_512_dense" else SIX_LAYER_512
    L = spec["layers"]
    if args.upper_bound:
        res = []
        for s in range(200, 200 + args.upper_bound):
            r = ps.self_seeded_run(spec, s)
            res.append(r)
            log(f"[bench] upper bound seed {s}: plain {r['plain']['passes']} passes "
                f"{r['plain']['time_to_zero_overflow_ms']} ms, own history per cell {r['learned']['passes']} "
                f"passes {r['learned']['time_to_zero_overflow_ms']} ms, per tile {r['tile_mean']['passes']} "
                f"passes {r['tile_mean']['time_to_zero_overflow_ms']} ms")
        tile = [dict(r, learned=r["tile_mean"]) for r in res]
        print(json.dumps({"metric": "time to 0 overflow seeded with the board's own final history / plain",
                          "boards": len(res), "per_cell": ps.compare(res), "per_tile": ps.compare(tile)}))
        return
    seeds = list(range(args.train_seeds))
    path = Path(args.dataset) if args.dataset else None
    t0 = time.perf_counter()
    if path is not None and path.exists():
        d = np.load(path, allow_pickle=False)
        x, y = d["x"], d["y"]
        plain_ms = d["plain_ms"]
        log(f"[bench] loaded {x.shape[0]} training samples from {path}")
    else:
        if args.collect_only and args.train_seed_start:
            seeds = list(range(args.train_seed_start, args.train_seeds))
        x, y, runs = ps.collect(spec, seeds, log=log)
        plain_ms = np.array([r["wall_ms"] for r in runs])
        if path is not None:
            path.parent.mkdir(parents=True, exist_ok=True)
            np.savez(path, x=x, y=y, plain_ms=plain_ms,
                     termination=np.array([r["termination"] for r in runs]),
                     passes=np.array([r["passes"] for r in runs]))
    collect_s = time.perf_counter() - t0
    if args.collect_only:
        print(json.dumps({"collected": int(x.shape[0]), "collect_s": collect_s,
                          "plain_ms_mean": float(np.mean(plain_ms))}))
        return
    n_val = max(1, len(seeds) // 10)
    n_fit = len(seeds) - n_val
    net, rep = ps.train(x[:n_fit], y[:n_fit], L, epochs=args.epochs, batch=32)
    log(f"[bench] trained on {n_fit} boards: {rep}")
    mse, zero_mse = ps.prediction_error(net, x[n_fit:], y[n_fit:])
    mean_map = y[:n_fit].mean(axis=0, keepdims=True)
    mean_mse = float(np.mean((mean_map - y[n_fit:]) ** 2))
    log(f"[bench] validation MSE {mse:.4f} (zero history {zero_mse:.4f}, mean map {mean_mse:.4f})")
    gains = (args.gain,) if args.gain > 0 else tuple(float(g) for g in args.gains.split(","))
    val = {}
    plain_val = {}
    for g in gains if len(gains) > 1 else ():
        res = []
        for s in seeds[n_fit:]:
            r = ps.seeded_run(net, spec, s, g, plain=plain_val.get(s))
            plain_val[s] = r["plain"]
            res.append(r)
        val[g] = ps.compare(res)
        log(f"[bench] validation gain {g}: {val[g]}")
    def score(c):
        return (c.get("learned_feasible", 0), -c.get("time_ratio", 9.0))
    gain = gains[0] if len(gains) == 1 else max(gains, key=lambda g: score(val[g]))
    held = []
    for s in range(200, 200 + args.heldout_seeds):
        r = ps.seeded_run(net, spec, s, gain)
        held.append(r)
        log(f"[bench] held-out seed {s}: plain {r['plain']['passes']} passes "
            f"{r['plain']['time_to_zero_overflow_ms']} ms, learned {r['learned']['passes']} passes "
            f"{r['learned']['time_to_zero_overflow_ms']} ms")
    cmp = ps.compare(held)
    full = []
    for s in range(200, 200 + args.full_size_boards):
        r = ps.seeded_run(net, SIX_LAYER_1536, s, gain, log=log, maximum_passes=8)
        full.append(r)
        log(f"[bench] six_layer_1536 seed {s}: plain {r['plain']['passes']} passes "
            f"{r['plain']['time_to_zero_overflow_ms']} ms, learned {r['learned']['passes']} passes "
            f"{r['learned']['time_to_zero_overflow_ms']} ms")
    line = {
        "metric": "learned-prior vs plain time to 0 overflow",
        "value": cmp.get("time_ratio"),
        "unit": "ratio (learned / plain wall time, boards where both reach zero overflow)",
        "n_gpus": 1,
        "higher_is_better": False,
        "vs_baseline": None,
        "dtype": "bf16 (network), u32 (allocator)",
        "data": "synthetic",
        "config": {"workload": f"{args.prior_family} allocator boards (one 32x32-pad BGA field, {spec['nets']} "
                               "nets of 2-6 pins, capacity 1), exact-order negotiation, 40-pass limit",
                   "train_seeds": f"0-{n_fit - 1}", "validation_seeds": f"{n_fit}-{len(seeds) - 1}",
                   "heldout_seeds": f"200-{199 + args.heldout_seeds}", "decode_gain": gain},
        "heldout": cmp,
        "validation": {str(g): v for g, v in val.items()},
        "prediction_mse": {"network": mse, "zero_history": zero_mse, "training_mean_map": mean_mse},
        "training": {k: (float(v) if isinstance(v, float) else v) for k, v in rep.items()},
        "collect_s": collect_s,
    }
    if full:
        line["six_layer_1536"] = ps.compare(full)
        line["six_layer_1536_runs"] = [{"seed": r["seed"], "plain": r["plain"]["passes"],
                                        "learned": r["learned"]["passes"],
                                        "plain_ms": r["plain"]["time_to_zero_overflow_ms"],
                                        "learned_ms": r["learned"]["time_to_zero_overflow_ms"]} for r in full]
    print(json.dumps(line))
    net.close()


def search_workload(args, workload, steps, warmup, cpu):
    """Times one search workload and returns its JSON object."""
    big = workload == "eight_layer_2048"
    spec = EIGHT_LAYER_2048 if big else FOUR_LAYER_1024
    cost = make_lattice(apgx, spec)
    lat = apgx.Lattice(cost, via_cost=spec["via"], cost_bytes=2)
    opts = dict(tile=tuple(int(v) for v in args.tile.split("x")), frontier_delta=args.delta,
                max_sweeps=args.max_sweeps, goal_bound=not args.no_goal_bound,
                sweep_schedule=args.sweep_schedule,
                use_graph=(1 if big else 0) if args.graph < 0 else args.graph)
    landmark_ms = None
    if args.landmarks:
        landmark_ms = lat.build_landmarks(args.landmarks)
        opts["landmarks"] = True
        opts["landmark_weight"] = args.landmark_weight
        log(f"[bench] {workload}: {args.landmarks} landmark fields built in {landmark_ms:.2f} ms (device)")
    tune = None
    if big:
        count = args.queries or spec["nets"]
        nets, pins = apgx.generate_nets(cost, count, spec["min_pins"], spec["max_pins"],
                                        spec["min_side"], spec["max_side"], spec["seed"])
        if not args.no_tune and args.tile == "0x0" and args.delta == 0 and args.max_sweeps == 0:
            t0 = time.perf_counter()
            opts, report = lat.autotune(nets, pins, sample_nets=256, **opts)
            tune = {"chosen": {"tile": "%dx%d" % opts["tile"], "frontier_delta": opts["frontier_delta"],
                               "max_sweeps": opts["max_sweeps"]},
                    "sample_nets": report["sample_nets"], "candidates": len(report["candidates"]),
                    "tune_s": time.perf_counter() - t0}
            log(f"[bench] autotune: {tune}")
        unit, what = "nets/s", "nets"

        def step(profile=False):
            return lat.route(nets, pins, profile_rounds=1 if profile else 0, **opts)
    else:
        count = args.queries or spec["queries"]
        pairs = apgx.generate_pairs(cost, count, 202)
        unit, what = "queries/s", "queries"

        def step(profile=False):
            return lat.route_pairs(pairs, profile_rounds=1 if profile else 0, **opts)
    log(f"[bench] {workload}: lattice {cost.shape} uploaded; {count} {what}")

    for i in range(warmup):
        _, st = step()
        log(f"[bench] {workload} warmup {i}: device {st.device_ms:.2f} ms rounds {st.rounds}")
    apgx.synchronize()
    t0 = time.perf_counter()
    dev_ms, all_stats = [], []
    reached = 0
    for i in range(steps):
        routes, st = step()
        dev_ms.append(st.device_ms)
        all_stats.append(st)
        reached = sum(r.status == apgx.NET_REACHED for r in routes)
        log(f"[bench] {workload} step {i}: device {st.device_ms:.2f} ms wall {st.wall_ms:.2f} ms "
            f"rounds {st.rounds} tiles {st.tile_visits}")
    apgx.synchronize()
    wall = (time.perf_counter() - t0) / steps
    ms = float(np.mean(dev_ms))

    # Roofline of the dominant kernel (the round kernel), each launch timed with
    # hipEvents on the search stream (profiled step, direct launches).
    _, prof = step(profile=True)
    launches = max(prof.rounds, 1)
    bytes_per_launch = prof.algorithmic_bytes / launches
    avg_launch_s = prof.round_kernel_ms / 1e3 / launches
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    st = all_stats[-1]
    L = cost.shape[0]
    cfg = run_config(workload, opts, args.landmarks)
    log(f"[bench] {workload} profile config: {json.dumps(cfg)}")
    traffic, tsrc = read_profile("traffic", cfg)
    traffic_bytes = None
    if traffic:
        traffic_bytes = traffic.get("traffic_bytes_per_launch")
        if traffic_bytes is None and "traffic_bytes_per_search" in traffic:
            traffic_bytes = traffic["traffic_bytes_per_search"] / launches
    visits = max(st.tile_visits, 1)
    line = {
        "metric": "queries/s & % HBM roofline",
        "value": count / (ms / 1e3),
        "unit": unit,
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "lattice": ("8x2048x2048" if big else "4x1024x1024")
                       + ", uint16 cost U[1,8], 15% keepouts (sides 4-64), vias 20",
            what: count,
            "tile": f"{st.tile_w}x{st.tile_h}",
            "graph": bool(opts["use_graph"]),
            "parallelism": "single GPU",
        },
        "wall_ms_per_step": wall * 1e3,
        f"wall_{what}_per_s": count / wall,
        "reached": reached,
        "rounds": st.rounds,
        "tile_visits": st.tile_visits,
        "tiles_touched": st.tiles_touched,
        "sweep_iterations_per_visit": st.inner_sweeps / visits,
        "sweeps_executed_per_visit": st.layer_sweeps / (visits * L),
        "full_sweeps_per_visit": st.full_sweeps / visits,
        "changing_full_sweep_fraction": st.changing_full_sweeps / max(st.full_sweeps, 1),
        "free_cell_fraction": st.free_cell_visits / max(st.cell_visits, 1),
        "relaxations_per_s": st.relaxations / (ms / 1e3),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_bytes,
            "traffic_source": tsrc,
            "kernel": (f"RegRoundKernel<{L}>" if (st.tile_w, st.tile_h) == (64, 64)
                       else f"RoundKernel<{st.tile_w},{st.tile_h}>"),
            "launches": launches,
            "avg_launch_us": avg_launch_s * 1e6,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
    }
    issue, isrc = read_profile("sq_counters", cfg)
    if issue:
        line["issue"] = {k: issue[k] for k in ("valu_busy", "lds_busy") if k in issue}
        line["issue"]["source"] = isrc
    if landmark_ms is not None:
        # Per-lattice preparation (like the upload): built once, reused by every
        # batch on the lattice.  The second rate charges it to a single step.
        line["landmarks"] = {"count": args.landmarks, "build_ms": landmark_ms,
                             f"{what}_per_s_with_build_in_one_step": count / ((ms + landmark_ms) / 1e3)}
        line["config"]["landmarks"] = args.landmarks
    if tune:
        line["autotune"] = tune
    if cpu:
        threads = args.cpu_threads or min(16, host_cpu()["usable_cpus"])
        if big:
            line["cpu_baseline"] = cpu_nets_baseline(cost, spec["via"], nets, pins,
                                                     args.cpu_sample or 32, threads)
        else:
            line["cpu_baseline"] = cpu_pairs_baseline(cost, spec["via"], pairs,
                                                      args.cpu_sample or 512, threads)
    lat.close()
    return line


def allocator(args):
    """Exact-order CS-RR-v1 on one six_layer_1536 board (2000 nets, capacity 1)
    until zero overflow or the pass limit, with a line per pass on stderr; with
    --prior-gain the same board again with its history seeded by the learned
    prior (trained on the --dataset runs of six_layer_512 boards).  The JSON
    line reports overuse and wall time per pass of each arm and, unless
    --no-cpu, the oracle's exact-order pass time on one thread, extrapolated
    from a 20 s sample whose trees are checked against the GPU's."""
    from apgar_amd import prior_study as ps
    from tests.lattice_cases import SIX_LAYER_1536, SIX_LAYER_512, make_board

    cost, nets, pins = make_board(apgx, SIX_LAYER_1536, args.board_seed)
    lat = apgx.Lattice(cost, via_cost=SIX_LAYER_1536["via"], cost_bytes=2)

    def run(arm, history=None, extra_ms=0.0):
        walls = []

        def on_pass(p, overuse, routed, ms):
            walls.append(ms)
            log(f"[bench] allocator {arm} pass {p}: overuse {overuse} routed {routed}/{len(nets)} wall {ms:.0f} ms")

        r = lat.negotiate(nets, pins, initial_history=history, maximum_passes=args.passes, batch_nets=0,
                          group_nets=args.group_nets, on_pass=on_pass)
        per_pass = [b - a for a, b in zip([0.0] + walls[:-1], walls)]
        return {"termination": ["feasible", "stalled", "pass_limit"][r["termination"]], "passes": r["passes"],
                "overuse_per_pass": r["overuse_per_pass"], "pass_wall_ms": per_pass,
                "mean_exact_order_pass_ms": float(np.mean(per_pass[1:])) if len(per_pass) > 1 else None,
                "wall_ms": r["wall_ms"] + extra_ms, "seed_ms": extra_ms,
                "time_to_zero_overflow_ms": (r["wall_ms"] + extra_ms)
                if r["termination"] == apgx.NEGOTIATION_FEASIBLE else None,
                "search_rounds": r["search_rounds"], "tile_visits": r["tile_visits"],
                "redone_nets": r["redone_nets"], "group_routed_nets": r["group_routed_nets"]}

    plain = run("plain")
    line = {"metric": "exact-order CS-RR-v1 pass time and overuse on six_layer_1536",
            "value": plain["mean_exact_order_pass_ms"], "unit": "ms per exact-order pass (passes 1..)",
            "n_gpus": 1, "higher_is_better": False, "vs_baseline": None, "dtype": "u32", "data": "synthetic",
            "config": {"workload": "six_layer_1536 allocator board (6 x 1536^2, five 32x32-pad BGA fields, "
                                   "2000 nets of 2-6 pins, capacity 1), exact net order",
                       "board_seed": args.board_seed, "pass_limit": args.passes,
                       "order": f"speculative groups of {args.group_nets}" if args.group_nets
                       else "window-disjoint levels"},
            "plain": plain}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_exact_pass_baseline(lat, cost, SIX_LAYER_1536["via"], nets, pins, 20.0)
        log(f"[bench] allocator cpu baseline: {line['cpu_baseline']['sample']}, trees equal: "
            f"{line['cpu_baseline']['sample_trees_equal_gpu']}")
    if args.prior_gain > 0:
        path = Path(args.dataset) if args.dataset else None
        if path is not None and path.exists():
            d = np.load(path, allow_pickle=False)
            x, y = d["x"], d["y"]
        else:
            x, y, _ = ps.collect(SIX_LAYER_512, list(range(args.train_seeds)), log=log)
            if path is not None:
                path.parent.mkdir(parents=True, exist_ok=True)
                np.savez(path, x=x, y=y)
        n_fit = x.shape[0] - max(1, x.shape[0] // 10)
        net, rep = ps.train(x[:n_fit], y[:n_fit], SIX_LAYER_512["layers"], epochs=args.epochs, batch=32)
        log(f"[bench] prior trained on {n_fit} six_layer_512 boards: {rep}")
        big = apgx.Prior(net.layers, cost.shape[1] // ps.TILE, cost.shape[2] // ps.TILE)
        big.set_params(net.get_params())
        t0 = time.perf_counter()
        hist = big.seed_history(lat, pins, gain=args.prior_gain)
        seed_ms = (time.perf_counter() - t0) * 1e3
        line["learned"] = run("learned", history=hist, extra_ms=seed_ms)
        line["learned"]["decode_gain"] = args.prior_gain
        line["learned"]["training_boards"] = int(n_fit)
        # Neither arm need reach zero overflow within the pass limit; the
        # comparison then uses the first wall time at which each arm's overuse
        # is at most the plain arm's final overuse.
        target = plain["overuse_per_pass"][-1]

        def time_to(arm):
            for overuse, t in zip(arm["overuse_per_pass"], np.cumsum(arm["pass_wall_ms"])):
                if overuse <= target:
                    return float(t) + arm["seed_ms"]
            return None

        tp, tl = time_to(plain), time_to(line["learned"])
        line["learned_vs_plain"] = {
            "target_overuse": target, "plain_ms": tp, "learned_ms": tl,
            "time_ratio": (tl / tp) if (tp and tl) else None,
            "zero_overflow_time_ratio": (line["learned"]["time_to_zero_overflow_ms"] / plain["time_to_zero_overflow_ms"])
            if plain["time_to_zero_overflow_ms"] and line["learned"]["time_to_zero_overflow_ms"] else None}
        big.close()
        net.close()
    lat.close()
    print(json.dumps(line))


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
    if args.workload == "learned_prior":
        learned_prior(args)
        return
    if args.workload == "allocator":
        allocator(args)
        return
    line = search_workload(args, args.workload, args.steps, args.warmup, not args.no_cpu)
    if args.workload == "four_layer_1024" and not args.no_sub:
        # The batched uneven persistent launch, timed in the same run.
        line["eight_layer_2048"] = search_workload(args, "eight_layer_2048", args.sub_steps, args.sub_warmup,
                                                   not args.no_cpu)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
