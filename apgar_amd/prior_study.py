"""Learned history-cost prior: training data, training and the held-out
comparison of time to zero overflow (learned seed vs zero-initialised history).

Everything that computes runs in libapgx on the GPU: the board generator's
boards are routed by apgx_negotiate (plain: zero history), their final history
planes are pooled into per-tile targets (apgx_prior_pool_history), the board
maps into input planes (apgx_prior_features), the network is trained by
apgx_prior_train and seeds the history of held-out boards through
apgx_prior_seed_history.  This module only orchestrates those calls.

Time to zero overflow of a run = wall time of apgx_negotiate when it ends
feasible (every net routed, zero overuse); the learned arm adds the wall time
of apgx_prior_seed_history (features, network, decoding, plane readback).
Both arms use the same policy and the same legality checks (the allocator's
own overuse accounting); a run that ends stalled or at the pass limit has no
time to zero overflow.
"""
from __future__ import annotations

import time

import numpy as np

import apgar_amd as apgx

TILE = 64
UNIT = 10.0  # the historical cost increment: history is learned in these units


def board(spec: dict, seed: int):
    p = {k: v for k, v in spec.items() if k != "via"}
    return apgx.generate_board(seed=seed, **p)


def negotiate(lat, nets, pins, history=None, maximum_passes=40, batch_nets=0):
    r = lat.negotiate(nets, pins, initial_history=history, maximum_passes=maximum_passes,
                      batch_nets=batch_nets)
    return r


def summary(r, extra_ms=0.0):
    feasible = r["termination"] == apgx.NEGOTIATION_FEASIBLE
    return {"termination": r["termination"], "passes": r["passes"], "routed": r["routed"],
            "overuse": r["overuse"], "overuse_per_pass": r["overuse_per_pass"],
            "wall_ms": r["wall_ms"] + extra_ms,
            "time_to_zero_overflow_ms": (r["wall_ms"] + extra_ms) if feasible else None}


def collect(spec: dict, seeds, log=print, **negotiate_kw):
    """Plain runs on `seeds`: input planes, targets and per-run summaries."""
    xs, ys, runs = [], [], []
    for s in seeds:
        cost, nets, pins = board(spec, s)
        lat = apgx.Lattice(cost, via_cost=spec["via"], cost_bytes=2)
        r = negotiate(lat, nets, pins, **negotiate_kw)
        xs.append(apgx.prior_features(lat, pins, TILE))
        ys.append(apgx.prior_targets(r["history"], TILE, UNIT))
        runs.append(summary(r))
        lat.close()
        log(f"[prior] plain seed {s}: termination {r['termination']} passes {r['passes']} "
            f"routed {r['routed']}/{len(nets)} wall {r['wall_ms']:.0f} ms")
    return np.stack(xs).astype(np.float32), np.stack(ys).astype(np.float32), runs


def train(x, y, layers, seed=1, **options):
    net = apgx.Prior(layers, x.shape[2], x.shape[3], seed=seed)
    report = net.train(x, y, **options)
    return net, report


def seeded_run(net, spec, seed, gain, plain=None, log=None, **negotiate_kw):
    """One board: plain and learned arms (same policy), learned seeding timed.
    `plain`: a summary of this board's plain run already measured (skips it);
    `log`: called with a line after each arm."""
    cost, nets, pins = board(spec, seed)
    lat = apgx.Lattice(cost, via_cost=spec["via"], cost_bytes=2)
    plain_summary = plain if plain is not None else summary(negotiate(lat, nets, pins, **negotiate_kw))
    if log:
        log(f"[prior] seed {seed} plain: {plain_summary['passes']} passes, termination "
            f"{plain_summary['termination']}, {plain_summary['wall_ms']:.0f} ms")
    prior_net = net
    if net.grid != (cost.shape[1] // TILE, cost.shape[2] // TILE):
        prior_net = apgx.Prior(net.layers, cost.shape[1] // TILE, cost.shape[2] // TILE)
        prior_net.set_params(net.get_params())
    t0 = time.perf_counter()
    hist = prior_net.seed_history(lat, pins, gain=gain)
    seed_ms = (time.perf_counter() - t0) * 1e3
    learned = negotiate(lat, nets, pins, history=hist, **negotiate_kw)
    if log:
        log(f"[prior] seed {seed} learned: {learned['passes']} passes, termination "
            f"{learned['termination']}, {learned['wall_ms'] + seed_ms:.0f} ms")
    lat.close()
    if prior_net is not net:
        prior_net.close()
    return {"seed": seed, "plain": plain_summary, "learned": summary(learned, seed_ms), "seed_ms": seed_ms,
            "seed_history_sum": int(hist.sum(dtype=np.uint64))}


def self_seeded_run(spec, seed, **negotiate_kw):
    """Upper bounds for any history prior on one board: the plain run, then the
    same board seeded with that run's own final history, per cell and as the
    64-cell tile mean (the resolution the network predicts)."""
    cost, nets, pins = board(spec, seed)
    lat = apgx.Lattice(cost, via_cost=spec["via"], cost_bytes=2)
    plain = negotiate(lat, nets, pins, **negotiate_kw)
    hist = plain["history"]
    L, H, W = hist.shape
    tiles = apgx.prior_targets(hist, TILE, UNIT)
    tile_mean = np.repeat(np.repeat(np.rint(UNIT * np.expm1(tiles)), TILE, axis=1), TILE, axis=2)
    tile_mean = tile_mean[:, :H, :W].astype(np.uint32)
    per_cell = negotiate(lat, nets, pins, history=hist, **negotiate_kw)
    per_tile = negotiate(lat, nets, pins, history=tile_mean, **negotiate_kw)
    lat.close()
    return {"seed": seed, "plain": summary(plain), "learned": summary(per_cell), "tile_mean": summary(per_tile)}


def compare(results):
    """Aggregate of seeded_run results: boards where both arms reach zero overflow."""
    both = [r for r in results
            if r["plain"]["time_to_zero_overflow_ms"] is not None
            and r["learned"]["time_to_zero_overflow_ms"] is not None]
    out = {"boards": len(results), "both_feasible": len(both),
           "plain_feasible": sum(r["plain"]["time_to_zero_overflow_ms"] is not None for r in results),
           "learned_feasible": sum(r["learned"]["time_to_zero_overflow_ms"] is not None for r in results)}
    if both:
        tp = sum(r["plain"]["time_to_zero_overflow_ms"] for r in both)
        tl = sum(r["learned"]["time_to_zero_overflow_ms"] for r in both)
        out.update({
            "plain_ms_total": tp, "learned_ms_total": tl, "time_ratio": tl / tp,
            "median_time_ratio": float(np.median([r["learned"]["time_to_zero_overflow_ms"]
                                                  / r["plain"]["time_to_zero_overflow_ms"] for r in both])),
            "plain_passes_mean": float(np.mean([r["plain"]["passes"] for r in both])),
            "learned_passes_mean": float(np.mean([r["learned"]["passes"] for r in both])),
        })
    return out


def prediction_error(net, x, y):
    """Held-out mean squared error of the network against two trivial priors:
    zero history and the training-set mean map."""
    pred = net.forward(x)
    return float(np.mean((pred - y) ** 2)), float(np.mean(y ** 2))
