"""Autotuner of the apply's launch shape: tile size (lanes per block of the build /
probe / commit passes), the probe's key merge (dedup_rounds: 0 = off), commit and
GC grid sizes, how many tiles a probe block packs (at the brick level and above),
the brick pass's key-merge rounds and slots per lane, and the pattern table's load
factor.

The search is exhaustive over tile x dedup rounds for every load factor, then
coordinate descent over the other fields from the best point.
Each point times `steps` applies (after `warmup`) of the caller's workload with
the device clock.
"""
import json
import os
import time

import torch

TUNING_FIELDS = ("tile", "dedup_rounds", "commit_grid", "gc_grid", "probe_blocks_bricks", "probe_blocks",
                 "brick_merge_rounds", "brick_rounds")
_PICKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def committed_pick(workload):
    """The autotuner's pick for a named workload from the package's tuning.json (the best
    point of a sweep on that workload: load factor and launch shape), or None."""
    if not os.path.exists(_PICKS):
        return None
    return json.load(open(_PICKS)).get(workload)


def apply_pick(interner, pick):
    """Set a pick's launch shape on an interner (its load factor is a constructor argument)."""
    interner.set_tuning(**{f: int(pick[f]) for f in TUNING_FIELDS if f in pick})


def _time(step, steps, warmup):
    for k in range(warmup):
        step(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step(warmup + k)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def autotune(make_case, load_factors=(0.5, 0.6, 0.7, 0.85), tiles=(128, 256, 512), dedups=(0, 2, 4, 8),
             commit_grids=(256, 1024, 4096), gc_grids=(256, 1024, 4096),
             probe_blocks_bricks=(0, 8192, 16384, 32768), probe_blocks=(0, 256, 512, 1024),
             brick_merge_rounds=(0, 2, 4, 8), brick_rounds=(0, 1, 2, 4), steps=6, warmup=2, log=None):
    """make_case(load_factor) -> (interner, step(k), close()); returns (best point, all points).

    A point is {"load_factor", "tile", "dedup_rounds", "commit_grid", "gc_grid",
    "probe_blocks_bricks", "probe_blocks", "brick_merge_rounds", "brick_rounds", "s_per_step"}."""
    points = []
    best = None
    for lf in load_factors:
        it, step, close = make_case(lf)
        try:
            for tile in tiles:
                for dd in dedups:
                    it.set_tuning(tile=tile, dedup_rounds=dd, commit_grid=1024, gc_grid=1024,
                                  probe_blocks_bricks=32768, probe_blocks=512, brick_merge_rounds=4, brick_rounds=0)
                    t = _time(step, steps, warmup)
                    pt = {"load_factor": lf, "tile": tile, "dedup_rounds": dd, "commit_grid": 1024,
                          "gc_grid": 1024, "probe_blocks_bricks": 32768, "probe_blocks": 512,
                          "brick_merge_rounds": 4, "brick_rounds": 0, "s_per_step": t}
                    points.append(pt)
                    if log:
                        log(pt)
                    if best is None or t < best["s_per_step"]:
                        best = dict(pt)
        finally:
            close()
    # coordinate descent over the grid sizes and the probe packing at the best
    # (load factor, tile, rounds)
    it, step, close = make_case(best["load_factor"])
    try:
        for field, choices in (("commit_grid", commit_grids), ("gc_grid", gc_grids),
                               ("probe_blocks_bricks", probe_blocks_bricks), ("probe_blocks", probe_blocks),
                               ("brick_merge_rounds", brick_merge_rounds), ("brick_rounds", brick_rounds)):
            for v in choices:
                cand = dict(best)
                cand[field] = v
                it.set_tuning(**{f: cand[f] for f in ("tile", "dedup_rounds", "commit_grid", "gc_grid",
                                                       "probe_blocks_bricks", "probe_blocks",
                                                       "brick_merge_rounds", "brick_rounds")})
                cand["s_per_step"] = _time(step, steps, warmup)
                points.append(cand)
                if log:
                    log(cand)
                if cand["s_per_step"] < best["s_per_step"]:
                    best = cand
    finally:
        close()
    return best, points
