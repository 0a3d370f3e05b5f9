#!/usr/bin/env python3
"""Voxel edits/s into one shared SVO-DAG on one MI355X.

Workload (the headline configuration): a 64x64 grid of 32^3 trees (MaxDepth 5,
u8 voxels) sharing one interner; a Perlin-dune height field (seeded 2-D gradient
noise, 4 octaves, lacunarity 2, gain 0.5, base period 128) with id 1 below the
surface and 2 at it; one batch per tree, all trees applied in one launch
sequence.  As in the reference's criterion harness (two prebuilt batches applied
alternately to the same trees), every step applies the whole grid again with the
ids swapped, so each step rebuilds all 4096 trees and reclaims the previous ones.

One JSON line: whole-job edits/s of the batched form, p50/p99 per-step latency,
the plain per-tree form on the same GPU, the dominant kernel against the HBM
roofline, and the oracle (literal CPU restatement of the reference) timed on the
host as the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voxel edits/s into shared DAG (batched vs plain, same GPU); p50/p99 batch latency; GB/s"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published peak


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--budget-mib", type=int, default=2048)
    ap.add_argument("--load-factor", type=float, default=0.5)
    ap.add_argument("--plain-trees", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    return ap.parse_args()


def kernel_bytes(name, launches, st, counts):
    """Algorithmic HBM bytes of all launches of one kernel over the timed region.

    st: interner counter deltas over the timed region; counts: per-step sizes.
    Terms are the bytes each unit of work must move at minimum (no re-reads).
    """
    steps = counts["steps"]
    slots1 = counts["brick_slots"] * steps
    touched = counts["touched_bricks"] * steps
    cand1 = counts["cand_bricks"] * steps
    cand_up = counts["cand_upper"] * steps
    slots_up = counts["upper_slots"] * steps
    s = counts["value_size"]
    d = counts["depth"]
    if name == "k_brick_scan":
        # set masks (2 B/slot) + touched values (8s) + old-tree path (8 B per level) + slot write
        return slots1 * 2 + touched * (8 * s + 8 * (d - 1)) + slots1 * 8
    if name == "k_brick_build":
        # masks + slot read/write; touched: values + base node children (64 B);
        # candidates: key (72 B) + hash (8) + slot index (8)
        return slots1 * (2 + 16) + touched * (8 * s + 64) + cand1 * 88
    if name in ("k_probe[bricks]", "k_probe[upper]"):
        c = cand1 if name == "k_probe[bricks]" else cand_up
        # key + hash + slot read (88), one table slot (8), the matched node's key (66),
        # result id / table slot / level slot writes (24)
        return c * (88 + 8 + 66 + 24)
    if name == "k_level_build":
        # 8 child slots (64 B) + slot write (8) per slot; candidate key + hash + slot (88)
        return slots_up * 72 + cand_up * 88
    if name == "k_gc_round":
        freed = st.get("recycled", 0)
        # per freed node: children (64) + refcount RMW of children (8 x 8) + columns (16)
        return freed * (64 + 64 + 16)
    return None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU (this build has no multi-GPU path); got --gpus {a.gpus}",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from voxelis_amd import api
    from voxelis_amd import workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    t_gen = time.perf_counter()
    grid = a.grid
    T = grid * grid
    mA, vA, edits, _ = W.terrain_batches(grid=grid, seed=a.seed)
    vB = vA.copy()
    vB[vA == 1] = 2
    vB[vA == 2] = 1
    E = int(edits.sum())
    t_gen = time.perf_counter() - t_gen
    print(f"[bench] workload: {T} trees, {E} edits per step (host generation {t_gen:.1f}s)",
          file=sys.stderr, flush=True)

    dev = torch.device("cuda", 0)
    mA_t = api.to_device(mA, dev)
    vA_t = api.to_device(vA, dev)
    vB_t = api.to_device(vB, dev)
    it = api.VoxInterner(1, a.budget_mib << 20, a.load_factor)
    f = api.Forest(it, [5] * T)
    viewsA = f.make_views([mA_t[i] for i in range(T)], [vA_t[i] for i in range(T)])
    viewsB = f.make_views([mA_t[i] for i in range(T)], [vB_t[i] for i in range(T)])

    for w in range(a.warmup):
        f.apply_views(viewsA if w % 2 == 0 else viewsB)
    it.sync()
    st_w = f.status.cpu().numpy()
    assert (st_w == 1).all(), f"warmup status {np.unique(st_w)}"
    print(f"[bench] warmup done ({a.warmup} steps)", file=sys.stderr, flush=True)

    counts_before = it.stats()
    if not a.no_profile:
        it.profile(True)
        it.profile_read()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(a.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        views = viewsA if (a.warmup + k) % 2 == 0 else viewsB
        evs[k][0].record()
        f.apply_views(views)
        evs[k][1].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    it.sync()
    prof = it.profile_read() if not a.no_profile else {}
    it.profile(False)
    st_t = f.status.cpu().numpy()
    assert (st_t == 1).all(), f"timed status {np.unique(st_t)}"
    counts_after = it.stats()
    st = {k: counts_after[k] - counts_before[k] for k in counts_after}
    lat = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    wall = t1 - t0
    value = E * a.steps / wall
    leaves, branches = it.counts()
    print(f"[bench] batched: {value / 1e9:.2f} G edits/s, {1e3 * wall / a.steps:.3f} ms/step, "
          f"p50 {np.percentile(lat, 50):.3f} ms p99 {np.percentile(lat, 99):.3f} ms; "
          f"patterns {leaves} leaves {branches} branches", file=sys.stderr, flush=True)

    # ---- roofline of the dominant kernel (HIP events on the library's stream)
    roofline = None
    kern = {}
    if prof:
        steps = a.steps
        counts = {
            "steps": steps, "value_size": 1, "depth": 5, "brick_slots": T * 4096,
            "touched_bricks": int(np.count_nonzero(mA[:, :, 0])),
            "cand_bricks": st["candidates"] and None,
        }
        # candidates per level are not split in the counters: the brick level is the
        # touched bricks that are not uniform; the rest are upper-level keys
        uni = int(np.count_nonzero((mA[:, :, 0] == 0xFF) & (vA == vA[:, :, :1]).all(axis=2)))
        counts["cand_bricks"] = counts["touched_bricks"] - uni
        counts["cand_upper"] = max(st["candidates"] // steps - counts["cand_bricks"], 0)
        counts["upper_slots"] = T * (512 + 64 + 8 + 1)
        dom = max(prof.items(), key=lambda kv: kv[1][0])
        for name, (ms, n) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
            kern[name] = {"ms_per_step": ms / steps, "launches_per_step": n / steps}
        name, (ms, n) = dom
        b = kernel_bytes(name, n, st, counts)
        if b is not None and ms > 0:
            per_launch_bytes = b / n
            avg_s = ms / n / 1e3
            ach = per_launch_bytes / avg_s / 1e9
            roofline = {"kernel": name, "bound": "hbm", "achieved": round(ach, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": None, "avg_launch_ms": round(ms / n, 4),
                        "algorithmic_bytes_per_launch": int(per_launch_bytes)}
        print(f"[bench] kernels: {json.dumps(kern)}", file=sys.stderr, flush=True)

    # ---- plain per-tree form on the same GPU (one apply call per tree)
    plain = None
    if a.plain_trees > 0:
        n_plain = min(a.plain_trees, T)
        it2 = api.VoxInterner(1, a.budget_mib << 20, a.load_factor)
        fs = []
        for i in range(n_plain):
            fi = api.Forest(it2, [5])
            fs.append((fi, fi.make_views([mA_t[i]], [vA_t[i]]), fi.make_views([mA_t[i]], [vB_t[i]])))
        for fi, va, vb in fs:
            fi.apply_views(va)
        it2.sync()
        torch.cuda.synchronize()
        p0 = time.perf_counter()
        for fi, va, vb in fs:
            fi.apply_views(vb)
        torch.cuda.synchronize()
        p1 = time.perf_counter()
        it2.sync()
        plain_edits = int(edits[:n_plain].sum())
        plain = plain_edits / (p1 - p0)
        print(f"[bench] plain: {plain / 1e6:.1f} M edits/s over {n_plain} trees", file=sys.stderr, flush=True)
        it2.close()

    # ---- CPU baseline: the oracle (single thread) on a bounded sample
    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        it_o = O.Interner(1, a.budget_mib << 20)
        done, tsum, esum = 0, 0.0, 0
        while done < T and tsum < a.cpu_seconds:
            t = O.Tree(5)
            ba = O.Batch(5)
            ba.load_dense(mA[done], vA[done])
            t.apply_batch(it_o, ba)
            bb = O.Batch(5)
            bb.load_dense(mA[done], vB[done])
            c0 = time.perf_counter()
            t.apply_batch(it_o, bb)
            tsum += time.perf_counter() - c0
            esum += int(edits[done])
            done += 1
        cpu = {"value": esum / tsum, "unit": "edits/s", "cores": 1, "kind": "port",
               "sample": f"oracle apply_batch replacing {done} of {T} terrain trees "
                         f"(ids swapped, shared interner), {esum} edits, {tsum:.1f}s"}
        print(f"[bench] cpu oracle: {cpu['value'] / 1e6:.1f} M edits/s ({cpu['sample']})",
              file=sys.stderr, flush=True)

    out = {
        "metric": METRIC, "value": value, "unit": "edits/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (seeded Perlin-dune terrain)",
        "config": {"workload": "perlin-dune terrain, 64x64 trees of 32^3 (MaxDepth 5), u8, "
                               "one batch per tree, all trees in one launch sequence, ids "
                               "alternating 1/2 and 2/1 per step",
                   "trees": T, "edits_per_step": E, "interner_budget_mib": a.budget_mib,
                   "load_factor": a.load_factor, "parallelism": "single GPU"},
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "plain_edits_per_s": plain,
        "batched_over_plain": (value / plain) if plain else None,
        "patterns": {"leaves": leaves, "branches": branches},
        "roofline": roofline,
        "kernels": kern,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
