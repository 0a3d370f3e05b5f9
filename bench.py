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
    ap.add_argument("--brick-rounds", type=int, default=None, choices=[0, 1, 2, 4],
                    help="uneven config: brick slots per lane of the brick pass (default: the library's)")
    ap.add_argument("--plain-trees", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--graphs", type=int, default=1,
                    help="replay the captured launch sequence of repeated applies (HIP graphs)")
    ap.add_argument("--tune-workload", default="terrain", choices=["terrain", "uneven"])
    ap.add_argument("--rates", default="200,1000,4000",
                    help="serve: Poisson arrival rates (requests/s) to sweep")
    ap.add_argument("--requests", type=int, default=1500, help="serve: requests per rate")
    ap.add_argument("--train-requests", type=int, default=1500,
                    help="serve: requests logged (with exact dry-run counts) to train the predictor")
    ap.add_argument("--config", default="terrain",
                    choices=["terrain", "plain", "uneven", "gets", "set", "vtm", "serve", "calib", "tune"],
                    help="terrain: the headline (64x64 Perlin-dune trees); plain: one 32^3 tree; "
                         "uneven: u16 trees of depth 5/6/7, dense or sparse; gets: batched reads; "
                         "set: single-voxel set/remove, one voxel per tree per call")
    ap.add_argument("--trees", type=int, default=4096)
    ap.add_argument("--queries", type=int, default=1 << 24)
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 1 GiB device copy (a known-size kernel for counter scaling)")
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
    lead1 = max(cand1 - st.get("merged", 0), 0)  # brick keys that reach the probe
    cand_up = counts["cand_upper"] * steps
    slots_up = counts["upper_slots"] * steps
    s = counts["value_size"]
    if name == "k_brick_build":
        # every brick slot: its set mask (2 B of the mask pair), its share of the recorded
        # height-2 node (1) and the slot write (8); touched bricks: 8 values; non-uniform
        # bricks: the brick-level node (8), its children (64) and their values (8s);
        # keys that reach the probe: candidate key + hash + slot (88)
        return slots1 * (2 + 1 + 8) + touched * 8 * s + cand1 * (8 + 64 + 8 * s) + lead1 * 88
    # a probed key: key + hash + slot read (88), one table slot (8), the matched node's
    # key (66), result id / table slot / level slot writes (24)
    probe_b = 88 + 8 + 66 + 24
    if name == "k_probe[bricks]":
        return lead1 * probe_b
    if name == "k_probe[upper]":
        return cand_up * probe_b
    if name == "k_level_build":
        # 8 child slots (64 B) + slot write (8) per slot; candidate key + hash + slot (88)
        return slots_up * 72 + cand_up * 88
    if name == "k_commit":
        owners = st.get("branch_misses", 0)
        # owners: list entry + table slot + id (24), key + hash (80), node write (64 + 2 + s),
        # table slot (8), 8 child values (8s) and refcount read-modify-writes (8 x 8)
        return owners * (24 + 80 + 66 + s + 8 + 8 * s + 64)
    if name == "k_gc_round":
        freed = st.get("recycled", 0)
        # per freed node: type + children read and clear (2 + 128), pattern slot read + tombstone (16),
        # child refcount read-modify-writes (64), value/ref/gen/free-stack writes (s + 10)
        return freed * (2 + 128 + 16 + 64 + s + 10)
    return None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU (this build has no multi-GPU path); got --gpus {a.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if a.config == "plain":
        return bench_plain(a)
    if a.config == "uneven":
        return bench_uneven(a)
    if a.config == "gets":
        return bench_gets(a)
    if a.config == "set":
        return bench_set(a)
    if a.config == "vtm":
        return bench_vtm(a)
    if a.config == "serve":
        return bench_serve(a)
    if a.config == "calib":
        return bench_calib(a)
    if a.config == "tune":
        return bench_tune(a)
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
    it.set_graphs(bool(a.graphs))
    f = api.Forest(it, [5] * T)
    viewsA = f.make_views([mA_t[i] for i in range(T)], [vA_t[i] for i in range(T)])
    viewsB = f.make_views([mA_t[i] for i in range(T)], [vB_t[i] for i in range(T)])

    for w in range(a.warmup):
        f.apply_views(viewsA if w % 2 == 0 else viewsB)
    it.sync()
    st_w = f.status.cpu().numpy()
    assert (st_w == 1).all(), f"warmup status {np.unique(st_w)}"
    print(f"[bench] warmup done ({a.warmup} steps)", file=sys.stderr, flush=True)

    # ---- timed steps (no per-kernel events: value and latency percentiles)
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
    st_t = f.status.cpu().numpy()
    assert (st_t == 1).all(), f"timed status {np.unique(st_t)}"
    lat = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    wall = t1 - t0
    value = E * a.steps / wall
    leaves, branches = it.counts()
    print(f"[bench] batched: {value / 1e9:.2f} G edits/s, {1e3 * wall / a.steps:.3f} ms/step, "
          f"p50 {np.percentile(lat, 50):.3f} ms p99 {np.percentile(lat, 99):.3f} ms; "
          f"patterns {leaves} leaves {branches} branches", file=sys.stderr, flush=True)

    # ---- the same steps again with per-kernel HIP events on the library's stream
    prof = {}
    counts_before = it.stats()
    if not a.no_profile:
        it.profile(True)
        for k in range(2):  # creates the event pool
            f.apply_views(viewsA if (a.warmup + a.steps + k) % 2 == 0 else viewsB)
        it.profile_read()
        counts_before = it.stats()
        for k in range(a.steps):
            f.apply_views(viewsA if (a.warmup + a.steps + k) % 2 == 0 else viewsB)
        prof = it.profile_read()
        it.profile(False)
    counts_after = it.stats()
    st = {k: counts_after[k] - counts_before[k] for k in counts_after}

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
                        "traffic": _committed_traffic(name, a.grid), "avg_launch_ms": round(ms / n, 4),
                        "algorithmic_bytes_per_launch": int(per_launch_bytes)}
        print(f"[bench] kernels: {json.dumps(kern)}", file=sys.stderr, flush=True)

    # ---- hash-consing passes against HBM, by the single algorithmic-bytes formula:
    # B_edit = E (4 + s); B_hc = N_cand (64+16+8) + N_leafcand (s+16+8)
    #        + N_new (64+s+8+4+2+16) + N_ref 8, over the probe / commit / build passes
    hc = None
    if prof:
        s_ = 1
        n_cand = st["candidates"] / a.steps
        n_leafcand = E  # every set voxel looks its leaf up
        n_new = (st["branch_misses"] + st["leaf_misses"]) / a.steps
        n_ref = 8 * st["branch_misses"] / a.steps + n_new  # child refs of new nodes + frees
        b_edit = E * (4 + s_)
        b_hc = n_cand * 88 + n_leafcand * (s_ + 24) + n_new * (64 + s_ + 8 + 4 + 2 + 16) + n_ref * 8
        hc_kern = ("k_brick_build", "k_brick_retry", "k_probe[bricks]", "k_level_build", "k_probe[upper]",
                   "k_commit")
        t_hc = sum(prof[k][0] for k in hc_kern if k in prof) / a.steps / 1e3
        if t_hc > 0:
            ach = (b_edit + b_hc) / t_hc / 1e9
            hc = {"bytes_edit": int(b_edit), "bytes_hash_consing": int(b_hc), "kernel_s": t_hc,
                  "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(ach / HBM_PEAK_GBS, 4), "kernels": list(hc_kern),
                  "n_cand": int(n_cand), "n_new": int(n_new)}
            print(f"[bench] hash-consing passes: {ach:.0f} GB/s = {100 * ach / HBM_PEAK_GBS:.1f} % of HBM peak "
                  f"({(b_edit + b_hc) / 1e9:.2f} GB in {1e3 * t_hc:.3f} ms)", file=sys.stderr, flush=True)

    # ---- atomic-throughput ceilings (project microbenchmark) next to the apply's rates
    atomics = None
    if not a.no_profile:
        import ctypes as C
        from voxelis_amd._lib import lib
        slots = 1 << 27
        buf = torch.zeros(slots, dtype=torch.int64, device=dev)
        ms = C.c_float(0)
        atomics = {}
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        for kind, name in ((0, "cas64_random"), (1, "add32_random"), (2, "add32_64_hot")):
            ops = 1 << 28
            lib().vx_bench_atomics(stream, kind, buf.data_ptr(), slots, ops, C.byref(ms))  # warm
            buf.zero_()
            rc = lib().vx_bench_atomics(stream, kind, buf.data_ptr(), slots, ops, C.byref(ms))
            if rc == 0 and ms.value > 0:
                atomics[name + "_ops_per_s"] = ops / (ms.value / 1e3)
        del buf
        t_probe = sum(prof[k][0] for k in ("k_probe[bricks]", "k_probe[upper]") if k in prof) / a.steps / 1e3
        if t_probe > 0:
            atomics["apply_probe_cas_per_s"] = (st["branch_misses"] + st["branch_dups"]) / a.steps / t_probe
        print(f"[bench] atomics: {json.dumps({k: f'{v:.3g}' for k, v in atomics.items()})}",
              file=sys.stderr, flush=True)

    if a.calib_copy:
        # the reference kernel for the memory counters: exactly 1 GiB read and written
        import ctypes as C
        from voxelis_amd._lib import lib
        src = torch.ones(1 << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        ms = C.c_float(0)
        lib().vx_bench_copy(C.c_void_p(torch.cuda.current_stream().cuda_stream), dst.data_ptr(), src.data_ptr(),
                            1 << 30, C.byref(ms))
        print(f"[bench] copy 1 GiB: {ms.value:.3f} ms = {2 * (1 << 30) / ms.value / 1e6:.0f} GB/s",
              file=sys.stderr, flush=True)
        del src, dst

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
        # the same steps as the GPU's: every tree rebuilt with the ids of the other phase,
        # step after step, until about cpu_seconds of timed applies (at least one step)
        trees, batches = [], []
        for i in range(T):
            t = O.Tree(5)
            ba, bb = O.Batch(5), O.Batch(5)
            ba.load_dense(mA[i], vA[i])
            bb.load_dense(mA[i], vB[i])
            t.apply_batch(it_o, ba)
            trees.append(t)
            batches.append((ba, bb))
        tsum, esum, nsteps, done = 0.0, 0, 0, 0
        while tsum < a.cpu_seconds:
            for i in range(T):
                c0 = time.perf_counter()
                trees[i].apply_batch(it_o, batches[i][1 - nsteps % 2])
                tsum += time.perf_counter() - c0
                esum += int(edits[i])
                done += 1
                if tsum >= a.cpu_seconds and nsteps > 0:
                    break
            nsteps += 1
        cpu = {"value": esum / tsum, "unit": "edits/s", "cores": 1, "kind": "port",
               "sample": f"oracle apply_batch on the {T} terrain trees (shared interner), ids alternating "
                         f"per step as on the GPU: {done} tree applies ({done / T:.1f} steps), "
                         f"{esum} edits, {tsum:.1f} s timed"}
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
                   "load_factor": a.load_factor, "parallelism": "single GPU",
                   "launch": "HIP-graph replay of the captured launch sequence" if a.graphs
                   else "kernels launched one by one"},
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "plain_edits_per_s": plain,
        "batched_over_plain": (value / plain) if plain else None,
        "patterns": {"leaves": leaves, "branches": branches},
        "roofline": roofline,
        "kernels": kern,
        "hash_consing_roofline": hc,
        "atomics": atomics,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def _committed_traffic(kernel, grid):
    """HBM bytes per launch of `kernel` on the default terrain workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_terrain_4096x32_u8.json")
    if grid != 64 or not os.path.exists(path):
        return None
    d = json.load(open(path))
    v = d["kernels"].get(kernel)
    return int(v["hbm_bytes_scaled"] / max(v["dispatches"], 1)) if v else None


def _timed_steps(f, views_cycle, steps, warmup, it, torch):
    for w in range(warmup):
        f.apply_views(views_cycle[w % len(views_cycle)])
    it.sync()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        evs[k][0].record()
        f.apply_views(views_cycle[(warmup + k) % len(views_cycle)])
        evs[k][1].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    it.sync()
    lat = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    return t1 - t0, lat


def bench_plain(a):
    """Config 1: one 32^3 u8 tree, v = ((x+y+z)&1)+1, apply then 32768 gets."""
    import torch
    from voxelis_amd import api
    from voxelis_amd import workloads as W
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    dense = W.plain_checker(5)
    mA, vA = W.dense_to_batch(dense, 5)
    vB = (3 - vA.astype(np.int64)).astype(np.uint8)  # the other checker phase
    it = api.VoxInterner(1, 256 << 20)
    f = api.Forest(it, [5])
    views = [f.make_views([api.to_device(mA)], [api.to_device(vA)]),
             f.make_views([api.to_device(mA)], [api.to_device(vB)])]
    wall, lat = _timed_steps(f, views, a.steps, a.warmup, it, torch)
    xyz = O.all_coords(5)
    chunk_t = torch.zeros(len(xyz), dtype=torch.int32, device="cuda")
    xyz_t = torch.from_numpy(xyz.reshape(-1)).cuda()
    out = torch.zeros(len(xyz), dtype=torch.uint8, device="cuda")
    pres = torch.zeros(len(xyz), dtype=torch.uint8, device="cuda")
    for _ in range(3):
        f.get_device(chunk_t, xyz_t, out, pres)
    torch.cuda.synchronize()
    g0 = time.perf_counter()
    for _ in range(a.steps):
        f.get_device(chunk_t, xyz_t, out, pres)
    torch.cuda.synchronize()
    gets = len(xyz) * a.steps / (time.perf_counter() - g0)
    # oracle: alternating applies on one tree and full reads
    it_o = O.Interner(1, 256 << 20)
    t_o = O.Tree(5)
    bA, bB = O.Batch(5), O.Batch(5)
    bA.load_dense(mA, vA)
    bB.load_dense(mA, vB)
    t_o.apply_batch(it_o, bA)
    c0 = time.perf_counter()
    n_o = 0
    while time.perf_counter() - c0 < 2.0:
        t_o.apply_batch(it_o, bB if n_o % 2 == 0 else bA)
        n_o += 1
    cpu_apply = 32768 * n_o / (time.perf_counter() - c0)
    c0 = time.perf_counter()
    t_o.get_many(it_o, xyz)
    cpu_get = len(xyz) / (time.perf_counter() - c0)
    print(json.dumps({
        "metric": METRIC, "config": {"workload": "one 32^3 tree (MaxDepth 5), u8, v=((x+y+z)&1)+1 "
                                                 "and its other phase applied alternately; 32768 gets"},
        "value": 32768 * a.steps / wall, "unit": "edits/s", "ms_per_step": 1e3 * wall / a.steps,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "gets_per_s": gets, "counts": it.counts(),
        "cpu_baseline": {"apply_edits_per_s": cpu_apply, "gets_per_s": cpu_get, "cores": 1,
                         "kind": "port"}}))


def bench_uneven(a):
    """Config 3: u16 trees of depth 5/6/7 (uniform), dense ((x+y+z) mod 65535)+1 fills with
    p = 0.5, else sparse uniform random edits at density 0.1/1/10 %; one launch per step;
    steps alternate the values and the values shifted by one, so every tree is rebuilt."""
    import torch
    from voxelis_amd import api
    from voxelis_amd import workloads as W
    depths, mk, vA_list, vB_list, edits, E = _uneven_setup(a, api, W, torch)
    it = api.VoxInterner(2, a.budget_mib << 20, a.load_factor)
    if a.brick_rounds is not None:
        it.set_tuning(brick_rounds=a.brick_rounds)
    f = api.Forest(it, depths)
    views = [f.make_views(mk, vA_list), f.make_views(mk, vB_list)]
    if not a.no_profile:
        it.profile(True)
    wall, lat = _timed_steps(f, views, a.steps, a.warmup, it, torch)
    prof = it.profile_read() if not a.no_profile else {}
    st = f.status.cpu().numpy()
    assert (st == 1).all(), np.unique(st)
    e57 = sum(e for e, d in zip(edits, depths) if d <= 6)
    out = {"metric": METRIC, "value": E * a.steps / wall, "unit": "edits/s",
           "ms_per_step": 1e3 * wall / a.steps, "p50_ms": float(np.percentile(lat, 50)),
           "p99_ms": float(np.percentile(lat, 99)), "dtype": "u16",
           "config": {"workload": "uneven batched edits: depth 5/6/7 trees, dense p=0.5 or sparse "
                                  "0.1/1/10 %, u16, one launch", "trees": a.trees, "edits_per_step": E,
                      "edits_depth_le6": int(e57), "depth7_trees": int(sum(1 for d in depths if d == 7))},
           "counts": it.counts(),
           "kernels": {k: {"ms_per_step": v[0] / (a.steps + a.warmup)} for k, v in prof.items()}}
    print(json.dumps(out))


def _uneven_setup(a, api, W, torch):
    """Device buffers of the uneven workload: per tree masks, values, values shifted by one."""
    spec = W.uneven_spec(a.trees, depths=(5, 6, 7), p_dense=0.5, seed=a.seed + 2)
    t_gen = time.perf_counter()
    dev = torch.device("cuda", 0)
    dense_buf = {}
    mk, vA_list, vB_list, edits = [], [], [], []
    for d, kind, dens, sd in spec:
        if kind == "dense":
            if d not in dense_buf:
                m, v = W.dense_to_batch(W.uneven_dense(d), d)
                vb = np.where(v == 65535, 1, v + 1).astype(np.uint16)
                dense_buf[d] = (api.to_device(m, dev), api.to_device(v, dev), api.to_device(vb, dev))
            m_t, vA_t, vB_t = dense_buf[d]
            edits.append(1 << (3 * d))
        else:
            xyz, vals = W.uneven_tree_edits(d, kind, dens, sd)
            from voxelis_amd.morton import apply_edits_to_dense, dense_batch_arrays
            m, v = dense_batch_arrays(d, np.uint16)
            apply_edits_to_dense(m, v, xyz, vals)
            vb = np.where(v == 0, 0, np.where(v == 65535, 1, v + 1)).astype(np.uint16)
            m_t, vA_t, vB_t = api.to_device(m, dev), api.to_device(v, dev), api.to_device(vb, dev)
            edits.append(int(np.count_nonzero(m[:, 0] != 0) and np.unpackbits(m[:, 0]).sum()))
        mk.append(m_t)
        vA_list.append(vA_t)
        vB_list.append(vB_t)
    E = int(sum(edits))
    depths = [d for d, _, _, _ in spec]
    t_gen = time.perf_counter() - t_gen
    print(f"[bench] uneven: {a.trees} trees, {E} edits per step, depths "
          f"{dict(zip(*np.unique(depths, return_counts=True)))} (gen {t_gen:.1f}s)", file=sys.stderr, flush=True)
    return depths, mk, vA_list, vB_list, edits, E


def bench_tune(a):
    """Autotuner sweep (tile x probe rounds x load factor, then grid sizes) on the terrain
    or the uneven workload; prints every point and the best."""
    import torch
    from voxelis_amd import api
    from voxelis_amd import workloads as W
    from voxelis_amd.autotune import autotune
    dev = torch.device("cuda", 0)
    if a.tune_workload == "terrain":
        T = a.grid * a.grid
        mA, vA, edits, _ = W.terrain_batches(grid=a.grid, seed=a.seed)
        vB = vA.copy()
        vB[vA == 1] = 2
        vB[vA == 2] = 1
        m_t, vA_t, vB_t = api.to_device(mA, dev), api.to_device(vA, dev), api.to_device(vB, dev)
        depths = [5] * T
        mk, va, vb = [m_t[i] for i in range(T)], [vA_t[i] for i in range(T)], [vB_t[i] for i in range(T)]
        vsize, E = 1, int(edits.sum())
    else:
        depths, mk, va, vb, _, E = _uneven_setup(a, api, W, torch)
        vsize = 2

    def make_case(lf):
        it = api.VoxInterner(vsize, a.budget_mib << 20, lf)
        f = api.Forest(it, depths)
        views = [f.make_views(mk, va), f.make_views(mk, vb)]
        return it, (lambda k: f.apply_views(views[k % 2])), it.close

    def log(pt):
        print(f"[tune] {json.dumps(pt)}  ({E / pt['s_per_step'] / 1e9:.2f} G edits/s)", file=sys.stderr,
              flush=True)

    best, points = autotune(make_case, steps=a.steps, warmup=a.warmup, log=log)
    print(json.dumps({"metric": "autotuned apply launch shape", "workload": a.tune_workload,
                      "edits_per_step": E, "best": best,
                      "best_edits_per_s": E / best["s_per_step"], "points": points}))


def bench_gets(a):
    """Batched VoxOpsRead::get over the terrain world: uniform random voxels of random trees."""
    import torch
    from voxelis_amd import api
    from voxelis_amd import workloads as W
    grid = a.grid
    T = grid * grid
    mA, vA, edits, _ = W.terrain_batches(grid=grid, seed=a.seed)
    it = api.VoxInterner(1, a.budget_mib << 20, a.load_factor)
    f = api.Forest(it, [5] * T)
    st = f.apply([api.to_device(mA[i]) for i in range(T)], [api.to_device(vA[i]) for i in range(T)])
    assert (st == 1).all()
    rng = np.random.default_rng(a.seed)
    Q = a.queries
    chunk = rng.integers(0, T, size=Q).astype(np.int32)
    xyz = rng.integers(0, 32, size=(Q, 3)).astype(np.int32)
    order = np.argsort(chunk, kind="stable")  # queries grouped by tree, as a caller batching per chunk would
    res = {}
    for name, idx in (("random", np.arange(Q)), ("grouped", order)):
        chunk_t = torch.from_numpy(chunk[idx]).cuda()
        xyz_t = torch.from_numpy(np.ascontiguousarray(xyz[idx]).reshape(-1)).cuda()
        out = torch.zeros(Q, dtype=torch.uint8, device="cuda")
        pres = torch.zeros(Q, dtype=torch.uint8, device="cuda")
        it.profile(True)
        for _ in range(3):
            f.get_device(chunk_t, xyz_t, out, pres)
        torch.cuda.synchronize()
        it.profile_read()
        g0 = time.perf_counter()
        for _ in range(a.steps):
            f.get_device(chunk_t, xyz_t, out, pres)
        torch.cuda.synchronize()
        wall = time.perf_counter() - g0
        prof = it.profile_read()
        ms, n = prof.get("k_get", (0.0, 1))
        res[name] = {"gets_per_s": Q * a.steps / wall, "kernel_ms": ms / n,
                     "kernel_gets_per_s": Q / (ms / n / 1e3)}
    print(json.dumps({"metric": "batched get/s", "config": {"workload": f"{Q} uniform random voxel reads "
                      f"over the 64x64 terrain world (4096 trees of 32^3)"}, "results": res}))


def bench_set(a):
    """Single-voxel VoxOpsWrite::set over the terrain world: every call sets (or, one time
    in four, removes) one uniform random voxel in each of the 4096 trees, values 1..3."""
    import torch
    from voxelis_amd import api
    from voxelis_amd import workloads as W
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    grid = a.grid
    T = grid * grid
    mA, vA, edits, _ = W.terrain_batches(grid=grid, seed=a.seed)
    it = api.VoxInterner(1, a.budget_mib << 20, a.load_factor)
    f = api.Forest(it, [5] * T)
    st = f.apply([api.to_device(mA[i]) for i in range(T)], [api.to_device(vA[i]) for i in range(T)])
    assert (st == 1).all()
    rng = np.random.default_rng(a.seed)
    K = a.warmup + a.steps
    xyz = rng.integers(0, 32, size=(K, T, 3)).astype(np.int32)
    vals = rng.integers(1, 4, size=(K, T)).astype(np.uint8)
    vals[rng.random((K, T)) < 0.25] = 0
    xyz_t = [torch.from_numpy(xyz[k].reshape(-1)).cuda() for k in range(K)]
    val_t = [torch.from_numpy(vals[k]).cuda() for k in range(K)]
    for k in range(a.warmup):
        f.set_voxels(xyz_t[k], val_t[k], sync=False)
    it.sync()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    changed = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j in range(a.steps):
        evs[j][0].record()
        f.set_voxels(xyz_t[a.warmup + j], val_t[a.warmup + j], sync=False)
        evs[j][1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    it.sync()
    lat = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    changed = int((f.status.cpu().numpy() == 1).sum())
    # the oracle (one thread) on the same trees and voxels, for about cpu_seconds
    it_o = O.Interner(1, a.budget_mib << 20)
    trees = []
    for i in range(T):
        t = O.Tree(5)
        b = O.Batch(5)
        b.load_dense(mA[i], vA[i])
        t.apply_batch(it_o, b)
        trees.append(t)
    done, tsum = 0, 0.0
    for k in range(K):  # one C call per round of T sets (no per-set call overhead)
        c0 = time.perf_counter()
        O.set_round(it_o, trees, xyz[k], vals[k])
        tsum += time.perf_counter() - c0
        done += T
        if tsum >= a.cpu_seconds:
            break
    res = {"sets_per_s": T * a.steps / wall, "ms_per_call": 1e3 * wall / a.steps,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "max_ms": float(lat.max()), "changed_in_last_call": changed,
           "cpu_baseline": {"value": done / tsum, "unit": "sets/s", "cores": 1, "kind": "port",
                            "sample": f"oracle VoxOpsWrite::set, {done} sets on the same trees and voxels "
                                      f"(rounds of {T} in one C call), {tsum:.2f} s"}}
    print(f"[bench] set: {res['sets_per_s'] / 1e6:.1f} M sets/s, {res['ms_per_call']:.3f} ms per call of {T} "
          f"sets (p50 {res['p50_ms']:.3f}, p99 {res['p99_ms']:.3f}); oracle {done / tsum / 1e6:.2f} M sets/s",
          file=sys.stderr, flush=True)
    print(json.dumps({"metric": "single-voxel sets/s", "config": {
        "workload": f"one uniform random voxel set (values 1..3, a quarter removes) in each of the {T} "
                    f"terrain trees (32^3, u8) per call"}, "results": res}))


def bench_vtm(a):
    """VTM snapshot (VoxModel::serialize / deserialize) of two resident worlds: the 64x64
    terrain (u8, few distinct nodes) and 1024 u16 trees of depth 5/6 with random values
    (almost every brick distinct).  Export = the data section on the host; import = a
    fresh interner rebuilt from it."""
    import torch
    from voxelis_amd import api
    from voxelis_amd import workloads as W
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    sys.path.insert(0, ROOT)
    from oracle import vtm_oracle as V
    res = {}
    worlds = []
    mA, vA, _, _ = W.terrain_batches(grid=a.grid, seed=a.seed)
    worlds.append(("terrain_4096x32_u8", 1, [5] * len(mA), list(mA), list(vA)))
    spec = W.uneven_spec(1024, depths=(5, 6), seed=a.seed)
    ub = W.uneven_batches(spec)
    worlds.append(("uneven_1024_u16", 2, [d for d, _, _, _ in spec], [m for m, _, _ in ub], [v for _, v, _ in ub]))
    for name, vs, depths, ms, vals in worlds:
        it = api.VoxInterner(vs, a.budget_mib << 20, a.load_factor)
        order = np.argsort(depths, kind="stable")
        # one VTM model has one depth: the deepest trees form the model
        d = max(depths)
        sel = [int(i) for i in order if depths[i] == d]
        f = api.Forest(it, [d] * len(sel))
        st = f.apply([api.to_device(ms[i]) for i in sel], [api.to_device(vals[i]) for i in sel])
        assert (st == 1).all()
        pos = np.array([[i % 64, 0, i // 64] for i in range(len(sel))], dtype=np.int32)
        leaves, branches = it.counts()
        f.serialize(pos)  # warm
        t0 = time.perf_counter()
        for _ in range(3):
            data = f.serialize(pos)
        t_exp = (time.perf_counter() - t0) / 3
        t_imp = 0.0
        for _ in range(3):
            it2 = api.VoxInterner(vs, a.budget_mib << 20, a.load_factor)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            f2, _ = api.Forest.deserialize(it2, data, d)
            t_imp += time.perf_counter() - t0
            assert f2.n == f.n
            del f2, it2
        t_imp /= 3
        nodes = leaves + branches - 1
        # the oracle's restatement (pure Python) on the parsed section, rebuilt once
        c0 = time.perf_counter()
        parsed = V.deserialize(data, vs)
        t_parse = time.perf_counter() - c0
        res[name] = {"trees": f.n, "depth": d, "nodes": nodes, "bytes": len(data),
                     "export_s": t_exp, "export_nodes_per_s": nodes / t_exp, "export_MB_per_s": len(data) / t_exp / 1e6,
                     "import_s": t_imp, "import_nodes_per_s": nodes / t_imp, "import_MB_per_s": len(data) / t_imp / 1e6,
                     "cpu_baseline": {"value": nodes / t_parse, "unit": "nodes/s parsed", "cores": 1, "kind": "port",
                                      "sample": f"oracle VoxModel::deserialize parse (pure Python) of the same "
                                                f"{len(data)} bytes, {t_parse:.2f} s"}}
        assert len(parsed[0]) + len(parsed[1]) == nodes
        print(f"[bench] vtm {name}: {nodes} nodes, {len(data) / 1e6:.1f} MB; export {1e3 * t_exp:.1f} ms "
              f"({nodes / t_exp / 1e6:.1f} M nodes/s), import {1e3 * t_imp:.1f} ms ({nodes / t_imp / 1e6:.1f} M nodes/s); "
              f"oracle parse {nodes / t_parse / 1e6:.2f} M nodes/s", file=sys.stderr, flush=True)
        del f, it
    print(json.dumps({"metric": "VTM export / import", "config": {"workload": "terrain 64x64 (u8) and 1024 "
                      "uneven u16 trees (the deepest of depth 5/6)"}, "results": res}))


# Published single-thread apply_batch / get rates of the reference (Apple M3 Max,
# i32, criterion mean; edits or gets per second derived from the batch size).
_PUBLISHED = {("set_uniform", 32): 1.418e9, ("set_uniform", 64): 1.366e9,
              ("set_sum", 32): 168.8e6, ("set_sum", 64): 170.6e6,
              ("set_checkerboard", 32): 140.4e6, ("set_checkerboard", 64): 139.8e6,
              ("set_sparse_fill", 32): 30.3e6, ("set_sparse_fill", 64): 27.5e6,
              ("get_full_sum", 32): 178.5e6, ("get_full_sum", 64): 158.6e6}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_calib(a):
    """The oracle (single thread) on the reference's own criterion shapes (i32, two
    prebuilt batches applied alternately to one tree, voxtree_bench.rs:553-579), next to
    the published M3 Max rates: how this host's CPU baseline relates to the reference's."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    def shape_batches(name, n):
        y, z, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)  # y, z, x loop order
        s = xyz.sum(axis=1).astype(np.int64)
        if name == "set_uniform":
            return xyz, np.ones(len(s)), np.full(len(s), 2)
        if name == "set_sum":
            return xyz, s + 1, s + 1000
        if name == "set_checkerboard":
            keep = s % 2 == 0
            return xyz[keep], np.ones(keep.sum()), np.full(keep.sum(), 2)
        keep = (xyz % 4 == 0).all(axis=1)  # set_sparse_fill: every 4th on each axis
        return xyz[keep], np.ones(keep.sum()), np.full(keep.sum(), 2)

    rows = []
    per = max(a.cpu_seconds / 10.0, 0.3)
    for name in ("set_uniform", "set_sum", "set_checkerboard", "set_sparse_fill", "get_full_sum"):
        for n in (32, 64):
            d = int(np.log2(n))
            it = O.Interner(4, 24 << 20)
            t = O.Tree(d)
            if name == "get_full_sum":
                xyz, v0, _ = shape_batches("set_sum", n)
                b = O.Batch(d)
                b.set_many(xyz, v0.astype(np.uint64))
                t.apply_batch(it, b)
                c0 = time.perf_counter()
                k = 0
                while time.perf_counter() - c0 < per:
                    t.get_many(it, xyz)
                    k += 1
                rate = k * len(xyz) / (time.perf_counter() - c0)
                edits = len(xyz)
            else:
                xyz, v0, v1 = shape_batches(name, n)
                bs = []
                for vals in (v0, v1):
                    b = O.Batch(d)
                    b.set_many(xyz, vals.astype(np.uint64))
                    bs.append(b)
                t.apply_batch(it, bs[0])
                k = 1
                c0 = time.perf_counter()
                while time.perf_counter() - c0 < per:
                    t.apply_batch(it, bs[k % 2])
                    k += 1
                rate = (k - 1) * len(xyz) / (time.perf_counter() - c0)
                edits = len(xyz)
            pub = _PUBLISHED[(name, n)]
            rows.append({"bench": name, "size": n, "per_batch": edits, "rate_per_s": rate,
                         "published_m3max_per_s": pub, "ratio_to_published": rate / pub})
            print(f"[calib] {name} {n}^3: {rate / 1e6:.1f} M/s (M3 Max {pub / 1e6:.1f} M/s, "
                  f"ratio {rate / pub:.2f})", file=sys.stderr, flush=True)
    ratios = [r["ratio_to_published"] for r in rows]
    print(json.dumps({"metric": "CPU-baseline calibration: oracle vs published reference rates",
                      "cpu": _cpu_model(), "nproc": os.cpu_count(), "threads": 1,
                      "geomean_ratio": float(np.exp(np.mean(np.log(ratios)))), "rows": rows}))


def _serve_world(api, serve, W, shape, budget):
    """u16 world of shape trees (32^3 each): dunes in the bottom layer, air above."""
    world = serve.World(2, budget, shape, depth=5)
    sx, sy, sz = shape
    heights = W.terrain_heights(grid=max(sx, sz), chunk=32, seed=5)
    trees = [world.tree_index(tx, 0, tz) for tz in range(sz) for tx in range(sx)]
    ms, vs = [], []
    for tz in range(sz):
        for tx in range(sx):
            m, v = W.dense_to_batch(W.terrain_dense(heights, tx, tz).astype(np.uint16), 5)
            ms.append(api.to_device(m))
            vs.append(api.to_device(v))
    sub, idx = world.sub(trees)
    sub.apply(ms, vs)
    world.forest.roots[idx] = sub.roots
    world.kind[trees] = 2
    world.it.sync()
    return world


def bench_serve(a):
    """Served stream: 16384-tree depth-5 u16 world resident in HBM (interner budget
    64 GiB); Poisson arrivals at swept rates; 70 % get batches of 1-65536 uniform
    random voxels, 25 % sphere edits (radius 2-40, random centre, u16 id or carving),
    5 % whole-tree fill/clear; admission by memory budget with the learned predictor
    (trained first on the same kind of stream, exact dry-run counts as targets)."""
    import torch
    from voxelis_amd import api, serve
    from voxelis_amd import predictor as PR
    from voxelis_amd import workloads as W
    shape = (32, 16, 32)
    budget = a.budget_mib << 20 if a.budget_mib != 2048 else 64 << 30
    t0 = time.perf_counter()
    world = _serve_world(api, serve, W, shape, budget)
    print(f"[serve] world {world.n} trees, {world.it.alive_nodes()} nodes "
          f"({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    # ---- predictor: log (features, exact growth) pairs, train on the GPU
    t0 = time.perf_counter()
    reqs = serve.make_requests(world, a.train_requests, rate=1e6, seed=a.seed + 50, carve=True, max_gets=1024)
    logger = serve.Server(world, log_exact=True)
    logger.serve(reqs, realtime=False)
    pred, margin, rep = PR.Predictor.fit_from_log(logger.log, holdout=0.3, quantile=0.995)
    print(f"[serve] predictor: {rep['pairs']} (request, tree) pairs from {rep['requests']} requests, "
          f"held-out MAE {rep['mae_nodes']:.1f} nodes (hard bound {rep['hard_bound_mae_nodes']:.1f}), "
          f"margin {margin:.2f}, held-out under-prediction {100 * rep['heldout_under_rate']:.2f} % "
          f"({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    results = []
    for ri, rate in enumerate(float(x) for x in a.rates.split(",")):
        reqs = serve.make_requests(world, a.requests, rate=rate, seed=a.seed + ri, carve=True)
        srv = serve.Server(world, predictor=pred, margin=margin)
        torch.cuda.synchronize()
        lat, wall = srv.serve(reqs, realtime=True)
        edits = sum(r.edits for r in reqs if r.kind in ("sphere", "fill"))
        gets = sum(len(r.payload[0]) for r in reqs if r.kind == "get")
        n_edit = sum(1 for r in reqs if r.kind in ("sphere", "fill"))
        span = max(r.done for r in reqs if r.done is not None)
        res = {"rate": rate, "requests": len(reqs), "p50_ms": 1e3 * float(np.percentile(lat, 50)),
               "p99_ms": 1e3 * float(np.percentile(lat, 99)), "edits_per_s": edits / span,
               "gets_per_s": gets / span, "cycles": srv.cycles, "wall_s": wall,
               "rolled_back_applies": srv.fallbacks,
               "fallback_rate": srv.fallback_requests / max(n_edit, 1), "exact_recounts": srv.recounts}
        results.append(res)
        print(f"[serve] rate {rate:.0f}/s: p50 {res['p50_ms']:.2f} ms p99 {res['p99_ms']:.2f} ms, "
              f"{res['edits_per_s'] / 1e6:.1f} M edits/s, {res['gets_per_s'] / 1e6:.1f} M gets/s, "
              f"{srv.cycles} cycles, fallback {100 * res['fallback_rate']:.2f} %", file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "served requests: p50/p99 latency, edits/s, gets/s at swept Poisson rates",
        "config": {"workload": "16384-tree depth-5 u16 world (32x16x32 trees, dunes in the bottom layer); "
                               "70% gets of 1-65536 voxels, 25% spheres r 2-40 (fill or carve), 5% "
                               "fill/clear; admission by memory budget with the learned predictor",
                   "interner_budget_gib": budget / (1 << 30), "requests_per_rate": a.requests},
        "predictor": rep,
        "results": results}))


if __name__ == "__main__":
    main()
