#!/usr/bin/env python3
"""Voxel edits/s into one shared SVO-DAG on one MI355X.

Headline workload (default, --config uneven): 4096 trees sharing one u16 interner, each
MaxDepth drawn uniformly from {5, 6, 7} (32^3 to 128^3); per tree either a dense fill
v = ((x+y+z) mod 65535) + 1 (probability 0.5) or sparse uniform random edits at a
density drawn from {0.1 %, 1 %, 10 %} with values uniform in [1, 65535]; one batch per
tree, all trees in one launch sequence (replayed from a captured HIP graph).  As the
reference's criterion harness applies two prebuilt batches alternately to the same
trees (voxtree_bench.rs:553-579), steps alternate the values v and v+1, so every step
rebuilds every tree and reclaims the previous one.  --config terrain is the 64x64 grid
of 32^3 u8 Perlin-dune trees.

One JSON line: whole-job edits/s of the batched form; p50/p99 per-launch latency over at
least --lat-steps launches; the same trees sliced into launches small enough for a p99
of at most --p99-ms, with their edits/s; the trees of depth <= 6 and of depth 7 on their
own; the plain per-tree form on the same GPU; every kernel's algorithmic bytes against
the HBM roofline, with the committed rocprofv3 traffic beside them; and the oracle (a
literal CPU restatement of the reference) on a bounded sample as the CPU baseline.
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
UNEVEN_WORKLOAD = ("uneven batched edits: 4096 u16 trees, MaxDepth uniform over {5,6,7}, dense "
                   "((x+y+z) mod 65535)+1 with p=0.5 else sparse uniform random at 0.1/1/10 %, one batch "
                   "per tree, all trees in one launch sequence, values alternating v and v+1 per step")
TERRAIN_WORKLOAD = ("perlin-dune terrain, 64x64 trees of 32^3 (MaxDepth 5), u8, one batch per tree, all "
                    "trees in one launch sequence, ids alternating 1/2 and 2/1 per step")
HC_KERNELS = ("k_brick_build", "k_brick_retry", "k_probe[bricks]", "k_level_build", "k_probe[upper]", "k_commit")
# rocprofv3 kernel symbol -> the library's kernel name
TRACE_NAMES = {"k_probe_bricks": "k_probe[bricks]", "k_probe": "k_probe[upper]"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--lat-steps", type=int, default=200,
                    help="launches timed one by one (HIP events) for the p50/p99 latency")
    ap.add_argument("--p99-ms", type=float, default=2.0, help="latency bound of the sliced form")
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--budget-mib", type=int, default=None,
                    help="interner budget (default 16384 for uneven, 2048 for terrain)")
    ap.add_argument("--load-factor", type=float, default=None,
                    help="pattern table load factor (default: the autotuner's pick, else 0.5)")
    ap.add_argument("--tuning", default="pick", choices=["pick", "default"],
                    help="pick: the autotuner's committed pick for the workload (voxelis_amd/tuning.json); "
                         "default: the library's built-in launch shape")
    ap.add_argument("--brick-rounds", type=int, default=None, choices=[0, 1, 2, 4],
                    help="brick slots per lane of the brick pass (default: the library's)")
    ap.add_argument("--plain-trees", type=int, default=96)
    ap.add_argument("--cpu-trees", type=int, default=256, help="uneven: trees in the CPU baseline's sample")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--quick", action="store_true",
                    help="timed steps only (for profiler runs): no latency loop, slices, subsets, plain or CPU legs")
    ap.add_argument("--graphs", type=int, default=1,
                    help="replay the captured launch sequence of repeated applies (HIP graphs)")
    ap.add_argument("--tune-workload", default="terrain", choices=["terrain", "uneven"])
    ap.add_argument("--rates", default="200,1000,4000",
                    help="serve: Poisson arrival rates (requests/s) to sweep")
    ap.add_argument("--requests", type=int, default=1500, help="serve: requests per rate")
    ap.add_argument("--train-requests", type=int, default=1500,
                    help="serve: requests logged (with exact dry-run counts) to train the predictor")
    ap.add_argument("--config", default="uneven",
                    choices=["uneven", "terrain", "plain", "gets", "set", "vtm", "serve", "calib", "tune",
                             "predictor", "edits"],
                    help="uneven: the headline (4096 u16 trees of depth 5/6/7, dense or sparse); terrain: "
                         "64x64 Perlin-dune u8 trees; plain: one 32^3 tree; gets: batched reads; set: "
                         "single-voxel set/remove, one voxel per tree per call")
    ap.add_argument("--trees", type=int, default=4096)
    ap.add_argument("--pairs", type=int, default=1 << 20,
                    help="predictor: (features, exact growth) pairs to log before training")
    ap.add_argument("--spheres", type=int, default=64, help="edits: sphere edits per apply")
    ap.add_argument("--admit-steps", type=int, default=20,
                    help="uneven: applies run under predictor-driven admission (0: none)")
    ap.add_argument("--queries", type=int, default=1 << 24)
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 1 GiB device copy (a known-size kernel for counter scaling)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarize two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE; csv output, each with "
                         "--calib-copy) into per-kernel HBM bytes, scaled by the copy kernel's readings")
    ap.add_argument("--trace-summary", metavar="TRACE_DIR",
                    help="summarize a rocprofv3 --kernel-trace --stats run (csv output) per kernel")
    ap.add_argument("--skip-applies", type=int, default=1,
                    help="summary modes: applies left out at the start (the first builds from empty trees)")
    ap.add_argument("--out", default=None, help="summary modes: also write the JSON there")
    a = ap.parse_args()
    a.pick = None
    wl_key = {"uneven": "uneven_4096_u16", "terrain": "terrain_4096x32_u8"}.get(a.config)
    if a.tuning == "pick" and wl_key and not (a.config == "uneven" and (a.trees != 4096 or a.seed != 1)) \
            and not (a.config == "terrain" and (a.grid != 64 or a.seed != 1)):
        from voxelis_amd.autotune import committed_pick
        a.pick = committed_pick(wl_key)
    if a.load_factor is None:
        a.load_factor = float(a.pick["load_factor"]) if a.pick else 0.5
    if a.budget_mib is None:
        big = a.config in ("uneven", "predictor", "edits") or (a.config == "tune" and a.tune_workload == "uneven")
        a.budget_mib = 16384 if big else 2048
    return a


# ---------------------------------------------------------------- algorithmic bytes

def algorithmic_bytes(U):
    """HBM bytes each kernel must move per step at minimum in this design (no re-reads),
    from the step's unit counts U (host-side workload sizes and the device counters).

    Every term is a read or write of a structure the kernel cannot avoid: batch masks and
    values, level results, candidate keys (72 B key + 8 B hash + 8 B slot), one table slot
    per probing key, the matched node's key on a hit (66 B), the owner's candidate on a
    duplicate (72 B), node writes of new nodes and the columns a free clears.  Reference
    count updates of children are not counted (their number is not counted on the device).
    """
    s = U["value_size"]
    n1 = U["brick_hits"] + U["brick_misses"] + U["brick_dups"]
    hu, mu, du = (U["branch_hits"] - U["brick_hits"], U["branch_misses"] - U["brick_misses"],
                  U["branch_dups"] - U["brick_dups"])
    nu = hu + mu + du
    lead1 = max(U["brick_keys"] - U["brick_merged"], 0)
    leadu = max((U["candidates"] - U["brick_keys"]) - (U["merged"] - U["brick_merged"]), 0)
    probe_miss = 2 + s + 8 + 8 * s  # generation, LOD value, owner entry, the children's values
    return {
        # per tree record and descent: the recorded height-2 nodes (64 B per level-3 slot)
        # and the nodes above them; u16: the leaf id of every value (map + generation + id)
        "k_prepare": U["level3_slots"] * 72 + U["trees"] * 48 + (65536 * 14 if s == 2 else 0),
        # per brick slot: the mask pair (2), its share of the recorded height-2 node (1), the
        # result (8); a touched brick: its 8 values; a mixed brick: the old brick-level node
        # (8, the leaf-brick check); a partially written one also its children (64, the
        # unwritten slots); a key that leaves the block: the candidate (88)
        "k_brick_build": (U["brick_slots"] * 11 + U["touched_bricks"] * 8 * s + U["mixed_bricks"] * 8
                          + U["partial_bricks"] * 64 + lead1 * 88),
        "k_probe[bricks]": n1 * 120 + U["brick_hits"] * 66 + U["brick_dups"] * 72 + U["brick_misses"] * probe_miss,
        # per slot: 8 child results, the recorded old node, the result; a leader's candidate
        "k_level_build": U["upper_slots"] * 80 + leadu * 88,
        "k_probe[upper]": nu * 120 + hu * 66 + du * 72 + mu * probe_miss,
        # per owner: list entry, table slot and id (24), key and hash (80), the node (64 + 2),
        # the table slot (8)
        "k_commit": U["branch_misses"] * 178,
        # per freed node: frontier entry, types/mask and generation (read and written), the
        # children (read), the table slot (read and written), free-stack entry
        "k_gc_round": U["recycled"] * (4 + 4 + 4 + 64 + 16 + 4),
    }


def _units(wl, st, steps):
    U = dict(wl["units"])
    for k in ("branch_hits", "branch_misses", "branch_dups", "candidates", "merged", "brick_keys",
              "brick_merged", "brick_hits", "brick_misses", "brick_dups", "recycled"):
        U[k] = st.get(k, 0) / steps
    return U


def _committed_traffic(profile):
    """Per-kernel HBM bytes per dispatch from a committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (copy-calibrated), or {} when the workload has none."""
    path = os.path.join(ROOT, "profiles", profile)
    if not profile or not os.path.exists(path):
        return {}
    d = json.load(open(path))
    return {k: v["hbm_bytes_scaled"] / max(v["dispatches"], 1) for k, v in d["kernels"].items()}


def _roofline(prof, st, steps, wl):
    """Per-kernel table, the dominant kernel's roofline and the hash-consing passes' roofline."""
    U = _units(wl, st, steps)
    B = algorithmic_bytes(U)
    traffic = _committed_traffic(wl["pmc_profile"])
    kern = {}
    for name, (ms, n) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
        row = {"ms_per_step": round(ms / steps, 4), "launches_per_step": n / steps}
        if name in B and ms > 0:
            row["algorithmic_bytes_per_step"] = int(B[name])
            row["achieved_gbs"] = round(B[name] / (ms / steps / 1e3) / 1e9, 1)
            row["frac"] = round(row["achieved_gbs"] / HBM_PEAK_GBS, 4)
        if name in traffic:
            t_step = traffic[name] * n / steps
            row["traffic_bytes_per_step"] = int(t_step)
            row["traffic_gbs"] = round(t_step / (ms / steps / 1e3) / 1e9, 1)
        kern[name] = row
    roofline = None
    if prof:
        name, (ms, n) = max(prof.items(), key=lambda kv: kv[1][0])
        if name in B and ms > 0:
            per_launch = B[name] * steps / n
            avg_s = ms / n / 1e3
            ach = per_launch / avg_s / 1e9
            roofline = {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": int(traffic[name]) if name in traffic else None,
                        "avg_launch_ms": round(ms / n, 4), "algorithmic_bytes_per_launch": int(per_launch)}
    hc = None
    t_hc = sum(prof[k][0] for k in HC_KERNELS if k in prof) / steps / 1e3
    if t_hc > 0:
        b_alg = sum(B[k] for k in HC_KERNELS if k in B)
        hc = {"kernels": [k for k in HC_KERNELS if k in prof], "kernel_s_per_step": round(t_hc, 6),
              "algorithmic": {"bytes_per_step": int(b_alg), "achieved_gbs": round(b_alg / t_hc / 1e9, 1),
                              "frac": round(b_alg / t_hc / 1e9 / HBM_PEAK_GBS, 4)},
              "target_frac": 0.40}
        if all(k in traffic for k in HC_KERNELS if k in prof):
            b_tr = sum(traffic[k] * prof[k][1] / steps for k in HC_KERNELS if k in prof)
            hc["traffic"] = {"bytes_per_step": int(b_tr), "achieved_gbs": round(b_tr / t_hc / 1e9, 1),
                             "frac": round(b_tr / t_hc / 1e9 / HBM_PEAK_GBS, 4),
                             "source": "profiles/" + wl["pmc_profile"]}
        else:
            hc["traffic"] = None
    total_alg = sum(B.values())
    return roofline, kern, hc, U, total_alg


# ---------------------------------------------------------------- workloads

def _uneven_workload(a, api, W, torch, features=False):
    """Device buffers of the uneven workload: per tree its own masks, values and values+1
    (and, with features, each tree's batch features for the growth predictor)."""
    from voxelis_amd.admission import batch_features
    spec = W.uneven_spec(a.trees, depths=(5, 6, 7), p_dense=0.5, seed=a.seed + 2)
    t_gen = time.perf_counter()
    dev = torch.device("cuda", 0)
    from voxelis_amd.morton import apply_edits_to_dense, dense_batch_arrays
    dense_host = {}
    mk, vA, vB, edits, feats = [], [], [], [], []
    touched = mixed = partial = 0
    for d, kind, dens, sd in spec:
        if kind == "dense":
            if d not in dense_host:
                m, v = W.dense_to_batch(W.uneven_dense(d), d)
                vb = np.where(v == 65535, 1, v + 1).astype(np.uint16)
                sm = m[:, 0]
                dense_host[d] = (api.to_device(m, dev), api.to_device(v, dev), api.to_device(vb, dev),
                                 int(np.count_nonzero(sm)),
                                 int(np.count_nonzero((sm != 0) & ((sm != 0xFF) | (v != v[:, :1]).any(axis=1)))),
                                 int(np.count_nonzero((sm != 0) & (sm != 0xFF))),
                                 batch_features(m, v, d, 2) if features else None)
            m_t, vA_t, vB_t, tc, mx, pt, ft = dense_host[d]
            feats.append(ft)
            # every tree reads its own copy of the batch, as distinct trees would
            m_t, vA_t, vB_t = m_t.clone(), vA_t.clone(), vB_t.clone()
            edits.append(1 << (3 * d))
        else:
            xyz, vals = W.uneven_tree_edits(d, kind, dens, sd)
            m, v = dense_batch_arrays(d, np.uint16)
            apply_edits_to_dense(m, v, xyz, vals)
            vb = np.where(v == 0, 0, np.where(v == 65535, 1, v + 1)).astype(np.uint16)
            sm = m[:, 0]
            tc = int(np.count_nonzero(sm))
            mx = int(np.count_nonzero((sm != 0) & ((sm != 0xFF) | (v != v[:, :1]).any(axis=1))))
            pt = int(np.count_nonzero((sm != 0) & (sm != 0xFF)))
            m_t, vA_t, vB_t = api.to_device(m, dev), api.to_device(v, dev), api.to_device(vb, dev)
            edits.append(int(np.unpackbits(sm).sum()))
            feats.append(batch_features(m, v, d, 2) if features else None)
        touched += tc
        mixed += mx
        partial += pt
        mk.append(m_t)
        vA.append(vA_t)
        vB.append(vB_t)
    depths = [d for d, _, _, _ in spec]
    t_gen = time.perf_counter() - t_gen
    units = {"value_size": 2, "trees": len(spec),
             "brick_slots": sum(8 ** (d - 1) for d in depths),
             "upper_slots": sum(sum(8 ** (d - h) for h in range(2, d + 1)) for d in depths),
             "level3_slots": sum(8 ** (d - 3) for d in depths),
             "touched_bricks": touched, "mixed_bricks": mixed, "partial_bricks": partial}
    print(f"[bench] uneven: {len(spec)} trees, {sum(edits)} edits per step, depths "
          f"{ {k: depths.count(k) for k in sorted(set(depths))} } (host generation and upload {t_gen:.1f}s)",
          file=sys.stderr, flush=True)
    return {"name": "uneven", "workload": UNEVEN_WORKLOAD, "dtype": "u16", "vsize": 2, "depths": depths,
            "kinds": [k for _, k, _, _ in spec], "spec": spec, "masks": mk, "vA": vA, "vB": vB,
            "edits": np.array(edits, dtype=np.int64), "units": units,
            "features": np.stack(feats) if features else None,
            "pmc_profile": "pmc_traffic_uneven_4096_u16.json" if a.trees == 4096 and a.seed == 1 else None,
            "data": "synthetic (seeded uniform random and set_sum-style dense fills)"}


def _uneven_host_batches(wl, idx):
    """Host (masks, values A, values B) of trees idx (regenerated from the spec)."""
    from voxelis_amd import workloads as W
    out = []
    for i, (m, v, _) in zip(idx, W.uneven_batches([wl["spec"][i] for i in idx])):
        vb = np.where(v == 0, 0, np.where(v == 65535, 1, v.astype(np.int64) + 1)).astype(np.uint16)
        out.append((m, v, vb))
    return out


def _terrain_workload(a, api, W, torch):
    t_gen = time.perf_counter()
    T = a.grid * a.grid
    mA, vA, edits, _ = W.terrain_batches(grid=a.grid, seed=a.seed)
    vB = vA.copy()
    vB[vA == 1] = 2
    vB[vA == 2] = 1
    dev = torch.device("cuda", 0)
    mA_t, vA_t, vB_t = api.to_device(mA, dev), api.to_device(vA, dev), api.to_device(vB, dev)
    sm = mA[:, :, 0]
    touched = int(np.count_nonzero(sm))
    uni = int(np.count_nonzero((sm == 0xFF) & (vA == vA[:, :, :1]).all(axis=2)))
    units = {"value_size": 1, "trees": T, "brick_slots": T * 4096, "upper_slots": T * (512 + 64 + 8 + 1),
             "level3_slots": T * 64, "touched_bricks": touched, "mixed_bricks": touched - uni,
             "partial_bricks": int(np.count_nonzero((sm != 0) & (sm != 0xFF)))}
    print(f"[bench] terrain: {T} trees, {int(edits.sum())} edits per step (host generation "
          f"{time.perf_counter() - t_gen:.1f}s)", file=sys.stderr, flush=True)
    return {"name": "terrain", "workload": TERRAIN_WORKLOAD, "dtype": "u8", "vsize": 1, "depths": [5] * T,
            "kinds": ["terrain"] * T, "masks": [mA_t[i] for i in range(T)], "vA": [vA_t[i] for i in range(T)],
            "vB": [vB_t[i] for i in range(T)], "edits": edits.astype(np.int64), "units": units,
            "host": (mA, vA, vB),
            "pmc_profile": "pmc_traffic_terrain_4096x32_u8.json" if a.grid == 64 and a.seed == 1 else None,
            "data": "synthetic (seeded Perlin-dune terrain)"}


# ---------------------------------------------------------------- measurement legs

def _forest(api, it, wl, idx):
    f = api.Forest(it, [wl["depths"][i] for i in idx])
    views = [f.make_views([wl["masks"][i] for i in idx], [wl["vA"][i] for i in idx]),
             f.make_views([wl["masks"][i] for i in idx], [wl["vB"][i] for i in idx])]
    return f, views


def _interner(api, a, wl):
    it = api.VoxInterner(wl["vsize"], a.budget_mib << 20, a.load_factor)
    it.set_graphs(bool(a.graphs))
    if a.pick:
        from voxelis_amd.autotune import apply_pick
        apply_pick(it, a.pick)
    if a.brick_rounds is not None:
        it.set_tuning(brick_rounds=a.brick_rounds)
    return it


def _check_status(f, what):
    st = f.status.cpu().numpy()
    assert (st == 1).all(), f"{what}: statuses {np.unique(st)}"


def _latencies(f, views, n, k0, torch):
    """n launches, each bracketed by HIP events on the current (the library's) stream."""
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(n):
        evs[k][0].record()
        f.apply_views(views[(k0 + k) % 2])
        evs[k][1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return np.array([e0.elapsed_time(e1) for e0, e1 in evs]), wall


def _sliced(api, a, wl, idx_all, torch, step_ms):
    """The trees in S interleaved slices, one launch sequence per slice, S doubling from an
    estimate until the p99 launch latency is within a.p99_ms (at least lat_steps launches)."""
    T = len(idx_all)
    S = 1
    while S < T and step_ms / S > 0.6 * a.p99_ms:
        S *= 2
    best = None
    while S <= T:
        it = _interner(api, a, wl)
        slices = [_forest(api, it, wl, idx_all[s::S]) for s in range(S)]
        e_slice = [int(wl["edits"][idx_all[s::S]].sum()) for s in range(S)]
        for w in range(3):  # fresh trees, then steady state; the last round captures the graphs
            for f, views in slices:
                f.apply_views(views[w % 2])
        it.sync()
        rounds = max(2, -(-max(a.lat_steps, 200) // S))
        evs = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(S)]
               for _ in range(rounds)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in range(rounds):
            for s, (f, views) in enumerate(slices):
                evs[r][s][0].record()
                f.apply_views(views[(r + 3) % 2])
                evs[r][s][1].record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        it.sync()
        for f, _ in slices:
            _check_status(f, f"slices of {S}")
        lat = np.array([e0.elapsed_time(e1) for row in evs for e0, e1 in row])
        rate = sum(e_slice) * rounds / wall
        res = {"slices": S, "trees_per_launch": -(-T // S), "launches": int(rounds * S),
               "edits_per_s": rate, "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
               "edits_per_launch_mean": sum(e_slice) / S}
        print(f"[bench] sliced x{S}: {rate / 1e9:.2f} G edits/s, p50 {res['p50_ms']:.3f} ms p99 "
              f"{res['p99_ms']:.3f} ms", file=sys.stderr, flush=True)
        it.close()
        del slices
        if res["p99_ms"] <= a.p99_ms:
            best = res
            break
        S *= 2
    return best


def _subset_rate(api, a, wl, idx, steps, torch):
    it = _interner(api, a, wl)
    f, views = _forest(api, it, wl, idx)
    for w in range(max(a.warmup, 3)):
        f.apply_views(views[w % 2])
    it.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        f.apply_views(views[(max(a.warmup, 3) + k) % 2])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    it.sync()
    _check_status(f, "subset")
    it.close()
    E = int(wl["edits"][idx].sum())
    return {"trees": len(idx), "edits_per_step": E, "edits_per_s": E * steps / wall, "ms_per_step": 1e3 * wall / steps}


def _plain(api, a, wl, idx, torch):
    """One apply call per tree (the plain per-tree GPU path), alternating phases."""
    it = _interner(api, a, wl)
    fs = [_forest(api, it, wl, [i]) for i in idx]
    for w in range(3):
        for f, views in fs:
            f.apply_views(views[w % 2])
    it.sync()
    s0 = it.stats()
    rounds = 2
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(rounds):
        for f, views in fs:
            f.apply_views(views[(r + 3) % 2])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    it.sync()
    s1 = it.stats()
    for f, _ in fs:
        _check_status(f, "plain")
    it.close()
    E = int(wl["edits"][idx].sum())
    st = {k: s1[k] - s0[k] for k in s1}
    sub = dict(wl["units"])
    frac = len(idx) / len(wl["depths"])
    tu = [_tree_units(wl, i) for i in idx]
    for k in ("brick_slots", "upper_slots", "level3_slots", "touched_bricks", "mixed_bricks", "partial_bricks"):
        sub[k] = sum(u[k] for u in tu)
    sub["trees"] = len(idx)
    b = sum(algorithmic_bytes(_units({"units": sub}, st, rounds)).values())
    ach = b / (wall / rounds) / 1e9
    return {"trees": len(idx), "calls_per_step": len(idx), "edits_per_step": E, "edits_per_s": E * rounds / wall,
            "ms_per_step": 1e3 * wall / rounds, "algorithmic_gbs": round(ach, 1),
            "frac": round(ach / HBM_PEAK_GBS, 5), "sample_share": round(frac, 4)}


def _tree_units(wl, i):
    d = wl["depths"][i]
    u = {"brick_slots": 8 ** (d - 1), "upper_slots": sum(8 ** (d - h) for h in range(2, d + 1)),
         "level3_slots": 8 ** (d - 3) if d >= 3 else 0}
    m = wl["masks"][i][:, 0].cpu().numpy()
    v = wl["vA"][i].cpu().numpy().view(np.uint16 if wl["vsize"] == 2 else np.uint8)
    u["touched_bricks"] = int(np.count_nonzero(m))
    u["mixed_bricks"] = int(np.count_nonzero((m != 0) & ((m != 0xFF) | (v != v[:, :1]).any(axis=1))))
    u["partial_bricks"] = int(np.count_nonzero((m != 0) & (m != 0xFF)))
    return u


def _cpu_baseline(a, wl):
    """The oracle (one thread) on a bounded sample of the same trees and steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    T = len(wl["depths"])
    if wl["name"] == "uneven":
        idx = list(range(0, T, max(1, T // a.cpu_trees)))[: a.cpu_trees]
        host = _uneven_host_batches(wl, idx)
    else:
        idx = list(range(T))
        mA, vA, vB = wl["host"]
        host = [(mA[i], vA[i], vB[i]) for i in idx]
    it_o = O.Interner(wl["vsize"], a.budget_mib << 20, extended_depth=True)
    trees, batches = [], []
    for i, (m, va, vb) in zip(idx, host):
        d = wl["depths"][i]
        t = O.Tree(d)
        ba, bb = O.Batch(d), O.Batch(d)
        ba.load_dense(m, va)
        bb.load_dense(m, vb)
        t.apply_batch(it_o, ba)  # the first step (fresh trees) is not timed
        trees.append(t)
        batches.append((ba, bb))
    tsum, esum, nsteps, done = 0.0, 0, 0, 0
    while tsum < a.cpu_seconds:
        for k, i in enumerate(idx):
            c0 = time.perf_counter()
            trees[k].apply_batch(it_o, batches[k][1 - nsteps % 2])
            tsum += time.perf_counter() - c0
            esum += int(wl["edits"][i])
            done += 1
            if tsum >= a.cpu_seconds and nsteps > 0:
                break
        nsteps += 1
    return {"value": esum / tsum, "unit": "edits/s", "cores": 1, "kind": "port",
            "cpu_model": _cpu_model(), "host_nproc": os.cpu_count(),
            "sample": f"oracle apply_batch on {len(idx)} of the {T} trees (every {max(1, T // len(idx))}th, shared "
                      f"interner), values alternating per step as on the GPU: {done} tree applies, {esum} edits, "
                      f"{tsum:.1f} s timed on one thread"}


def _atomics(torch):
    import ctypes as C
    from voxelis_amd._lib import lib
    buf = torch.zeros(1 << 27, dtype=torch.int64, device="cuda")
    ms = C.c_float(0)
    out = {}
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for kind, name, ops in ((0, "cas64_random", 1 << 28), (1, "add32_random", 1 << 28),
                            (3, "add32_one_counter_per_wavefront", 1 << 26)):
        lib().vx_bench_atomics(stream, kind, buf.data_ptr(), 1 << 27, ops, C.byref(ms))  # warm
        buf.zero_()
        rc = lib().vx_bench_atomics(stream, kind, buf.data_ptr(), 1 << 27, ops, C.byref(ms))
        if rc == 0 and ms.value > 0:
            out[name + "_ops_per_s"] = (ops / 64 if kind == 3 else ops) / (ms.value / 1e3)
    del buf
    return out


def bench_headline(a, wl_name):
    import torch
    from voxelis_amd import api
    from voxelis_amd import workloads as W
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    admit = wl_name == "uneven" and not a.quick and a.admit_steps > 0 and os.path.exists(PREDICTOR_FILE)
    wl = (_uneven_workload(a, api, W, torch, features=admit) if wl_name == "uneven"
          else _terrain_workload(a, api, W, torch))
    T = len(wl["depths"])
    E = int(wl["edits"].sum())
    idx_all = np.arange(T)
    it = _interner(api, a, wl)
    f, views = _forest(api, it, wl, idx_all)
    for w in range(a.warmup):
        f.apply_views(views[w % 2])
    it.sync()
    _check_status(f, "warmup")
    print(f"[bench] warmup done ({a.warmup} steps)", file=sys.stderr, flush=True)

    # ---- the timed steps: whole-job throughput
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        f.apply_views(views[(a.warmup + k) % 2])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    it.sync()
    _check_status(f, "timed")
    value = E * a.steps / wall
    leaves, branches = it.counts()
    print(f"[bench] batched: {value / 1e9:.2f} G edits/s, {1e3 * wall / a.steps:.3f} ms/step; patterns "
          f"{leaves} leaves {branches} branches", file=sys.stderr, flush=True)
    k0 = a.warmup + a.steps
    lat = None
    if not a.quick:
        lat, lat_wall = _latencies(f, views, a.lat_steps, k0, torch)
        k0 += a.lat_steps
        it.sync()
        _check_status(f, "latency")
        print(f"[bench] latency over {a.lat_steps} launches: p50 {np.percentile(lat, 50):.3f} ms, p99 "
              f"{np.percentile(lat, 99):.3f} ms", file=sys.stderr, flush=True)

    # ---- per-kernel HIP events on the library's stream (kernels launched one by one)
    roofline = kern = hc = None
    U = total_alg = None
    st = {}
    if not a.no_profile:
        it.profile(True)
        for k in range(2):
            f.apply_views(views[(k0 + k) % 2])
        it.profile_read()
        s0 = it.stats()
        psteps = max(2, min(a.steps, 6))
        for k in range(psteps):
            f.apply_views(views[(k0 + 2 + k) % 2])
        prof = it.profile_read()
        it.profile(False)
        s1 = it.stats()
        k0 += 2 + psteps
        st = {k: s1[k] - s0[k] for k in s1}
        roofline, kern, hc, U, total_alg = _roofline(prof, st, psteps, wl)
        print(f"[bench] kernels: {json.dumps({k: v['ms_per_step'] for k, v in kern.items()})}", file=sys.stderr,
              flush=True)
        if hc:
            print(f"[bench] hash-consing passes: algorithmic {hc['algorithmic']['frac'] * 100:.1f} % of HBM peak, "
                  f"traffic {hc['traffic']['frac'] * 100 if hc['traffic'] else float('nan'):.1f} %",
                  file=sys.stderr, flush=True)
    table = it.table_stats()
    it.close()
    del f, views
    if a.calib_copy:
        # the reference kernel of the memory counters: exactly 1 GiB read and 1 GiB written
        import ctypes as C
        from voxelis_amd._lib import lib
        src = torch.ones(1 << 30, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        ms = C.c_float(0)
        lib().vx_bench_copy(C.c_void_p(torch.cuda.current_stream().cuda_stream), dst.data_ptr(), src.data_ptr(),
                            1 << 30, C.byref(ms))
        print(f"[bench] copy 1 GiB: {ms.value:.3f} ms = {2 * (1 << 30) / ms.value / 1e6:.0f} GB/s",
              file=sys.stderr, flush=True)
        del src, dst

    out = {
        "metric": METRIC, "value": value, "unit": "edits/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": wl["dtype"], "data": wl["data"],
        "config": {"workload": wl["workload"], "trees": T, "edits_per_step": E,
                   "interner_budget_mib": a.budget_mib, "load_factor": a.load_factor, "seed": a.seed,
                   "parallelism": "single GPU",
                   "launch": "HIP-graph replay of the captured launch sequence" if a.graphs
                   else "kernels launched one by one",
                   "tuning": dict(a.pick, source="voxelis_amd/tuning.json (autotuner pick)") if a.pick
                   else "library defaults"},
        "p50_ms": float(np.percentile(lat, 50)) if lat is not None else None,
        "p99_ms": float(np.percentile(lat, 99)) if lat is not None else None,
        "latency_launches": a.lat_steps if lat is not None else 0,
        "patterns": {"leaves": leaves, "branches": branches},
        "pattern_table": {"slots": table[0], "tombstones": table[1], "rebuilds": table[2]},
        "roofline": roofline, "kernels": kern, "hash_consing_roofline": hc,
        "all_kernels_algorithmic_gbs": round(total_alg / (wall / a.steps) / 1e9, 1) if total_alg else None,
        "counters_per_step": {k: v / max(2, min(a.steps, 6)) for k, v in st.items()} if st else None,
    }
    if wl_name == "uneven":
        d = np.array(wl["depths"])
        out["config"]["trees_by_depth"] = {int(k): int((d == k).sum()) for k in (5, 6, 7)}
        out["config"]["edits_depth_le6"] = int(wl["edits"][d <= 6].sum())
        out["config"]["edits_depth7"] = int(wl["edits"][d == 7].sum())
        out["config"]["p_dense"] = 0.5
    if a.quick:
        print(json.dumps(out))
        return
    step_ms = 1e3 * wall / a.steps
    out["atomics"] = _atomics(torch)
    sl = _sliced(api, a, wl, idx_all, torch, step_ms)
    out["p99_bounded"] = sl
    if wl_name == "uneven":
        d = np.array(wl["depths"])
        out["by_depth"] = {"depth_le6": _subset_rate(api, a, wl, idx_all[d <= 6], a.steps, torch),
                           "depth7": _subset_rate(api, a, wl, idx_all[d == 7], a.steps, torch)}
        print(f"[bench] depth<=6: {out['by_depth']['depth_le6']['edits_per_s'] / 1e9:.2f} G edits/s, depth 7: "
              f"{out['by_depth']['depth7']['edits_per_s'] / 1e9:.2f} G edits/s", file=sys.stderr, flush=True)
    plain = None
    if a.plain_trees > 0:
        n = min(a.plain_trees, T)
        plain = _plain(api, a, wl, list(range(0, T, max(1, T // n)))[:n], torch)
        print(f"[bench] plain: {plain['edits_per_s'] / 1e9:.3f} G edits/s over {plain['trees']} trees",
              file=sys.stderr, flush=True)
    out["plain"] = plain
    out["plain_edits_per_s"] = plain["edits_per_s"] if plain else None
    out["batched_over_plain"] = value / plain["edits_per_s"] if plain else None
    out["p99_bounded_over_plain"] = sl["edits_per_s"] / plain["edits_per_s"] if (plain and sl) else None
    if admit:
        # admission by memory budget with the committed growth predictor
        from voxelis_amd.predictor import GrowthMLP, Predictor
        d = np.load(PREDICTOR_FILE)
        pred = Predictor(GrowthMLP(params=d["params"]), d["mu"], d["sd"])
        adm = _admitted_steps(api, a, wl, pred, float(d["margin"]), a.admit_steps, torch)
        adm["predictor"] = "voxelis_amd/growth_batch.npz (bench.py --config predictor)"
        out["admission"] = adm
        print(f"[bench] admission: {adm['fallbacks']} fallback recounts in {adm['applies']} applies, reserved/used "
              f"{adm['reserved_over_used']:.3f}", file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu:
        cpu = _cpu_baseline(a, wl)
        print(f"[bench] cpu oracle: {cpu['value'] / 1e6:.1f} M edits/s ({cpu['sample']})", file=sys.stderr,
              flush=True)
    out["cpu_baseline"] = cpu
    print(json.dumps(out))


def main():
    a = parse()
    if a.pmc_summary:
        return pmc_summary(a)
    if a.trace_summary:
        return trace_summary(a)
    if a.gpus != 1:
        print(f"bench.py measures one GPU (this build has no multi-GPU path); got --gpus {a.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if a.config in ("uneven", "terrain"):
        return bench_headline(a, a.config)
    return {"plain": bench_plain, "gets": bench_gets, "set": bench_set, "vtm": bench_vtm,
            "serve": bench_serve, "calib": bench_calib, "tune": bench_tune,
            "predictor": bench_predictor, "edits": bench_edits}[a.config](a)


# ---------------------------------------------------------------- profiler summaries

def _kernel_short(name):
    import re
    m = re.search(r"vx::(k_\w+)", name)
    base = m.group(1) if m else name.split("(")[0].strip()
    return TRACE_NAMES.get(base, base)


def _csv_rows(d, suffix):
    import csv
    import glob
    rows = []
    for p in glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True):
        with open(p, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def _steady_from(rows, skip):
    """Dispatch id where the apply after the first `skip` applies begins (applies are
    delimited by their k_prepare dispatches), or 0 when there are not that many."""
    starts = sorted({int(r["Dispatch_Id"]) for r in rows if _kernel_short(r["Kernel_Name"]) == "k_prepare"})
    return starts[skip] if len(starts) > skip else 0


def pmc_summary(a):
    """Per-kernel HBM bytes per dispatch from two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE),
    each scaled by what the counter reads for the 1 GiB copy kernel of the same run.  The
    first --skip-applies applies (the build from empty trees) are left out."""
    res = {}
    for counter, d in (("FETCH_SIZE", a.pmc_summary[0]), ("WRITE_SIZE", a.pmc_summary[1])):
        rows = [r for r in _csv_rows(d, "counter_collection.csv") if r.get("Counter_Name") == counter]
        first = _steady_from(rows, a.skip_applies)
        per = {}
        for r in rows:
            k = _kernel_short(r["Kernel_Name"])
            disp = int(r["Dispatch_Id"])
            if k != "k_copy_bench" and disp < first:
                continue
            e = per.setdefault(k, {})
            e[disp] = e.get(disp, 0.0) + float(r["Counter_Value"])
        res[counter] = per
    copy_bytes = 1 << 30
    fetch_copy = sum(res["FETCH_SIZE"].get("k_copy_bench", {}).values())
    write_copy = sum(res["WRITE_SIZE"].get("k_copy_bench", {}).values())
    fs = copy_bytes / 1024 / fetch_copy if fetch_copy else 1.0
    ws = copy_bytes / 1024 / write_copy if write_copy else 1.0
    applies = max(len(res["FETCH_SIZE"].get("k_prepare", {})), 1)
    out = {"copy_kernel_bytes": copy_bytes, "fetch_kib_for_copy": fetch_copy, "write_kib_for_copy": write_copy,
           "fetch_scale": fs, "write_scale": ws, "applies": applies, "skipped_applies": a.skip_applies,
           "kernels": {}}
    for k in sorted(set(res["FETCH_SIZE"]) | set(res["WRITE_SIZE"])):
        if k == "k_copy_bench":
            continue
        fd, wd = res["FETCH_SIZE"].get(k, {}), res["WRITE_SIZE"].get(k, {})
        n = max(len(fd), len(wd), 1)
        fk, wk = sum(fd.values()), sum(wd.values())
        out["kernels"][k] = {"fetch_kib": fk, "write_kib": wk, "dispatches": n,
                             "hbm_bytes_scaled": (fk * fs + wk * ws) * 1024}
    s = json.dumps(out, indent=1)
    print(s)
    if a.out:
        open(a.out, "w").write(s + "\n")


def trace_summary(a):
    """Per-kernel dispatches and durations from a rocprofv3 --kernel-trace run, after the
    first --skip-applies applies."""
    rows = _csv_rows(a.trace_summary, "kernel_trace.csv")
    first = _steady_from(rows, a.skip_applies)
    per = {}
    for r in rows:
        if int(r["Dispatch_Id"]) < first:
            continue
        k = _kernel_short(r["Kernel_Name"])
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3  # us
        per.setdefault(k, []).append(dur)
    applies = max(len(per.get("k_prepare", [])), 1)
    out = {k: {"dispatches": len(v), "per_apply": round(len(v) / applies, 2), "total_us": round(sum(v), 1),
               "us_per_apply": round(sum(v) / applies, 1), "avg_us": round(sum(v) / len(v), 2),
               "median_us": round(float(np.median(v)), 2), "max_us": round(max(v), 2)}
           for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1]))}
    out = {"applies": applies, "skipped_applies": a.skip_applies, "kernels": out}
    s = json.dumps(out, indent=1)
    print(s)
    if a.out:
        open(a.out, "w").write(s + "\n")


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
        wl = _uneven_workload(a, api, W, torch)
        depths, mk, va, vb, E = wl["depths"], wl["masks"], wl["vA"], wl["vB"], int(wl["edits"].sum())
        vsize = 2

    def make_case(lf):
        it = api.VoxInterner(vsize, a.budget_mib << 20, lf)
        it.set_graphs(bool(a.graphs))
        f = api.Forest(it, depths)
        views = [f.make_views(mk, va), f.make_views(mk, vb)]
        return it, (lambda k: f.apply_views(views[k % 2])), it.close

    def log(pt):
        print(f"[tune] {json.dumps(pt)}  ({E / pt['s_per_step'] / 1e9:.2f} G edits/s)", file=sys.stderr,
              flush=True)

    best, points = autotune(make_case, steps=a.steps, warmup=a.warmup, log=log)
    out = {"metric": "autotuned apply launch shape", "workload": a.tune_workload, "edits_per_step": E,
           "graphs": bool(a.graphs), "best": best, "best_edits_per_s": E / best["s_per_step"], "points": points}
    print(json.dumps(out))
    if a.out:
        open(a.out, "w").write(json.dumps(out, indent=1) + "\n")


def bench_gets(a):
    """Batched VoxOpsRead::get over the 64x64 terrain world (4096 u8 trees of 32^3):
    --queries uniform random voxels of uniform random trees, read three ways: in arrival
    order and grouped by tree with the per-query walk from the root (k_get), and grouped
    by tree with each run's top three levels staged in LDS (k_get_grouped).  Kernel times
    are HIP events on the library's stream; the grouped kernel's roofline counts the
    query's coordinates, value and presence byte plus the run offsets as its HBM bytes
    (the DAG's nodes are L2/MALL-resident: the kernel is bound by dependent-load latency,
    which is what the LDS staging removes)."""
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
    Q = a.queries
    chunk = rng.integers(0, T, size=Q).astype(np.int32)
    xyz = rng.integers(0, 32, size=(Q, 3)).astype(np.int32)
    order = np.argsort(chunk, kind="stable")  # queries grouped by tree, as a caller batching per chunk would
    offsets = np.searchsorted(chunk[order], np.arange(T + 1)).astype(np.int64)
    res, outs = {}, {}
    for name, idx, kern in (("random", np.arange(Q), "k_get"), ("grouped_walk", order, "k_get"),
                            ("grouped_lds", order, "k_get_grouped")):
        chunk_t = torch.from_numpy(chunk[idx]).cuda()
        xyz_t = torch.from_numpy(np.ascontiguousarray(xyz[idx]).reshape(-1)).cuda()
        off_t = torch.from_numpy(offsets).cuda()
        out = torch.zeros(Q, dtype=torch.uint8, device="cuda")
        pres = torch.zeros(Q, dtype=torch.uint8, device="cuda")

        def run():
            if kern == "k_get":
                f.get_device(chunk_t, xyz_t, out, pres)
            else:
                f.get_grouped_device(off_t, xyz_t, out, pres, nq=Q)
        it.profile(True)
        for _ in range(3):
            run()
        torch.cuda.synchronize()
        it.profile_read()
        g0 = time.perf_counter()
        for _ in range(a.steps):
            run()
        torch.cuda.synchronize()
        wall = time.perf_counter() - g0
        prof = it.profile_read()
        it.profile(False)
        it.sync()
        ms, n = prof.get(kern, (0.0, 1))
        res[name] = {"kernel": kern, "gets_per_s": Q * a.steps / wall, "kernel_ms": ms / n,
                     "kernel_gets_per_s": Q / (ms / n / 1e3)}
        if name != "random":
            outs[name] = (out.clone(), pres.clone())
        print(f"[gets] {name}: {Q * a.steps / wall / 1e9:.1f} G gets/s, {kern} {ms / n:.3f} ms per launch "
              f"({Q / (ms / n / 1e3) / 1e9:.1f} G gets/s)", file=sys.stderr, flush=True)
    # the two grouped forms read the same queries: their results agree
    same = bool(torch.equal(outs["grouped_walk"][0], outs["grouped_lds"][0]) and
                torch.equal(outs["grouped_walk"][1], outs["grouped_lds"][1]))
    assert same, "grouped reads differ between the staged and the walking kernel"
    kms = res["grouped_lds"]["kernel_ms"]
    per_launch = Q * (12 + 1 + 1) + 8 * (T + 1)
    ach = per_launch / (kms / 1e3) / 1e9
    roofline = {"kernel": "k_get_grouped", "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_ms": round(kms, 4), "algorithmic_bytes_per_launch": per_launch,
                "bytes_per_get": "12 (x, y, z) + 1 (value) + 1 (present) + run offsets"}
    # the oracle on a bounded sample: 256 of the trees and their queries
    it_o = O.Interner(1, a.budget_mib << 20)
    sample = list(range(0, T, T // 256))[:256]
    trees_o = {}
    for i in sample:
        t = O.Tree(5)
        b = O.Batch(5)
        b.load_dense(mA[i], vA[i])
        t.apply_batch(it_o, b)
        trees_o[i] = t
    sel = np.isin(chunk, sample)
    qs = [(i, xyz[chunk == i]) for i in sample]
    done, tsum = 0, 0.0
    while tsum < a.cpu_seconds:
        for i, q in qs:
            c0 = time.perf_counter()
            trees_o[i].get_many(it_o, q)
            tsum += time.perf_counter() - c0
            done += len(q)
    print(json.dumps({"metric": "batched get/s", "value": res["grouped_lds"]["gets_per_s"], "unit": "gets/s",
                      "config": {"workload": f"{Q} uniform random voxel reads over the 64x64 terrain world "
                                             f"(4096 u8 trees of 32^3), grouped by tree"},
                      "results": res, "roofline": roofline,
                      "cpu_baseline": {"value": done / tsum, "unit": "gets/s", "cores": 1, "kind": "port",
                                       "cpu_model": _cpu_model(), "host_nproc": os.cpu_count(),
                                       "sample": f"oracle VoxOpsRead::get on 256 of the trees, their "
                                                 f"{int(sel.sum())} queries grouped by tree, {done} gets in "
                                                 f"{tsum:.1f} s on one thread"}}))


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


PREDICTOR_FILE = os.path.join(ROOT, "voxelis_amd", "growth_batch.npz")


def _admitted_steps(api, a, wl, pred, margin, n, torch):
    """n applies of the whole workload (values alternating per step) under
    predictor-driven admission; returns (admission report, edits/s)."""
    from voxelis_amd.admission import BatchAdmission, with_state
    it = _interner(api, a, wl)
    idx = np.arange(len(wl["depths"]))
    f, views = _forest(api, it, wl, idx)
    adm = BatchAdmission(pred, margin, wl["vsize"])
    last = np.zeros(len(idx))
    static = wl["features"]
    hard = float(np.expm1(static[:, 15].astype(np.float64)).sum())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(n):
        empty = (f.roots.cpu().numpy() == 0)
        per = adm.apply(f, views[k % 2], with_state(static, empty, last))
        last = per
    wall = time.perf_counter() - t0
    _check_status(f, "admitted")
    rep = adm.report()
    rep["hard_bound_over_used"] = hard / max(float(np.mean(adm.used)), 1.0)
    rep["edits_per_s"] = float(wl["edits"].sum()) * n / wall
    rep["note"] = ("each apply synchronizes to read its counts and statuses (the admission decision "
                   "needs them); the first apply builds the trees from empty")
    it.close()
    return rep


def bench_predictor(a):
    """The learned growth predictor for batched admission.  Logs (features, exact per-tree
    growth) pairs with dry runs (vx_apply_opts.counts) from the plain checker tree, the
    64x64 terrain grid and the 4096-tree uneven workload: random subsets of trees with
    random value phases, applied for real after their dry run so the worlds evolve, and
    reset now and then so builds from empty trees are logged too.  Trains the MLP on the
    GPU with the library's kernels, then runs the uneven workload under predictor-driven
    admission (node_limit from the prediction, the exact dry-run recount when the library
    reports VX_E_BUDGET) and saves the predictor (--out, .npz)."""
    import torch
    from voxelis_amd import api
    from voxelis_amd import workloads as W
    from voxelis_amd.admission import batch_features, fit, with_state
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(a.seed + 77)
    t_all = time.perf_counter()
    # ---- the three worlds: per tree its depth, masks, value phases (device) and batch features
    worlds = {}
    wl = _uneven_workload(a, api, W, torch, features=True)
    def shift(v):  # u16 (held as int16 bits) v -> v + 1, 65535 -> 1, 0 stays 0
        u = v.to(torch.int32) & 0xFFFF
        u = torch.where(u == 0, u, (u % 65535) + 1)
        return torch.where(u > 32767, u - 65536, u).to(v.dtype)
    v2 = [shift(v) for v in wl["vB"]]
    worlds["uneven"] = {"vsize": 2, "depths": np.array(wl["depths"]), "masks": wl["masks"],
                        "phases": [wl["vA"], wl["vB"], v2], "static": wl["features"], "budget": a.budget_mib}
    mA, vA, _, _ = W.terrain_batches(grid=a.grid, seed=a.seed)
    T = len(mA)
    m_t, v_t = api.to_device(mA, dev), api.to_device(vA, dev)
    tphases = []
    for below, surface in ((1, 2), (2, 1), (1, 3), (3, 2)):
        x = torch.where(v_t == 1, torch.full_like(v_t, below), torch.where(v_t == 2, torch.full_like(v_t, surface), v_t))
        tphases.append([x[i] for i in range(T)])
    worlds["terrain"] = {"vsize": 1, "depths": np.full(T, 5), "masks": [m_t[i] for i in range(T)],
                         "phases": tphases, "budget": 2048,
                         "static": np.stack([batch_features(mA[i], vA[i], 5, 1) for i in range(T)])}
    mp, vp = W.dense_to_batch(W.plain_checker(5), 5)
    P = 64
    mp_t, vp_t = api.to_device(mp, dev), api.to_device(vp, dev)
    pph = [[torch.where(vp_t == 1, torch.full_like(vp_t, c), torch.full_like(vp_t, d))] * P
           for c, d in ((1, 2), (2, 1), (3, 1), (1, 4))]
    worlds["plain"] = {"vsize": 1, "depths": np.full(P, 5), "masks": [mp_t] * P, "phases": pph, "budget": 256,
                       "static": np.stack([batch_features(mp, vp, 5, 1)] * P)}
    print(f"[predictor] worlds ready ({time.perf_counter() - t_all:.1f}s)", file=sys.stderr, flush=True)

    def reset(w):
        if "it" in w:
            w["it"].close()
        w["it"] = api.VoxInterner(w["vsize"], w["budget"] << 20, a.load_factor)
        n = len(w["depths"])
        w["roots"] = torch.zeros(n, dtype=torch.int64, device=dev)
        w["last"] = np.zeros(n)
        w["rounds"] = 0

    for w in worlds.values():
        reset(w)
    # ---- log pairs
    Xs, ys, gs = [], [], []
    npairs = rnd = 0
    t_log = time.perf_counter()
    names, weights = ["uneven", "terrain", "plain"], [0.6, 0.35, 0.05]
    while npairs < a.pairs:
        name = names[int(rng.choice(3, p=weights))]
        w = worlds[name]
        if w["rounds"] >= 40:
            reset(w)
        n = len(w["depths"])
        k = int(rng.integers(max(1, n // 64), n + 1))
        idx = np.sort(rng.choice(n, size=k, replace=False))
        vals = w["phases"][int(rng.integers(len(w["phases"])))]
        f = api.Forest(w["it"], w["depths"][idx])
        idx_t = torch.from_numpy(idx).to(dev)
        f.roots.copy_(w["roots"][idx_t])
        views = f.make_views([w["masks"][i] for i in idx], [vals[i] for i in idx])
        X = with_state(w["static"][idx], (f.roots.cpu().numpy() == 0), w["last"][idx])
        f.apply_views(views, dry_run=True, counts=True)
        w["it"].sync()
        nb, nl, per = f.counts_host()
        f.apply_views(views)
        w["it"].sync()
        st = f.status.cpu().numpy()
        ok = st == 1
        Xs.append(X[ok])
        ys.append(per[ok].astype(np.float64))
        gs.append(np.full(int(ok.sum()), rnd))
        w["roots"][idx_t] = f.roots
        w["last"][idx] = np.where(ok, per, w["last"][idx])
        w["rounds"] += 1
        npairs += int(ok.sum())
        rnd += 1
        if rnd % 50 == 0:
            print(f"[predictor] {rnd} applies, {npairs} pairs ({time.perf_counter() - t_log:.1f}s)",
                  file=sys.stderr, flush=True)
    for w in worlds.values():
        w["it"].close()
    t_log = time.perf_counter() - t_log
    X, y, g = np.concatenate(Xs), np.concatenate(ys), np.concatenate(gs)
    # ---- train on the GPU (vx_mlp_gradient / vx_mlp_adam) and evaluate on held-out applies
    t0 = time.perf_counter()
    pred, margin, rep = fit(X, y, g)
    t_fit = time.perf_counter() - t0
    print(f"[predictor] {len(y)} pairs from {rnd} applies ({t_log:.1f}s); trained in {t_fit:.1f}s; held-out "
          f"MAE {rep['mae_new_branches_per_apply']:.0f} new branches per apply (mean "
          f"{rep['mean_new_branches_per_apply']:.0f}), margin {margin:.3f}", file=sys.stderr, flush=True)
    if a.out:
        np.savez(a.out, params=pred.model.p.cpu().numpy(), mu=pred.mu, sd=pred.sd, margin=np.float64(margin))
    # ---- deployed: the uneven workload's steps under predictor-driven admission
    a.pick = None
    adm = _admitted_steps(api, a, wl, pred, margin, max(a.admit_steps, 4), torch)
    print(f"[predictor] admission: {adm['fallbacks']} fallbacks in {adm['applies']} applies, MAE "
          f"{adm['mae_new_branches_per_apply']:.0f} new branches per apply, reserved/used "
          f"{adm['reserved_over_used']:.3f} (hard bound {adm['hard_bound_over_used']:.1f})", file=sys.stderr,
          flush=True)
    print(json.dumps({"metric": "learned growth predictor: held-out error and admission fallback rate",
                      "config": {"workload": "pairs logged from the plain checker tree, the 64x64 terrain grid "
                                             "and the 4096-tree uneven workload (random tree subsets, value "
                                             "phases); deployed in the uneven workload's admission",
                                 "network": "FP32 MLP 16-64-64-1, trained with vx_mlp_gradient / vx_mlp_adam"},
                      "log_s": t_log, "train_s": t_fit, "training": rep, "admission": adm}))


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


def bench_edits(a):
    """Edit lists through vx_trees_apply_edits: the served stream's sphere edits (radius
    uniform 2-40 voxels, random centres, u16 ids, a quarter carving) on the 16384-tree
    depth-5 u16 world, --spheres per apply.  The device radix sort of the edits by (tree,
    Morton path), the scatter into Batch buffers and the apply are timed per kernel (HIP
    events on the library's stream)."""
    import torch
    from voxelis_amd import api, serve
    from voxelis_amd import workloads as W
    world = _serve_world(api, serve, W, (32, 16, 32), a.budget_mib << 20)
    f, n = world.forest, world.n
    rng = np.random.default_rng(a.seed + 11)
    size = np.array(world.shape) * world.side
    batches = []
    for k in range(a.warmup + a.steps):
        per_tree = {}
        for _ in range(a.spheres):
            c = rng.integers(0, size)
            r = float(rng.uniform(2, 40))
            v = 0 if rng.random() < 0.25 else int(rng.integers(1, 65536))
            for t, (loc, vals) in serve.sphere_edits(world, c, r, v).items():
                per_tree.setdefault(t, []).append((loc, vals))
        counts = np.zeros(n, dtype=np.int64)
        xyz, val = [], []
        for t in range(n):
            for loc, vals in per_tree.get(t, ()):
                xyz.append(loc)
                val.append(vals)
                counts[t] += len(vals)
        offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
        batches.append((offs, api.to_device(np.concatenate(xyz).astype(np.int32)),
                        api.to_device(np.concatenate(val).astype(np.uint16)), int(offs[-1]), len(per_tree)))
    for k in range(a.warmup):
        f.apply_edits(batches[k][0], batches[k][1], batches[k][2], sync=False)
    world.it.sync()
    world.it.profile(True)
    world.it.profile_read()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.warmup, a.warmup + a.steps):
        f.apply_edits(batches[k][0], batches[k][1], batches[k][2], sync=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = world.it.profile_read()
    world.it.profile(False)
    world.it.sync()
    E = sum(b[3] for b in batches[a.warmup:])
    kern = {k: round(v[0] / a.steps, 4) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])}
    sort_ms = sum(prof[k][0] for k in ("k_sort_hist", "k_sort_scan", "k_sort_scatter") if k in prof) / a.steps
    passes = prof.get("k_sort_scatter", (0, 0))[1] / a.steps
    res = {"edits_per_apply": E / a.steps, "trees_per_apply": float(np.mean([b[4] for b in batches[a.warmup:]])),
           "edits_per_s": E / wall, "ms_per_apply": 1e3 * wall / a.steps,
           "sort_ms_per_apply": sort_ms, "sort_passes": passes,
           "sort_keys_per_s": (E / a.steps) / (sort_ms / 1e3) if sort_ms else None,
           "sort_bytes_per_key_per_pass": 2 * (8 + 4) + 8,
           "kernels_ms_per_apply": kern,
           "note": "each apply runs with profiling events around every kernel (launched one by one)"}
    print(f"[edits] {E / a.steps / 1e6:.2f} M edits per apply over {res['trees_per_apply']:.0f} trees: "
          f"{res['edits_per_s'] / 1e6:.1f} M edits/s, {res['ms_per_apply']:.3f} ms per apply; sort {sort_ms:.3f} ms "
          f"({passes:.0f} passes, {res['sort_keys_per_s'] / 1e9:.2f} G keys/s)", file=sys.stderr, flush=True)
    print(json.dumps({"metric": "edit lists: sorted on the device and applied", "config": {
        "workload": f"{a.spheres} sphere edits per apply (radius 2-40, u16 ids, a quarter carving) on the "
                    f"16384-tree depth-5 u16 world"}, "results": res}))


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
