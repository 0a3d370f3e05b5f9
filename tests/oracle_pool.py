"""Oracle runs in worker processes (test infrastructure only).

A forest too large for one oracle thread within a test's time is split over worker
processes: each applies its own subset of the trees in its own oracle interner and
returns per-tree canonical hashes, sampled reads and its live patterns (leaf values,
branch canonical hashes).  Trees are independent and a node is live in the shared
interner exactly when some tree reaches it, so the union of the workers' patterns is the
pattern set of one shared interner (the configurations here have no reference-count
drift: no fills, no split leaves).
"""
import multiprocessing as mp

import numpy as np

import oracle_lib as O


def next_phase(v):
    """The other phase of the alternating benchmark steps: v -> v + 1 (65535 -> 1), 0 kept."""
    v = np.asarray(v)
    return np.where(v == 0, 0, np.where(v == 65535, 1, v.astype(np.int64) + 1)).astype(np.uint16)


def uneven_tree_batch(spec_row, dense_cache=None):
    """(masks, values) of one tree of the uneven workload (host, phase A)."""
    from voxelis_amd import workloads as W
    from voxelis_amd.morton import apply_edits_to_dense, dense_batch_arrays
    d, kind, dens, sd = spec_row
    if kind == "dense":
        if dense_cache is not None and d in dense_cache:
            return dense_cache[d]
        out = W.dense_to_batch(W.uneven_dense(d), d)
        if dense_cache is not None:
            dense_cache[d] = out
        return out
    xyz, vals = W.uneven_tree_edits(d, kind, dens, sd)
    m, v = dense_batch_arrays(d, np.uint16)
    apply_edits_to_dense(m, v, xyz, vals)
    return m, v


def read_coords(depth, n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 1 << depth, size=(n, 3)).astype(np.int32)


def _uneven_worker(args):
    spec, idx, budget, phases, n_reads, read_seed = args
    it = O.Interner(2, budget, extended_depth=True)
    cache = {}
    hashes, reads, status = [], [], []
    for i in idx:
        d = spec[i][0]
        m, v = uneven_tree_batch(spec[i], cache)
        t = O.Tree(d)
        st = []
        for ph in range(phases):
            b = O.Batch(d)
            b.load_dense(m, v if ph % 2 == 0 else next_phase(v))
            st.append(bool(t.apply_batch(it, b)))
        status.append(st)
        hashes.append(t.canonical_hash(it))
        vals, pres = t.get_many(it, read_coords(d, n_reads, read_seed + i))
        reads.append(np.where(pres == 1, vals, 0).astype(np.uint64))
    lv, bc = it.live_patterns()
    return list(idx), hashes, reads, status, lv, bc


def run_uneven(spec, workers=8, budget=4 << 30, phases=2, n_reads=64, read_seed=1000):
    """Oracle results of the uneven workload's trees over `workers` processes:
    (hashes[T], reads[T], status[T], leaf count, branch count incl. EMPTY)."""
    T = len(spec)
    jobs = [(spec, list(range(w, T, workers)), budget, phases, n_reads, read_seed) for w in range(workers)]
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        parts = pool.map(_uneven_worker, jobs)
    hashes, reads, status = [None] * T, [None] * T, [None] * T
    leaves, branches = set(), set()
    for idx, hs, rs, st, lv, bc in parts:
        for k, i in enumerate(idx):
            hashes[i], reads[i], status[i] = hs[k], rs[k], st[k]
        leaves.update(int(x) for x in lv)
        branches.update((int(a), int(b)) for a, b in bc)
    return hashes, reads, status, len(leaves), len(branches) + 1
