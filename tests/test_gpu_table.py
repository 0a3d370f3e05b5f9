"""Pattern-table churn.  Frees leave tombstones in the open-addressing table; a key
reuses one only when its scan meets it before an empty slot, so without maintenance
the table fills with tombstones and scans lengthen until no empty slot is left.  Many
applies of fresh random keys on a small interner must stay error-free with short
scans, rebuild the table by themselves, and match the oracle throughout."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_helpers import assert_same

pytestmark = pytest.mark.gpu

MB = 1024 * 1024


@pytest.fixture(scope="module")
def vx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    from voxelis_amd import api
    return api


def _step(api, f, it_o, trees_o, rng, depth):
    """Every voxel of every tree rewritten with random values from a small alphabet: all
    brick keys are new (their predecessors are freed), the leaves stay few."""
    xyz = O.all_coords(depth)
    vals = [rng.integers(1, 201, size=len(xyz)).astype(np.uint16) for _ in trees_o]
    offs = np.arange(len(trees_o) + 1, dtype=np.uint64) * len(xyz)
    st = f.apply_edits(offs, api.to_device(np.tile(xyz, (len(trees_o), 1))), api.to_device(np.concatenate(vals)))
    assert (st == 1).all(), st
    for t, v in zip(trees_o, vals):
        b = O.Batch(depth)
        b.set_many(xyz, v.astype(np.uint64))
        assert t.apply_batch(it_o, b)


def test_table_churn_rebuilds(vx):
    api = vx
    depth, n_trees = 4, 8
    budget = 2 * MB  # u16: 26214 nodes, a 65536-slot table at load factor 0.5
    it = api.VoxInterner(2, budget)
    it_o = O.Interner(2, budget)
    f = api.Forest(it, [depth] * n_trees)
    trees_o = [O.Tree(depth) for _ in range(n_trees)]
    rng = np.random.default_rng(7)
    s0 = it.stats()
    for _ in range(200):
        _step(api, f, it_o, trees_o, rng, depth)
    s1 = it.stats()
    slots, tomb, rebuilds = it.table_stats()
    assert slots == 65536
    leaders = (s1["branch_hits"] - s0["branch_hits"]) + (s1["branch_misses"] - s0["branch_misses"])
    probes = s1["probes"] - s0["probes"]
    assert probes < 16 * leaders, (probes, leaders)  # scans stay short
    assert rebuilds >= 1 or tomb < slots // 16, (tomb, rebuilds)
    assert_same(it, f, it_o, trees_o, full_reads=False, what="after churn")
    # a forced rebuild leaves no tombstone and changes nothing observable
    assert it.rebuild_table() == rebuilds + 1
    assert it.table_stats()[1] == 0
    assert_same(it, f, it_o, trees_o, full_reads=True, what="after rebuild")
    for _ in range(20):
        _step(api, f, it_o, trees_o, rng, depth)
    assert_same(it, f, it_o, trees_o, full_reads=True, what="after rebuild and more churn")


def test_table_rebuilds_when_applies_are_queued(vx):
    """Applies queued back to back with no synchronization (the host far ahead of the
    device, as a benchmark's latency loop runs): the rebuild decision is made on the
    device in stream order, so the table keeps empty slots and scans stay short however
    many applies are queued, and the trees still match the oracle."""
    api = vx
    from voxelis_amd import workloads as W
    from gpu_helpers import oracle_apply_dense
    depth, n_trees, K, N = 4, 8, 6, 300
    budget = 2 * MB
    it = api.VoxInterner(2, budget)
    it.set_graphs(True)
    f = api.Forest(it, [depth] * n_trees)
    rng = np.random.default_rng(9)
    side = 1 << depth
    host, dev, views = [], [], []
    for k in range(K):
        dense = rng.integers(1, 201, size=(n_trees, side, side, side)).astype(np.uint16)
        mv = [W.dense_to_batch(dense[t], depth) for t in range(n_trees)]
        host.append(mv)
        # the views hold raw device pointers: their tensors stay referenced
        dev.append(([api.to_device(m) for m, _ in mv], [api.to_device(v) for _, v in mv]))
        views.append(f.make_views(*dev[-1]))
    s0 = it.stats()
    for j in range(N):
        f.apply_views(views[j % K])
    it.sync()
    assert (f.status.cpu().numpy() == 1).all()
    s1 = it.stats()
    slots, tomb, rebuilds = it.table_stats()
    leaders = (s1["branch_hits"] - s0["branch_hits"]) + (s1["branch_misses"] - s0["branch_misses"])
    probes = s1["probes"] - s0["probes"]
    assert probes < 16 * leaders, (probes, leaders)
    assert rebuilds >= 1 and tomb < slots, (tomb, rebuilds)
    it_o = O.Interner(2, budget)
    trees_o = [O.Tree(depth) for _ in range(n_trees)]
    for j in range(N):
        for t, (m, v) in zip(trees_o, host[j % K]):
            ok, _ = oracle_apply_dense(it_o, t, m, v)
            assert ok
    assert_same(it, f, it_o, trees_o, full_reads=True, what=f"after {N} queued applies")
    it.close()
