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
