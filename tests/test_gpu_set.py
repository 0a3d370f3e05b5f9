"""Single-voxel set / remove on the GPU (vx_trees_set) against the oracle's
VoxOpsWrite::set (set_at_depth_iterative / remove_at_depth, voxtree.rs:162-705).

Bar: per call, every tree's changed flag (or panic) equals the oracle's; after
every few rounds the canonical DAG hash of every tree, the pattern counts and
(for small trees) every voxel read are equal; the root reference counts are equal;
clearing every tree at the end leaves the oracle's pattern counts (so reference
counts never drifted).  The value set is small so that bricks fill up and collapse,
leaves split and removes empty whole branches.
"""
import numpy as np
import pytest

import oracle_lib as O
from gpu_helpers import assert_same, oracle_apply_dense

pytestmark = pytest.mark.gpu

MB = 1024 * 1024


@pytest.fixture(scope="module")
def vx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    from voxelis_amd import api, workloads
    return torch, api, workloads


def _round(torch, api, f, it_o, trees_o, xyz, vals, vsize):
    """One vx_trees_set call and the same sets in the oracle; returns the status."""
    dt = np.uint8 if vsize == 1 else np.uint16
    st = f.set_voxels(torch.from_numpy(xyz.astype(np.int32).reshape(-1)).cuda(),
                      api.to_device(vals.astype(dt)))
    for t, t_o in enumerate(trees_o):
        x, y, z = (int(c) for c in xyz[t])
        try:
            ch = t_o.set(it_o, x, y, z, int(vals[t]))
            assert st[t] == (1 if ch else 0), (t, st[t], ch, (x, y, z), int(vals[t]))
        except O.OraclePanic as e:
            assert (e.kind, st[t]) == ("RANGE", -4), (t, e.kind, st[t])
    return st


@pytest.mark.parametrize("vsize", [1, 2])
def test_random_sets_and_removes_match_oracle(vx, vsize):
    torch, api, W = vx
    rng = np.random.default_rng(7 + vsize)
    depths = [1, 2, 3, 4, 5, 3, 2, 6] * 3
    n = len(depths)
    it = api.VoxInterner(vsize, 128 * MB)
    it_o = O.Interner(vsize, 128 * MB)
    f = api.Forest(it, depths)
    trees_o = [O.Tree(d) for d in depths]
    # starting content: empty, a whole-tree fill (a leaf root to split), or a dense batch
    fills = np.array([2 if t % 3 == 1 else 0 for t in range(n)], dtype=np.uint64)
    f.fill(fills)
    for t, t_o in enumerate(trees_o):
        if fills[t]:
            t_o.fill(it_o, int(fills[t]))
    dense_trees = [t for t in range(n) if t % 3 == 2 and depths[t] >= 2]
    if dense_trees:
        sub = api.Forest(it, [depths[t] for t in dense_trees])
        sub.roots.copy_(f.roots[dense_trees])
        bm, bv = [], []
        for t in dense_trees:
            s = 1 << depths[t]
            dense = rng.integers(0, 3, size=(s, s, s)).astype(np.uint8 if vsize == 1 else np.uint16)
            dense[: s // 2] = 1  # a uniform lower half: collapsed leaves above the bricks
            m, v = W.dense_to_batch(dense, depths[t])
            bm.append(api.to_device(m))
            bv.append(api.to_device(v))
            oracle_apply_dense(it_o, trees_o[t], m, v)
        sub.apply(bm, bv)
        f.roots[dense_trees] = sub.roots
    assert_same(it, f, it_o, trees_o, full_reads=False, what="start")
    for rnd in range(40):
        xyz = np.zeros((n, 3), dtype=np.int64)
        vals = np.zeros(n, dtype=np.int64)
        for t, d in enumerate(depths):
            s = 1 << d
            xyz[t] = rng.integers(0, s, 3)
            r = rng.random()
            vals[t] = 0 if r < 0.3 else int(rng.integers(1, 4))
        if rnd == 5:
            xyz[3] = [1 << depths[3], 0, 0]  # out of range: the reference's assert (RANGE)
        _round(torch, api, f, it_o, trees_o, xyz, vals, vsize)
        if rnd % 8 == 7 or rnd == 39:
            assert_same(it, f, it_o, trees_o, full_reads=(rnd == 39), what=f"round {rnd}")
            roots = f.roots_host()
            for t, t_o in enumerate(trees_o):
                if t_o.root != 0:
                    assert it.get_ref(int(roots[t])) == it_o.get_ref(t_o.root), (rnd, t)
    assert it_o.stats()["transient_free"] == 0
    # clearing everything leaves the oracle's counts: no reference count drifted
    f.fill(np.zeros(n, dtype=np.uint64))
    for t_o in trees_o:
        t_o.fill(it_o, 0)
    assert it.counts() == it_o.counts() == (0, 1)


def test_equal_paths_share_nodes(vx):
    """Two empty trees setting the same voxel in one call get the same root (hash-consed
    across trees), with two references; a third, different voxel gets another root."""
    torch, api, W = vx
    it = api.VoxInterner(1, 64 * MB)
    it_o = O.Interner(1, 64 * MB)
    f = api.Forest(it, [4, 4, 4])
    trees_o = [O.Tree(4) for _ in range(3)]
    xyz = np.array([[3, 5, 7], [3, 5, 7], [3, 5, 6]])
    vals = np.array([9, 9, 9])
    st = _round(torch, api, f, it_o, trees_o, xyz, vals, 1)
    assert (st == 1).all()
    r = f.roots_host()
    assert r[0] == r[1] and r[0] != r[2]
    assert it.get_ref(int(r[0])) == 2 == it_o.get_ref(trees_o[0].root)
    assert_same(it, f, it_o, trees_o, what="shared")


def test_collapse_and_remove_to_empty(vx):
    """Setting all 8 voxels of a brick to one value collapses it into the leaf (and the
    parents above, one by one, when the whole tree is filled voxel by voxel); removing
    every voxel again empties the tree."""
    torch, api, W = vx
    it = api.VoxInterner(1, 64 * MB)
    it_o = O.Interner(1, 64 * MB)
    f = api.Forest(it, [2])
    t_o = [O.Tree(2)]
    coords = O.all_coords(2)
    for c in coords:
        _round(torch, api, f, it_o, t_o, c[None, :], np.array([5]), 1)
    assert_same(it, f, it_o, t_o, what="filled")
    assert f.roots_host()[0] >> 63 == 1  # the whole tree collapsed into one leaf
    for c in coords[::-1]:
        _round(torch, api, f, it_o, t_o, c[None, :], np.array([0]), 1)
        assert_same(it, f, it_o, t_o, full_reads=False, what="removing")
    assert f.roots_host()[0] == 0 and t_o[0].root == 0
    assert it.counts() == it_o.counts()


def test_set_beyond_capacity_rolls_back(vx):
    """A set call that needs more nodes than the interner holds is rolled back as a whole
    (VX_E_OOM on every tree, nothing changed); a smaller call then matches the oracle."""
    torch, api, W = vx
    n = 12
    xyz = np.stack([np.arange(n) * 2, np.arange(n), 31 - np.arange(n)], axis=1)
    vals = np.full(n, 7)
    # the nodes the call creates (shared bricks and paths counted once), from the oracle
    probe = O.Interner(1, 64 * MB)
    for t in range(n):
        O.Tree(5).set(probe, *(int(c) for c in xyz[t]), 7)
    leaves, branches = probe.counts()
    cap = 1 + leaves + (branches - 1) - 1  # the EMPTY branch, and one node short
    it = api.VoxInterner(1, cap * (78 + 1))
    assert it.capacity() == cap
    it_o = O.Interner(1, cap * (78 + 1))
    f = api.Forest(it, [5] * n)
    trees_o = [O.Tree(5) for _ in range(n)]
    before = (it.counts(), it.alive_nodes(), f.roots_host())
    st = f.set_voxels(torch.from_numpy(xyz.astype(np.int32).reshape(-1)).cuda(), api.to_device(vals.astype(np.uint8)))
    assert (st == -2).all(), st
    assert (it.counts(), it.alive_nodes(), f.roots_host()) == before
    # the first half fits
    h = n // 2
    fh = api.Forest(it, [5] * h)
    _round(torch, api, fh, it_o, trees_o[:h], xyz[:h], vals[:h], 1)
    assert_same(it, fh, it_o, trees_o[:h], what="after rollback")
