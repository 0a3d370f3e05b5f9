"""Pins the oracle's to_vec (utils/common.rs:158-246): at full resolution it equals a
get() of every voxel (a separate traversal), and at LOD l a voxel whose cube is one
value everywhere keeps that value (a uniform cube is a single leaf or its LOD is that
value)."""
import numpy as np

import oracle_lib as O


def _tree(depth, seed):
    rng = np.random.default_rng(seed)
    it = O.Interner(2, 64 << 20)
    t = O.Tree(depth)
    b = O.Batch(depth)
    n = 1 << depth
    k = n ** 3 // 3
    xyz = rng.integers(0, n, (k, 3)).astype(np.int32)
    b.set_many(xyz, rng.integers(1, 4, k).astype(np.uint64))
    # a uniform 4^3 cube at the origin and an 8^3 one at the far corner
    c = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.int32)
    b.set_many(c, np.full(len(c), 9, dtype=np.uint64))
    c8 = (np.stack(np.meshgrid(*[np.arange(8)] * 3, indexing="ij"), -1).reshape(-1, 3) + n - 8).astype(np.int32)
    b.set_many(c8, np.full(len(c8), 11, dtype=np.uint64))
    t.apply_batch(it, b)
    return it, t


def test_to_vec_equals_reads():
    for d in (3, 5):
        it, t = _tree(d, d)
        xyz = O.all_coords(d)  # y, z, x order
        vals, pres = t.get_many(it, xyz)
        assert np.array_equal(t.to_vec(it), np.where(pres == 1, vals, 0))


def test_to_vec_lod_keeps_uniform_cubes():
    it, t = _tree(5, 7)
    full = t.to_vec(it).reshape(32, 32, 32)
    for lod in (1, 2):
        s = 1 << lod
        red = t.to_vec(it, lod).reshape(32 // s, 32 // s, 32 // s)
        cubes = full.reshape(32 // s, s, 32 // s, s, 32 // s, s)
        uniform = (cubes == cubes[:, :1, :, :1, :, :1]).all(axis=(1, 3, 5))
        first = cubes[:, 0, :, 0, :, 0]
        assert uniform.any()
        assert np.array_equal(red[uniform], first[uniform])
    assert t.to_vec(it, 2).reshape(8, 8, 8)[0, 0, 0] == 9
    assert t.to_vec(it, 3).reshape(4, 4, 4)[3, 3, 3] == 11
