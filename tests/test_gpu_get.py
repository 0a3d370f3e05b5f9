"""Reads grouped by tree (vx_trees_get_grouped: each run of one tree's queries walks the
tree's top three levels once into LDS, every query continues from there) against the
oracle's VoxOpsRead::get (voxtree.rs:145-159, utils/common.rs:122-156) and against the
ungrouped kernel.  The trees cover every depth from 2 to 7, a leaf root, a leaf one
level down, empty regions and an empty tree; the query runs cover no query, runs below
the staging thresholds (32 and 256 queries), runs spanning several blocks, and voxels
outside their tree."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_helpers import oracle_apply_dense

pytestmark = pytest.mark.gpu

MB = 1024 * 1024


@pytest.fixture(scope="module")
def vx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    from voxelis_amd import api, workloads
    return torch, api, workloads


def _dense_trees(vsize, rng):
    """(depth, dense [N,N,N] (y,z,x)) per tree."""
    vmax = {1: 255, 2: 65535, 4: 0xFFFFFFFF}[vsize]
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[vsize]

    def rnd(shape, p):
        v = rng.integers(1, vmax, size=shape, endpoint=True).astype(dt)
        return np.where(rng.random(shape) < p, v, 0).astype(dt)

    trees = [(5, np.full((32, 32, 32), 7, dtype=dt))]                    # a leaf root
    d = rnd((32, 32, 32), 0.05)
    d[:16, :16, :16] = 3                                                  # a leaf at depth 1
    trees.append((5, d))
    n = 64
    y, z, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    trees.append((6, ((x + y + z) % min(vmax, 65535) + 1).astype(dt)))   # dense, no uniform region
    trees.append((7, rnd((128, 128, 128), 0.001)))                        # mostly empty regions
    trees.append((2, rnd((4, 4, 4), 0.7)))
    trees.append((2, rnd((4, 4, 4), 0.3)))  # (a depth-1 batch is the reference's panic, H9)
    trees.append((5, np.zeros((32, 32, 32), dtype=dt)))                  # an empty tree
    trees.append((3, rnd((8, 8, 8), 0.5)))
    one = np.zeros((16, 16, 16), dtype=dt)
    one[5, 9, 2] = 11
    trees.append((4, one))
    trees.append((6, rnd((64, 64, 64), 0.1)))
    return trees


def _build(api, W, vsize, seed):
    rng = np.random.default_rng(seed)
    trees = _dense_trees(vsize, rng)
    it = api.VoxInterner(vsize, 256 * MB)
    f = api.Forest(it, [d for d, _ in trees])
    it_o = O.Interner(vsize, 256 * MB, extended_depth=True)
    trees_o, ms, vs, oks = [], [], [], []
    for d, dense in trees:
        m, v = W.dense_to_batch(dense, d)
        ms.append(api.to_device(m))
        vs.append(api.to_device(v))
        t = O.Tree(d)
        ok, _ = oracle_apply_dense(it_o, t, m, v)
        oks.append(int(ok))
        trees_o.append(t)
    # a tree with nothing to set gets a batch without patches, as Batch::has_patches says
    st = f.apply(ms, vs, has_patches=[bool(dense.any()) for _, dense in trees])
    assert list(st) == oks, (st, oks)
    return rng, trees, f, it, it_o, trees_o


@pytest.mark.parametrize("vsize", [1, 2, 4])
def test_grouped_reads_match_oracle(vx, vsize):
    torch, api, W = vx
    rng, trees, f, it, it_o, trees_o = _build(api, W, vsize, 11 + vsize)
    # queries per tree: none, below both thresholds, between them, several blocks
    counts = [0, 3000, 40, 5000, 10, 300, 256, 700, 33, 1200]
    assert len(counts) == len(trees)
    xyz = []
    for (d, _), n in zip(trees, counts):
        xyz.append(rng.integers(0, 1 << d, size=(n, 3)).astype(np.int32))
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    q = np.concatenate(xyz)
    gv, gp = f.get_grouped(offsets, q)
    chunk = np.repeat(np.arange(len(trees), dtype=np.uint32), counts)
    pv, pp = f.get(chunk, q)
    assert np.array_equal(gp, pp) and np.array_equal(gv, pv)
    for i, t in enumerate(trees_o):
        a, b = int(offsets[i]), int(offsets[i + 1])
        ov, op = t.get_many(it_o, q[a:b])
        assert np.array_equal(gp[a:b], op), i
        assert np.array_equal(np.where(op == 1, gv[a:b], 0), np.where(op == 1, ov, 0)), i
    # every voxel of the small trees, one run each
    for i, (d, _) in enumerate(trees):
        if d > 4:
            continue
        allq = O.all_coords(d)
        off = np.zeros(len(trees) + 1, dtype=np.uint64)
        off[i + 1:] = len(allq)
        gv, gp = f.get_grouped(off, allq)
        ov, op = trees_o[i].get_many(it_o, allq)
        assert np.array_equal(gp, op) and np.array_equal(np.where(op == 1, gv, 0), np.where(op == 1, ov, 0)), i
    it.close()


def test_grouped_reads_out_of_range(vx):
    torch, api, W = vx
    rng, trees, f, it, it_o, trees_o = _build(api, W, 2, 5)
    counts = [100] * len(trees)
    xyz = [rng.integers(0, 1 << d, size=(n, 3)).astype(np.int32) for (d, _), n in zip(trees, counts)]
    xyz[2][7] = [64, 0, 0]      # outside the depth-6 tree
    xyz[5][0] = [0, -1, 0]     # (tree 5 has depth 2)
    q = np.concatenate(xyz)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    offsets[-1] -= 5            # the last five queries belong to no tree
    gv, gp = f.get_grouped(offsets, q)
    assert gp[200 + 7] == 0xFF and gp[500] == 0xFF
    assert (gp[-5:] == 0xFF).all()
    bad = {207, 500} | set(range(len(q) - 5, len(q)))
    ok = np.array([i not in bad for i in range(len(q))])
    assert (gp[ok] != 0xFF).all()
    chunk = np.repeat(np.arange(len(trees), dtype=np.uint32), counts)
    pv, pp = f.get(chunk[:-5], q[:-5])
    sel = ok[:-5]
    assert np.array_equal(gp[:-5][sel], pp[sel]) and np.array_equal(gv[:-5][sel], pv[sel])
    it.close()
