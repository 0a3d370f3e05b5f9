"""Shared helpers of the GPU parity tests: run the same batches through the oracle
and through the HIP path, then compare reads, pattern counts, canonical hashes."""
import numpy as np

import oracle_lib as O


def oracle_apply_dense(it_o, tree_o, masks, values, fill=None):
    """Feed dense (masks, values) to the oracle through its own Batch::just_set."""
    b = tree_o.create_batch()
    if fill is not None:
        b.just_fill(fill)
    set_mask = masks[:, 0]
    p, slot = np.nonzero(np.unpackbits(set_mask[:, None], axis=1, bitorder="little"))
    if len(p):
        full = (p.astype(np.int64) << 3) | slot
        xyz = _decode(full)
        b.set_many(xyz, values[p, slot].astype(np.uint64))
    return tree_o.apply_batch(it_o, b), b


def _decode(full):
    x = np.zeros(len(full), dtype=np.int64)
    y = np.zeros(len(full), dtype=np.int64)
    z = np.zeros(len(full), dtype=np.int64)
    for b in range(10):
        x |= ((full >> (3 * b)) & 1) << b
        y |= ((full >> (3 * b + 1)) & 1) << b
        z |= ((full >> (3 * b + 2)) & 1) << b
    return np.stack([x, y, z], axis=1).astype(np.int32)


def full_reads_gpu(forest, tree_index, depth):
    xyz = O.all_coords(depth)
    chunk = np.full(len(xyz), tree_index, dtype=np.uint32)
    vals, pres = forest.get(chunk, xyz)
    return np.where(pres == 1, vals, 0)


def full_reads_oracle(it_o, tree_o):
    xyz = O.all_coords(tree_o.depth)
    vals, pres = tree_o.get_many(it_o, xyz)
    return np.where(pres == 1, vals, 0)


def assert_same(it_g, forest, it_o, trees_o, full_reads=True, what=""):
    g_hash = forest.canonical_hashes()
    roots = forest.roots_host()
    for i, t in enumerate(trees_o):
        o_hash = t.canonical_hash(it_o)
        assert g_hash[i] == o_hash, f"{what} tree {i}: canonical hash differs"
        assert (roots[i] == 0) == (t.root == 0), f"{what} tree {i}: emptiness differs"
        assert (roots[i] >> 63) == (t.root >> 63), f"{what} tree {i}: leaf-root differs"
        if full_reads:
            g = full_reads_gpu(forest, i, t.depth)
            o = full_reads_oracle(it_o, t)
            assert np.array_equal(g, o), f"{what} tree {i}: reads differ"
    assert it_g.counts() == it_o.counts(), f"{what}: pattern counts differ"
