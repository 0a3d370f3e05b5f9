"""Closed-form known answers that pin the oracle's apply_batch beyond the ported tests.

Config 1 (one 32^3 chunk, u8, v = ((x+y+z)&1)+1): every brick carries the same
pattern [1,2,2,1,2,1,1,2], so the DAG is a chain of one node per level.  The
reference's sequential allocation order fixes the node ids exactly.

Dense set_sum (v = x+y+z+1): a node at height h is determined by the sum of its
base corner coordinates, so it has 3*2^(D-h)-2 distinct values per height:
3*2^D-2 leaves and 3(2^D-1)-2D branches.
"""
import numpy as np

import oracle_lib as O

MB = 1024 * 1024


def config1(it):
    t = O.Tree(5)
    b = t.create_batch()
    xyz = O.all_coords(5)
    b.set_many(xyz, ((xyz.sum(axis=1)) & 1) + 1)
    assert t.apply_batch(it, b)
    return t


def test_config1_counts_refs_ids():
    it = O.Interner(1, 256 * MB)
    t = config1(it)
    assert it.counts() == (2, 6)  # 2 leaves; 5 branches + the EMPTY pattern
    assert t.root == 0x007F800000000007  # new_branch(7, 0, types=0, mask=0xFF)
    # walk the chain: every level is one node with 8 identical children
    node, level_ids = t.root, []
    for _ in range(4):
        ch = it.children(node)
        assert len(set(ch)) == 1
        node = ch[0]
        level_ids.append(node)
    brick = level_ids[-1]
    assert brick == 0x7FFF800000000003  # types=mask=0xFF, index 3
    leaves = it.children(brick)
    assert [it.get_value(c) for c in leaves] == [1, 2, 2, 1, 2, 1, 1, 2]
    assert it.get_ref(t.root) == 1
    for nid in level_ids:
        assert it.get_ref(nid) == 8
    assert sorted({it.get_ref(c) for c in leaves}) == [4]
    assert {c & 0xFFFFFFFF for c in leaves} == {1, 2}


def test_config1_reads():
    it = O.Interner(1, 256 * MB)
    t = config1(it)
    xyz = O.all_coords(5)
    vals, pres = t.get_many(it, xyz)
    assert pres.all()
    assert np.array_equal(vals, ((xyz.sum(axis=1)) & 1) + 1)


def set_sum_tree(it, d, offset=1):
    t = O.Tree(d)
    b = t.create_batch()
    xyz = O.all_coords(d)
    b.set_many(xyz, xyz.sum(axis=1) + offset)
    assert t.apply_batch(it, b)
    return t


def test_set_sum_closed_form_depths():
    for d in (2, 3, 4, 5):
        it = O.Interner(2, 64 * MB)
        set_sum_tree(it, d)
        leaves, branches = it.counts()
        assert leaves == 3 * 2 ** d - 2, d
        assert branches - 1 == 3 * (2 ** d - 1) - 2 * d, d


def test_set_sum_union_is_deepest():
    it = O.Interner(2, 64 * MB)
    for d in (3, 5, 4, 2, 5):
        set_sum_tree(it, d)
    leaves, branches = it.counts()
    assert leaves == 94 and branches - 1 == 83


def test_alternating_batches_reclaim():
    # the reference bench method (voxtree_bench.rs:553-579): two prebuilt batches
    # applied alternately to one tree; the old tree is reclaimed each time
    it = O.Interner(4, 64 * MB)
    t = O.Tree(4)
    xyz = O.all_coords(4)
    b0, b1 = t.create_batch(), t.create_batch()
    b0.set_many(xyz, xyz.sum(axis=1) + 1)
    b1.set_many(xyz, xyz.sum(axis=1) + 1000)
    for i in range(6):
        assert t.apply_batch(it, b0 if i % 2 == 0 else b1)
        leaves, branches = it.counts()
        assert leaves == 3 * 16 - 2 and branches - 1 == 3 * 15 - 8
        assert it.get_ref(t.root) == 1
    st = it.stats()
    assert st["transient_free"] == 0 and st["h4_undercount"] == 0 and st["h5_collapse_cloned"] == 0


def test_canonical_hash_is_id_independent():
    # the same content reached through different allocation orders hashes equal
    it1 = O.Interner(2, 64 * MB)
    it2 = O.Interner(2, 64 * MB)
    for d in (2, 3):  # pre-populate it2 so every id differs
        set_sum_tree(it2, d, offset=500)
    t1 = set_sum_tree(it1, 4)
    t2 = set_sum_tree(it2, 4)
    assert t1.root != t2.root
    assert t1.canonical_hash(it1) == t2.canonical_hash(it2)
    # one voxel changed -> different hash
    b = t2.create_batch()
    b.set(3, 2, 1, 999)
    assert t2.apply_batch(it2, b)
    assert t1.canonical_hash(it1) != t2.canonical_hash(it2)
    # batch and single-set paths that build the same canonical DAG agree
    it3 = O.Interner(2, 64 * MB)
    t3 = O.Tree(3)
    for x, y, z in O.all_coords(3):
        t3.set(it3, int(x), int(y), int(z), int(x + y + z + 1))
    t4 = set_sum_tree(it1, 3)
    assert t3.canonical_hash(it3) == t4.canonical_hash(it1)
