"""The reference's voxtree.rs unit tests (voxtree.rs:1128-2202), ported to the oracle.

Each test names the reference test it restates.  Random tests use a fixed seed
where the reference used the unseeded thread RNG (voxtree.rs:2096, 2148).
"""
import numpy as np
import pytest

import oracle_lib as O

MB = 1024 * 1024


def shadow_expect(tree, it, expect):
    """Exhaustive read-back against a dense shadow (y, z, x) array of values (0 = None)."""
    d = tree.depth
    xyz = O.all_coords(d)
    vals, pres = tree.get_many(it, xyz)
    n = 1 << d
    exp = expect[xyz[:, 1], xyz[:, 2], xyz[:, 0]] if expect.ndim == 3 else expect
    exp = np.asarray(exp, dtype=np.uint64)
    assert np.array_equal(pres.astype(bool), exp != 0)
    assert np.array_equal(np.where(pres.astype(bool), vals, 0), exp)
    return n


def grid(d):
    n = 1 << d
    y, z, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return x, y, z


def test_create():  # voxtree.rs:1128-1134
    t = O.Tree(3)
    assert t.is_empty() and t.voxels_per_axis() == 8


def test_set_and_get():  # voxtree.rs:1155-1198
    it = O.Interner(1, 2048)
    t = O.Tree(3)
    assert t.set(it, 0, 0, 0, 42) and t.get(it, 0, 0, 0) == 42
    assert t.set(it, 0, 0, 0, 24) and t.get(it, 0, 0, 0) == 24
    assert t.get(it, 1, 1, 1) is None
    assert t.set(it, 7, 7, 7, 99) and t.get(it, 7, 7, 7) == 99
    t.clear(it)
    pos = [(0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 0, 1), (1, 1, 0), (1, 1, 1)]
    for i, p in enumerate(pos):
        t.set(it, *p, i + 1)
    for i, p in enumerate(pos):
        assert t.get(it, *p) == i + 1


def test_is_empty():  # voxtree.rs:1200-1214
    it = O.Interner(1, 1024)
    t = O.Tree(3)
    assert t.is_empty()
    assert t.set(it, 0, 0, 0, 1)
    assert not t.is_empty()
    t.clear(it)
    assert t.is_empty()


def test_clear():  # voxtree.rs:1216-1243
    it = O.Interner(1, 2048)
    t = O.Tree(3)
    pos = [(0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 0, 1), (1, 1, 0), (1, 1, 1)]
    for i, p in enumerate(pos):
        t.set(it, *p, i + 1)
    t.clear(it)
    assert t.is_empty()
    for p in pos:
        assert t.get(it, *p) is None
    assert it.counts() == (0, 1)  # everything was reclaimed


def test_no_default_leaf_nodes():  # voxtree.rs:1245-1262
    it = O.Interner(1, 1024)
    t = O.Tree(3)
    assert t.set(it, 0, 0, 0, 42) and t.get(it, 0, 0, 0) == 42 and not t.is_empty()
    assert t.set(it, 0, 0, 0, 0)
    assert t.get(it, 0, 0, 0) is None and t.is_empty()


def test_dirty_flag():  # voxtree.rs:1264-1282
    it = O.Interner(1, 1024)
    t = O.Tree(3)
    assert not t.dirty
    assert t.set(it, 0, 0, 0, 1) and t.dirty
    t.dirty = False
    t.clear(it)
    assert t.dirty


def test_shared_interner_uniqueness():  # voxtree.rs:1284-1304
    it = O.Interner(1, 1024)
    t1, t2 = O.Tree(3), O.Tree(3)
    assert t1.set(it, 0, 0, 0, 42)
    assert t1.get(it, 0, 0, 0) == 42 and t2.get(it, 0, 0, 0) is None
    assert t2.set(it, 0, 0, 0, 24) and t2.get(it, 0, 0, 0) == 24
    assert t1.root != t2.root


def test_shared_interner_deduplication():  # voxtree.rs:1306-1321
    it = O.Interner(1, 1024)
    t1, t2 = O.Tree(3), O.Tree(3)
    assert t1.set(it, 0, 0, 0, 42) and t2.set(it, 0, 0, 0, 42)
    assert t1.root == t2.root


def test_set_behaviour():  # voxtree.rs:1323-1343
    it = O.Interner(1, MB)
    t = O.Tree(5)
    assert t.set(it, 0, 0, 0, 3) and t.get(it, 0, 0, 0) == 3
    assert t.set(it, 0, 0, 0, 4) and t.get(it, 0, 0, 0) == 4
    assert not t.set(it, 0, 0, 0, 4) and t.get(it, 0, 0, 0) == 4


def test_batch_double_apply():  # voxtree.rs:1345-1362
    it = O.Interner(1, MB)
    t = O.Tree(5)
    b = t.create_batch()
    b.set(0, 0, 0, 3)
    assert t.apply_batch(it, b)
    assert not t.apply_batch(it, b)


def _check_expand_shared_leaf(t, it, fill, v):
    x, y, z = grid(5)
    exp = np.full(x.shape, fill, dtype=np.uint64)
    exp[0, 0, 0] = v
    shadow_expect(t, it, exp)


def test_patterns_set_expand_shared_leaf():  # voxtree.rs:1364-1397
    it = O.Interner(1, MB)
    t = O.Tree(5)
    for value in range(1, 6):
        t.fill(it, value * 10)
        t.set(it, 0, 0, 0, value)
        _check_expand_shared_leaf(t, it, value * 10, value)
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def test_patterns_batch_expand_shared_leaf():  # voxtree.rs:1399-1434
    it = O.Interner(1, MB)
    t = O.Tree(5)
    b = t.create_batch()
    b.fill(1)
    b.set(0, 0, 0, 2)
    assert t.apply_batch(it, b)
    _check_expand_shared_leaf(t, it, 1, 2)
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1
    assert it.stats()["h3_leaks"] == 1  # Phase-0 fill leaf kept alongside the patches


def test_patterns_set_checkerboard():  # voxtree.rs:1436-1472
    it = O.Interner(1, MB)
    t = O.Tree(5)
    x, y, z = grid(5)
    for yy in range(32):
        for zz in range(32):
            for xx in range(32):
                if (xx + yy + zz) % 2 == 0:
                    assert t.set(it, xx, yy, zz, 2)
    shadow_expect(t, it, np.where((x + y + z) % 2 == 0, 2, 0))
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def test_patterns_batch_checkerboard():  # voxtree.rs:1474-1511
    it = O.Interner(1, MB)
    t = O.Tree(5)
    b = t.create_batch()
    xyz = O.all_coords(5)
    vals = np.where((xyz.sum(axis=1) % 2) == 0, 2, 1)
    b.set_many(xyz, vals)
    assert t.apply_batch(it, b)
    x, y, z = grid(5)
    shadow_expect(t, it, np.where((x + y + z) % 2 == 0, 2, 1))
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def test_patterns_set_solid_fill_one_by_one():  # voxtree.rs:1513-1544
    it = O.Interner(1, MB)
    t = O.Tree(5)
    for x, y, z in O.all_coords(5):
        assert t.set(it, int(x), int(y), int(z), 3)
    shadow_expect(t, it, np.full((32, 32, 32), 3))
    assert not t.is_empty() and t.is_leaf() and it.get_ref(t.root) == 1


def test_patterns_batch_solid_fill_one_by_one():  # voxtree.rs:1546-1580
    it = O.Interner(1, MB)
    t = O.Tree(5)
    b = t.create_batch()
    xyz = O.all_coords(5)
    b.set_many(xyz, np.full(len(xyz), 3))
    assert t.apply_batch(it, b)
    shadow_expect(t, it, np.full((32, 32, 32), 3))
    assert not t.is_empty() and t.is_leaf() and it.get_ref(t.root) == 1


def test_patterns_set_solid_fill_half_one_by_one():  # voxtree.rs:1582-1625
    it = O.Interner(1, MB)
    t = O.Tree(5)
    x, y, z = grid(5)
    xyz = O.all_coords(5)
    half = xyz[xyz[:, 1] < 16]
    for value in range(1, 6):
        for p in half:
            assert t.set(it, int(p[0]), int(p[1]), int(p[2]), value)
        shadow_expect(t, it, np.where(y < 16, value, 0))
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def test_patterns_batch_solid_fill_half_one_by_one():  # voxtree.rs:1627-1673
    it = O.Interner(1, MB)
    t = O.Tree(5)
    x, y, z = grid(5)
    xyz = O.all_coords(5)
    half = xyz[xyz[:, 1] < 16]
    for value in range(1, 6):
        b = t.create_batch()
        b.set_many(half, np.full(len(half), value))
        assert t.apply_batch(it, b)
        shadow_expect(t, it, np.where(y < 16, value, 0))
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1
    # only the current tree survives: 1 leaf + root (+ EMPTY pattern)
    assert it.counts() == (1, 2)


def test_patterns_set_solid_fill_fill_op():  # voxtree.rs:1675-1699
    it = O.Interner(1, MB)
    t = O.Tree(5)
    t.fill(it, 3)
    shadow_expect(t, it, np.full((32, 32, 32), 3))
    assert not t.is_empty() and t.is_leaf() and it.get_ref(t.root) == 1


def test_patterns_batch_solid_fill_fill_op():  # voxtree.rs:1701-1729
    it = O.Interner(1, MB)
    t = O.Tree(5)
    b = t.create_batch()
    b.fill(3)
    assert t.apply_batch(it, b)
    shadow_expect(t, it, np.full((32, 32, 32), 3))
    assert not t.is_empty() and t.is_leaf() and it.get_ref(t.root) == 1


def _sparse():
    x, y, z = grid(5)
    return np.where((x % 4 == 0) & (y % 4 == 0) & (z % 4 == 0), 3, 0)


def test_patterns_set_sparse_fill():  # voxtree.rs:1731-1768
    it = O.Interner(1, MB)
    t = O.Tree(5)
    for yy in range(0, 32, 4):
        for zz in range(0, 32, 4):
            for xx in range(0, 32, 4):
                assert t.set(it, xx, yy, zz, 3)
    shadow_expect(t, it, _sparse())
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def test_patterns_batch_sparse_fill():  # voxtree.rs:1770-1810
    it = O.Interner(1, MB)
    t = O.Tree(5)
    b = t.create_batch()
    for yy in range(0, 32, 4):
        for zz in range(0, 32, 4):
            for xx in range(0, 32, 4):
                b.set(xx, yy, zz, 3)
    assert t.apply_batch(it, b)
    shadow_expect(t, it, _sparse())
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def test_patterns_set_gradient_fill():  # voxtree.rs:1812-1851
    it = O.Interner(1, MB)
    t = O.Tree(5)
    for xx in range(32):
        v = xx % 256
        for yy in range(32):
            for zz in range(32):
                assert t.set(it, xx, yy, zz, v) == (v > 0)
    x, y, z = grid(5)
    shadow_expect(t, it, x % 256)
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def test_patterns_batch_gradient_fill():  # voxtree.rs:1853-1892
    it = O.Interner(1, MB)
    t = O.Tree(5)
    b = t.create_batch()
    xyz = O.all_coords(5)
    b.set_many(xyz, xyz[:, 0] % 256)  # x == 0 is a clear, ignored by apply (H8)
    assert t.apply_batch(it, b)
    x, y, z = grid(5)
    shadow_expect(t, it, x % 256)
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def _hollow():
    x, y, z = grid(5)
    e = (x == 0) | (x == 31) | (y == 0) | (y == 31) | (z == 0) | (z == 31)
    return np.where(e, 3, 0)


def test_patterns_set_hollow_cube():  # voxtree.rs:1894-1948
    it = O.Interner(1, MB)
    t = O.Tree(5)
    h = _hollow()
    for yy in range(32):
        for zz in range(32):
            for xx in range(32):
                if h[yy, zz, xx]:
                    assert t.set(it, xx, yy, zz, 3)
                else:
                    assert t.get(it, xx, yy, zz) is None
    shadow_expect(t, it, h)
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def test_patterns_batch_hollow_cube():  # voxtree.rs:1950-2005
    it = O.Interner(1, MB)
    t = O.Tree(5)
    b = t.create_batch()
    h = _hollow()
    xyz = O.all_coords(5)
    sel = h[xyz[:, 1], xyz[:, 2], xyz[:, 0]] != 0
    b.set_many(xyz[sel], np.full(int(sel.sum()), 3))
    assert t.apply_batch(it, b)
    shadow_expect(t, it, h)
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def _diag():
    x, y, z = grid(5)
    return np.where((x == y) & (x == z), 3, 0)


def test_patterns_set_diagonal():  # voxtree.rs:2007-2042
    it = O.Interner(1, MB)
    t = O.Tree(5)
    for i in range(32):
        assert t.set(it, i, i, i, 3) and t.get(it, i, i, i) == 3
    shadow_expect(t, it, _diag())
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def test_patterns_batch_diagonal():  # voxtree.rs:2044-2082
    it = O.Interner(1, MB)
    t = O.Tree(5)
    b = t.create_batch()
    for i in range(32):
        b.set(i, i, i, 3)
    assert t.apply_batch(it, b)
    shadow_expect(t, it, _diag())
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def _noise(seed):
    rng = np.random.default_rng(seed)
    pts = rng.integers(0, 32, size=(1000, 3)).astype(np.int32)
    vals = rng.integers(1, 256, size=1000)
    dense = np.zeros((32, 32, 32), dtype=np.uint64)  # [y, z, x]
    for (x, y, z), v in zip(pts, vals):
        dense[y, z, x] = v
    return pts, vals, dense


def test_patterns_set_random_noise():  # voxtree.rs:2084-2132
    it = O.Interner(1, MB)
    t = O.Tree(5)
    pts, vals, dense = _noise(2084)
    for (x, y, z), v in zip(pts, vals):
        t.set(it, int(x), int(y), int(z), int(v))
        assert t.get(it, int(x), int(y), int(z)) == v
    shadow_expect(t, it, dense)
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def test_patterns_batch_random_noise():  # voxtree.rs:2134-2185
    it = O.Interner(1, MB)
    t = O.Tree(5)
    pts, vals, dense = _noise(2134)
    b = t.create_batch()
    b.set_many(pts, vals)  # duplicate coordinates: last write wins (H11)
    assert t.apply_batch(it, b)
    shadow_expect(t, it, dense)
    assert not t.is_empty() and not t.is_leaf() and it.get_ref(t.root) == 1


def test_max_depth_zero():  # voxtree.rs:2187-2202
    it = O.Interner(1, MB)
    t = O.Tree(0)
    assert t.set(it, 0, 0, 0, 1) and t.get(it, 0, 0, 0) == 1
    assert not t.is_empty() and t.is_leaf() and it.get_ref(t.root) == 1


def test_child_index_range():  # voxtree.rs:1136-1153 (via the Morton digit)
    for d in range(1, 7):
        for x, y, z in O.all_coords(d)[::97]:
            p = O.encode_path(int(x), int(y), int(z))
            for depth in range(d):
                assert (p >> (3 * (d - depth - 1))) & 7 < 8


# ---- panics the reference raises on this path (mapped to error codes) ----

def test_range_panics():  # voxtree.rs:146-148, 164-166
    it = O.Interner(1, MB)
    t = O.Tree(3)
    with pytest.raises(O.OraclePanic) as e:
        t.get(it, 8, 0, 0)
    assert e.value.kind == "RANGE"
    with pytest.raises(O.OraclePanic):
        t.set(it, -1, 0, 0, 1)


def test_same_root_panics():  # voxtree.rs:311 (H6)
    it = O.Interner(1, MB)
    t = O.Tree(3)
    b = t.create_batch()
    b.fill(4)
    assert t.apply_batch(it, b)
    with pytest.raises(O.OraclePanic) as e:
        t.apply_batch(it, b)
    assert e.value.kind == "SAME_ROOT"


def test_leaf_brick_panics():  # voxtree.rs:829-834 + block_id.rs:273-276 (H1)
    it = O.Interner(1, MB)
    t = O.Tree(3)
    b = t.create_batch()
    xyz = O.all_coords(1)  # one full 2x2x2 brick at the origin, uniform
    b.set_many(xyz, np.full(8, 5))
    b.set(7, 7, 7, 1)  # keeps the root a branch
    assert t.apply_batch(it, b)
    b2 = t.create_batch()
    b2.set(0, 0, 0, 6)  # partial edit of the collapsed brick leaf
    with pytest.raises(O.OraclePanic) as e:
        t.apply_batch(it, b2)
    assert e.value.kind == "LEAF_TYPES"


def test_batch_depth_below_two_panics():  # voxtree.rs:792, 913 underflow (H9)
    it = O.Interner(1, MB)
    t = O.Tree(1)
    b = t.create_batch()
    b.set(0, 0, 0, 1)
    with pytest.raises(O.OraclePanic) as e:
        t.apply_batch(it, b)
    assert e.value.kind == "DEPTH"


def test_depth_seven_rejected_unless_extended():  # max_depth.rs:77-83
    it = O.Interner(2, MB)
    t = O.Tree(7)
    b = t.create_batch()
    b.set(0, 0, 0, 1)
    with pytest.raises(O.OraclePanic):
        t.apply_batch(it, b)
    it2 = O.Interner(2, MB, extended_depth=True)
    assert t.apply_batch(it2, b) and t.get(it2, 0, 0, 0) == 1


def test_snapshot_restore_undoes_a_panicking_apply():
    """The comparison mode for reference panics: an apply that mutates the interner and
    then hits H1 is undone completely (counts, refcounts, allocator, later behaviour)."""
    def build(it):
        t = O.Tree(4)
        b = O.Batch(4)
        g = np.stack(np.meshgrid(*[np.arange(14, 16)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.int32)
        b.set_many(g, np.full(len(g), 5, dtype=np.uint64))  # a uniform brick: a leaf at brick level
        b.set_many(np.array([[13, 15, 15], [9, 9, 9], [12, 3, 7]], dtype=np.int32),
                   np.array([3, 2, 3], dtype=np.uint64))  # its parent stays a branch
        t.apply_batch(it, b)
        return t

    def later(it, t):
        b = O.Batch(4)
        b.set_many(np.array([[11, 12, 12], [1, 14, 2]], dtype=np.int32), np.array([7, 8], dtype=np.uint64))
        t.apply_batch(it, b)

    it_a, it_b = O.Interner(1, 1 << 20), O.Interner(1, 1 << 20)
    ta, tb = build(it_a), build(it_b)
    before = (it_a.counts(), it_a.alive_nodes(), it_a.get_ref(ta.root), list(ta.to_vec(it_a)))
    bad = O.Batch(4)
    # bricks earlier in Morton order create new nodes, then the last brick is a partial
    # edit of the collapsed one
    bad.set_many(np.array([[0, 0, 0], [2, 2, 2], [15, 15, 15]], dtype=np.int32),
                 np.array([4, 6, 9], dtype=np.uint64))
    # without the snapshot the panicking call leaves mutations behind
    it_c = O.Interner(1, 1 << 20)
    tc = build(it_c)
    with pytest.raises(O.OraclePanic):
        tc.apply_batch(it_c, bad)
    assert (it_c.counts(), it_c.alive_nodes()) != before[:2]
    root0 = ta.root
    with pytest.raises(O.OraclePanic) as e:
        ta.apply_batch_atomic(it_a, bad)
    assert e.value.kind == "LEAF_TYPES"
    assert ta.root == root0
    assert (it_a.counts(), it_a.alive_nodes(), it_a.get_ref(ta.root), list(ta.to_vec(it_a))) == before
    later(it_a, ta)
    later(it_b, tb)
    assert it_a.counts() == it_b.counts() and it_a.alive_nodes() == it_b.alive_nodes()
    assert ta.root == tb.root and list(ta.to_vec(it_a)) == list(tb.to_vec(it_b))
