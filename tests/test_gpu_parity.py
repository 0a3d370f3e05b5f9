"""Parity of the HIP path (through the C ABI) against the CPU oracle.

Bar: bit-exact reads, equal leaf/branch pattern counts, equal canonical DAG hash
per tree, equal reference counts where the reference's tests assert them.
"""
import numpy as np
import pytest

import oracle_lib as O
from gpu_helpers import assert_same, full_reads_gpu, full_reads_oracle, oracle_apply_dense

pytestmark = pytest.mark.gpu

MB = 1024 * 1024


@pytest.fixture(scope="module")
def vx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    import voxelis_amd
    from voxelis_amd import api, workloads
    return api, workloads


def test_config1_plain(vx):
    api, W = vx
    it = api.VoxInterner(1, 256 * MB)
    dense = W.plain_checker(5)
    masks, values = W.dense_to_batch(dense, 5)
    import torch
    f = api.Forest(it, [5])
    m = api.to_device(masks)
    v = api.to_device(values)
    st = f.apply([m], [v])
    assert st[0] == 1
    assert it.counts() == (2, 6)
    root = f.roots_host()[0]
    assert it.get_ref(root) == 1
    assert (root >> 47) & 0xFFFF == 0x00FF  # types 0x00, mask 0xFF (Appendix C root shape)
    reads = full_reads_gpu(f, 0, 5)
    assert np.array_equal(reads, O.all_coords(5).sum(axis=1) % 2 + 1)
    it_o = O.Interner(1, 256 * MB)
    t_o = O.Tree(5)
    oracle_apply_dense(it_o, t_o, masks, values)
    assert_same(it, f, it_o, [t_o], what="config1")
    # every non-root node of the chain holds 8 refs, the two leaves 4 each
    node = root
    for _ in range(4):
        ch = it.get_children(node)
        assert len(set(ch)) == 1
        node = ch[0]
        assert it.get_ref(node) == 8
    leaves = it.get_children(node)
    assert sorted(it.get_value(c) for c in leaves) == [1, 1, 1, 1, 2, 2, 2, 2]
    assert {it.get_ref(c) for c in leaves} == {4}


def _run_tree_case(api, depth, steps, value_size=1):
    """steps: list of ('batch', fill, xyz, values) | ('fill', v); same on both sides."""
    it = api.VoxInterner(value_size, MB)
    it_o = O.Interner(value_size, MB)
    t = api.VoxTree(depth)
    t_o = O.Tree(depth)
    for st in steps:
        if st[0] == "batch":
            _, fill, xyz, vals = st
            b = t.create_batch(it)
            b_o = t_o.create_batch()
            if fill is not None:
                b.just_fill(fill)
                b_o.just_fill(fill)
            if len(xyz):
                b.set_many(xyz, vals)
                b_o.set_many(xyz, np.asarray(vals, dtype=np.uint64))
            err_o = err_g = None
            try:
                ch_o = t_o.apply_batch(it_o, b_o)
            except O.OraclePanic as e:
                err_o = e.kind
            try:
                ch_g = t.apply_batch(it, b)
            except api.VxError as e:
                err_g = e.kind
            if err_o is not None:
                return it, t, it_o, t_o, (err_o, err_g)
            assert err_g is None and ch_o == ch_g
        else:
            t.fill(it, st[1])
            t_o.fill(it_o, st[1])
    return it, t, it_o, t_o, None


def _check_tree(api, it, t, it_o, t_o, refs=True):
    f = api.Forest(it, [t.max_depth])
    f.set_roots([t.root])
    assert_same(it, f, it_o, [t_o])
    if refs and t.root != 0:
        assert it.get_ref(t.root) == it_o.get_ref(t_o.root)


def _grid(d):
    n = 1 << d
    y, z, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


def _xyz(d, sel=None):
    xyz = O.all_coords(d)
    return xyz if sel is None else xyz[sel]


REFERENCE_BATCH_CASES = {
    # voxtree.rs test name -> (depth, steps builder)
    "batch_double_apply": lambda: (5, [("batch", None, np.array([[0, 0, 0]]), [3])] * 2),
    "batch_expand_shared_leaf": lambda: (5, [("batch", 1, np.array([[0, 0, 0]]), [2])]),
    "batch_checkerboard": lambda: (5, [("batch", None, _xyz(5), np.where(_xyz(5).sum(1) % 2 == 0, 2, 1))]),
    "batch_solid_fill_one_by_one": lambda: (5, [("batch", None, _xyz(5), np.full(32768, 3))]),
    "batch_solid_fill_half": lambda: (5, [("batch", None, _xyz(5, _xyz(5)[:, 1] < 16),
                                            np.full(16384, v)) for v in range(1, 6)]),
    "batch_solid_fill_fill_op": lambda: (5, [("batch", 3, np.zeros((0, 3)), [])]),
    "batch_sparse_fill": lambda: (5, [("batch", None, _xyz(5, (_xyz(5) % 4 == 0).all(1)), np.full(512, 3))]),
    "batch_gradient_fill": lambda: (5, [("batch", None, _xyz(5), _xyz(5)[:, 0] % 256)]),
    "batch_hollow_cube": lambda: (5, [("batch", None, _xyz(5, ((_xyz(5) == 0) | (_xyz(5) == 31)).any(1)),
                                       np.full(int(((_xyz(5) == 0) | (_xyz(5) == 31)).any(1).sum()), 3))]),
    "batch_diagonal": lambda: (5, [("batch", None, np.array([[i, i, i] for i in range(32)]), np.full(32, 3))]),
    "batch_random_noise": lambda: (5, [("batch", None, np.random.default_rng(2134).integers(0, 32, (1000, 3)),
                                        np.random.default_rng(2135).integers(1, 256, 1000))]),
}


@pytest.mark.parametrize("name", sorted(REFERENCE_BATCH_CASES))
def test_reference_batch_cases(vx, name):
    api, _ = vx
    depth, steps = REFERENCE_BATCH_CASES[name]()
    it, t, it_o, t_o, err = _run_tree_case(api, depth, steps)
    assert err is None
    _check_tree(api, it, t, it_o, t_o)


def test_double_apply_returns_false(vx):
    api, _ = vx
    it = api.VoxInterner(1, MB)
    t = api.VoxTree(5)
    b = t.create_batch(it)
    b.set(0, 0, 0, 3)
    assert t.apply_batch(it, b)
    assert not t.apply_batch(it, b)


def test_errors_map_reference_panics(vx):
    api, _ = vx
    # H6: applying a fill-only batch that reproduces the root (voxtree.rs:311)
    it, t, it_o, t_o, err = _run_tree_case(api, 3, [("batch", 4, np.zeros((0, 3)), [])] * 2)
    assert err == ("SAME_ROOT", "SAME_ROOT")
    # the reference's Phase 0 took a reference before panicking; the HIP path
    # leaves the tree untouched, so only shape and counts are compared
    _check_tree(api, it, t, it_o, t_o, refs=False)
    # H1: partial edit of a collapsed brick leaf (block_id.rs:273-276)
    xyz = O.all_coords(1)
    steps = [("batch", None, np.concatenate([xyz, [[7, 7, 7]]]), np.concatenate([np.full(8, 5), [1]])),
             ("batch", None, np.array([[0, 0, 0]]), [6])]
    it, t, it_o, t_o, err = _run_tree_case(api, 3, steps)
    assert err == ("LEAF_TYPES", "LEAF_BRICK")
    _check_tree(api, it, t, it_o, t_o)  # the tree is left as it was
    # H9: batch on a depth-1 tree
    it, t, it_o, t_o, err = _run_tree_case(api, 1, [("batch", None, np.array([[0, 0, 0]]), [1])])
    assert err == ("DEPTH", "DEPTH")


def test_fill_and_clear(vx):
    api, _ = vx
    steps = [("fill", 7), ("batch", None, np.array([[1, 2, 3]]), [9]), ("fill", 7), ("fill", 0),
             ("batch", None, _xyz(3), np.arange(512) % 7 + 1), ("fill", 0)]
    it, t, it_o, t_o, err = _run_tree_case(api, 3, steps)
    assert err is None
    _check_tree(api, it, t, it_o, t_o)
    assert it.counts() == (0, 1)


def test_set_sum_mixed_depths_one_launch(vx):
    api, W = vx
    import torch
    depths = [2, 5, 3, 6, 4, 5, 2]
    it = api.VoxInterner(2, 64 * MB)
    it_o = O.Interner(2, 64 * MB)
    f = api.Forest(it, depths)
    ms, vs, trees_o = [], [], []
    for d in depths:
        n = 1 << d
        y, z, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        dense = (x + y + z + 1).astype(np.uint16)
        m, v = W.dense_to_batch(dense, d)
        ms.append(api.to_device(m))
        vs.append(api.to_device(v))
        t_o = O.Tree(d)
        oracle_apply_dense(it_o, t_o, m, v)
        trees_o.append(t_o)
    st = f.apply(ms, vs)
    assert (st == 1).all()
    assert it.counts() == (190, 177 + 1)  # the deepest depth's closed form (D = 6)
    assert_same(it, f, it_o, trees_o, full_reads=False)


def test_terrain_small_grid_and_alternation(vx):
    api, W = vx
    import torch
    grid = 4
    mA, vA, eA, _ = W.terrain_batches(grid=grid, seed=11)
    mB, vB, eB, _ = W.terrain_batches(grid=grid, seed=11, below=2, surface=1)
    it = api.VoxInterner(1, 64 * MB)
    it_o = O.Interner(1, 64 * MB)
    n = grid * grid
    f = api.Forest(it, [5] * n)
    trees_o = [O.Tree(5) for _ in range(n)]
    tA = ([api.to_device(mA[i]) for i in range(n)], [api.to_device(vA[i]) for i in range(n)])
    tB = ([api.to_device(mB[i]) for i in range(n)], [api.to_device(vB[i]) for i in range(n)])
    for step in range(4):
        (m, v), (mh, vh) = (tA, (mA, vA)) if step % 2 == 0 else (tB, (mB, vB))
        st = f.apply(m, v)
        assert (st == 1).all(), st
        for i in range(n):
            assert oracle_apply_dense(it_o, trees_o[i], mh[i], vh[i])[0]
        assert_same(it, f, it_o, trees_o, full_reads=(step < 2), what=f"step {step}")
    s = it.stats()
    assert s["h4_undercount"] == 0 and s["h5_collapse_cloned"] == 0


def test_gc_follows_small_fronts_in_one_launch(vx):
    """Replaced sparse paths are freed by blocks that follow their share of the roots
    down the tree.  Alone in its apply, a depth-6 tree with 210 random voxels frees
    1, 8, 59, 178, 206 and 209 nodes per height, so its block stages more than 512
    freed nodes before one free-stack update; three trees of 419 to 3277 voxels in
    another apply share one block, which passes more than 256 nodes of one height on
    to the next launch.  Alternating values replace every path; counts, hashes and
    reads must stay the oracle's."""
    api, W = vx
    spec = [(6, "sparse", 0.0008, 21), (6, "sparse", 0.0016, 22), (5, "sparse", 0.1, 23), (6, "sparse", 0.01, 24)]
    batches = W.uneven_batches(spec)
    it = api.VoxInterner(2, 128 * MB)
    it_o = O.Interner(2, 128 * MB)
    groups = [[0], [1, 2, 3]]
    forests = [api.Forest(it, [spec[i][0] for i in g]) for g in groups]
    trees_o = [O.Tree(d) for d, _, _, _ in spec]
    for step in range(4):
        vs = [v if step % 2 == 0 else np.where(v == 65535, 1, v + 1).astype(v.dtype) for _, v, _ in batches]
        for f, g in zip(forests, groups):
            st = f.apply([api.to_device(batches[i][0]) for i in g], [api.to_device(vs[i]) for i in g])
            assert (st == 1).all(), st
        for t_o, (m, _, _), v in zip(trees_o, batches, vs):
            b = O.Batch(t_o.depth)
            b.load_dense(m, v)
            assert t_o.apply_batch(it_o, b)
        for f, g in zip(forests, groups):
            sub = [trees_o[i] for i in g]
            g_hash = f.canonical_hashes()
            for k, t_o in enumerate(sub):
                assert g_hash[k] == t_o.canonical_hash(it_o), f"step {step} tree {g[k]}"
                if step >= 2:
                    assert np.array_equal(full_reads_gpu(f, k, t_o.depth), full_reads_oracle(it_o, t_o)), g[k]
        assert it.counts() == it_o.counts(), f"step {step}: pattern counts differ"


@pytest.mark.parametrize("brick_rounds", [0, 1])
def test_uneven_u16_mixed_depths_with_depth7(vx, brick_rounds):
    """Config 3 in miniature: u16, depths 5/6/7, dense set_sum-style and sparse random
    trees in one launch; depth 7 is compared with the oracle's extended depth.  Its
    635 k brick slots take four slots per lane by size (0), or one when forced."""
    api, W = vx
    spec = [(5, "dense", 0, 1), (6, "sparse", 0.01, 2), (7, "sparse", 0.001, 3), (6, "dense", 0, 4),
            (5, "sparse", 0.1, 5), (7, "dense", 0, 6), (5, "sparse", 0.001, 7), (6, "sparse", 0.1, 8)]
    batches = W.uneven_batches(spec)
    it = api.VoxInterner(2, 256 * MB)
    it.set_tuning(brick_rounds=brick_rounds)
    it_o = O.Interner(2, 256 * MB, extended_depth=True)
    f = api.Forest(it, [d for d, _, _, _ in spec])
    trees_o = []
    for (d, _, _, _), (m, v, _) in zip(spec, batches):
        t_o = O.Tree(d)
        b = O.Batch(d)
        b.load_dense(m, v)
        assert t_o.apply_batch(it_o, b)
        trees_o.append(t_o)
    st = f.apply([api.to_device(m) for m, _, _ in batches], [api.to_device(v) for _, v, _ in batches])
    assert (st == 1).all(), st
    assert_same(it, f, it_o, trees_o, full_reads=False, what="uneven")
    # full reads on the depth-5/6 trees, a random sample on the depth-7 ones
    rng = np.random.default_rng(0)
    for i, t_o in enumerate(trees_o):
        if t_o.depth <= 6:
            g = full_reads_gpu(f, i, t_o.depth)
            o = O.all_coords(t_o.depth)
            ov, op = t_o.get_many(it_o, o)
            assert np.array_equal(g, np.where(op == 1, ov, 0)), i
        else:
            xyz = rng.integers(0, 128, size=(20000, 3)).astype(np.int32)
            gv, gp = f.get(np.full(len(xyz), i, dtype=np.uint32), xyz)
            ov, op = t_o.get_many(it_o, xyz)
            assert np.array_equal(np.where(gp == 1, gv, 0), np.where(op == 1, ov, 0)), i
    # replace every tree with a shifted dense fill (exercises GC of mixed depths)
    batches2 = [W.dense_to_batch((W.uneven_dense(d) % 7 + 1).astype(np.uint16), d) for d, _, _, _ in spec]
    for t_o, (m, v) in zip(trees_o, batches2):
        b = O.Batch(t_o.depth)
        b.load_dense(m, v)
        assert t_o.apply_batch(it_o, b)
    st = f.apply([api.to_device(m) for m, _ in batches2], [api.to_device(v) for _, v in batches2])
    assert (st == 1).all(), st
    assert_same(it, f, it_o, trees_o, full_reads=False, what="uneven replaced")


def test_edit_lists_sorted_on_device(vx):
    """vx_trees_apply_edits against the oracle's Batch::just_set in submission order:
    duplicates (last write wins), default values (clears, ignored by apply), trees
    without edits, mixed depths, u16, and an out-of-range coordinate."""
    api, W = vx
    rng = np.random.default_rng(42)
    depths = [5, 2, 6, 5, 3, 7, 5, 4]
    it = api.VoxInterner(2, 256 * MB)
    it_o = O.Interner(2, 256 * MB, extended_depth=True)
    f = api.Forest(it, depths)
    trees_o = [O.Tree(d) for d in depths]
    for rnd in range(3):
        xyzs, vals, offs = [], [], [0]
        expect = []
        for t, d in enumerate(depths):
            n = 0 if (t == 1 and rnd == 0) else int(rng.integers(1, 3000))
            n_side = 1 << d
            xyz = rng.integers(0, n_side, size=(n, 3)).astype(np.int32)
            if n:
                xyz[n // 2:] = xyz[: n - n // 2]  # many duplicate voxels
            v = rng.integers(0, 9, size=n).astype(np.uint16)  # includes 0 = clear
            bad = (t == 4 and rnd == 1)
            if bad:
                xyz[n // 3] = [n_side, 0, 0]
            xyzs.append(xyz)
            vals.append(v)
            offs.append(offs[-1] + n)
            expect.append(-4 if bad else None)
        xyz_all = np.concatenate(xyzs)
        val_all = np.concatenate(vals)
        st = f.apply_edits(offs, api.to_device(xyz_all), api.to_device(val_all))
        for t, d in enumerate(depths):
            if expect[t] == -4:
                assert st[t] == -4
                continue
            if len(xyzs[t]) == 0:
                assert st[t] == 0
                continue
            b = O.Batch(d)
            b.set_many(xyzs[t], vals[t].astype(np.uint64))
            try:
                ch = trees_o[t].apply_batch(it_o, b)
                assert st[t] == (1 if ch else 0), (rnd, t, st[t], ch)
            except O.OraclePanic as e:
                assert (e.kind, st[t]) in (("SAME_ROOT", -6), ("LEAF_TYPES", -5)), (e.kind, st[t])
        assert_same(it, f, it_o, trees_o, full_reads=False, what=f"edits round {rnd}")
        for i, t_o in enumerate(trees_o):
            if t_o.depth <= 5:
                assert np.array_equal(full_reads_gpu(f, i, t_o.depth),
                                      np.where(t_o.get_many(it_o, O.all_coords(t_o.depth))[1] == 1,
                                               t_o.get_many(it_o, O.all_coords(t_o.depth))[0], 0))


def test_graph_replay_matches_direct_launches(vx):
    """Repeated batched applies replayed from captured HIP graphs give the same trees,
    counts and statuses as launching every kernel directly, also with another call in
    between whose larger workspace invalidates the captured sequences."""
    api, W = vx
    grid = 4
    n = grid * grid
    mA, vA, _, _ = W.terrain_batches(grid=grid, seed=31)
    vB = vA.copy()
    vB[vA == 1] = 2
    vB[vA == 2] = 1
    mC, vC, _, _ = W.terrain_batches(grid=2 * grid, seed=32)  # 4x the trees: workspace grows
    dev = lambda arrs: [api.to_device(a) for a in arrs]
    tA, tB = (dev(mA), dev(vA)), (dev(mA), dev(vB))
    tC = (dev(mC), dev(vC))
    runs = []
    for graphs in (False, True):
        it = api.VoxInterner(1, 64 * MB)
        it.set_graphs(graphs)
        f = api.Forest(it, [5] * n)
        fc = api.Forest(it, [5] * (4 * n))
        vA_, vB_ = f.make_views(*tA), f.make_views(*tB)
        vC_ = fc.make_views(*tC)
        log = []
        for step in range(12):
            if step == 6:
                fc.apply_views(vC_)
                it.sync()
                log.append(("C", fc.status.cpu().numpy().tolist(), fc.canonical_hashes()))
            f.apply_views(vA_ if step % 2 == 0 else vB_)
            it.sync()
            log.append((step, f.status.cpu().numpy().tolist(), f.canonical_hashes(), it.counts()))
        runs.append(log)
        it.close()
    assert runs[0] == runs[1]


@pytest.mark.parametrize("tile,dedup,grid_blocks,pb,br", [(128, 0, 64, 0, 4), (512, 2, 4096, 1, 1), (256, 8, 1024, 3, 4),
                                                        (256, 4, 1024, 0, 0), (256, 8, 1024, 0, 2)])
def test_launch_shapes_match_oracle(vx, tile, dedup, grid_blocks, pb, br):
    """Every tile size / probe schedule / grid size / probe tile grouping the
    autotuner may pick gives the oracle's trees (two applies, so the second replaces
    and reclaims the first).  probe_blocks 1 packs 16 tiles into every probe block
    (several passes per block), 3 packs 16 at the brick level and fewer above; br is
    the brick slots per lane of the brick pass (0: chosen by size)."""
    api, W = vx
    grid = 4
    n = grid * grid
    mA, vA, _, _ = W.terrain_batches(grid=grid, seed=41)
    mB = mA  # same shapes, other ids: no partial edit of a collapsed brick (a reference panic)
    vB = vA.copy()
    vB[vA == 1] = 3
    vB[vA == 2] = 1
    it = api.VoxInterner(1, 64 * MB)
    it.set_tuning(tile=tile, dedup_rounds=dedup, commit_grid=grid_blocks, gc_grid=grid_blocks,
                  probe_blocks_bricks=pb, probe_blocks=pb, brick_merge_rounds=dedup, brick_rounds=br)
    assert it.get_tuning() == {"tile": tile, "dedup_rounds": dedup, "commit_grid": grid_blocks,
                               "gc_grid": grid_blocks, "probe_blocks_bricks": pb, "probe_blocks": pb,
                               "brick_merge_rounds": dedup, "brick_rounds": br}
    it_o = O.Interner(1, 64 * MB)
    f = api.Forest(it, [5] * n)
    trees_o = [O.Tree(5) for _ in range(n)]
    for m, v in ((mA, vA), (mB, vB)):
        st = f.apply([api.to_device(x) for x in m], [api.to_device(x) for x in v])
        assert (st == 1).all(), st
        for i in range(n):
            oracle_apply_dense(it_o, trees_o[i], m[i], v[i])
    assert_same(it, f, it_o, trees_o, full_reads=True, what=f"tile {tile}")


@pytest.mark.parametrize("lod", [0, 1, 3])
def test_to_vec_with_lod_matches_oracle(vx, lod):
    """Dense extraction of every tree (full resolution and reduced LOD depths, where
    branches contribute their LOD values) against the oracle's to_vec."""
    api, W = vx
    grid = 3
    n = grid * grid
    m, v, _, _ = W.terrain_batches(grid=grid, seed=51, below=3, surface=7)
    spec = [(5, "dense", 0, 1), (6, "sparse", 0.01, 2), (6, "dense", 0, 4), (5, "sparse", 0.1, 5)]
    ub = W.uneven_batches(spec)
    for vsize, depths, batches in ((1, [5] * n, [(m[i], v[i]) for i in range(n)]),
                                   (2, [d for d, _, _, _ in spec], [(mm, vv) for mm, vv, _ in ub])):
        it = api.VoxInterner(vsize, 64 * MB)
        it_o = O.Interner(vsize, 64 * MB)
        f = api.Forest(it, depths)
        st = f.apply([api.to_device(a) for a, _ in batches], [api.to_device(b) for _, b in batches])
        assert (st == 1).all()
        got = f.to_vec(lod)
        for i, (d, (mm, vv)) in enumerate(zip(depths, batches)):
            t = O.Tree(d)
            oracle_apply_dense(it_o, t, mm, vv)
            exp = t.to_vec(it_o, lod)
            assert np.array_equal(got[i].reshape(-1).astype(np.uint64), exp), (vsize, i, lod)
        it.close()


def test_edit_lists_many_rounds_per_sort_block(vx):
    """Edit lists large enough that every block of the device radix sort walks several
    rounds of its range (1.9 M edits over 64 trees), with many duplicate voxels whose
    last write must win (the sort is stable, core/batch.rs:145-175)."""
    api, W = vx
    rng = np.random.default_rng(314)
    depths = [5] * 48 + [4] * 16
    it = api.VoxInterner(2, 512 * MB)
    it_o = O.Interner(2, 512 * MB)
    f = api.Forest(it, depths)
    trees_o = [O.Tree(d) for d in depths]
    xyzs, vals, offs = [], [], [0]
    for d in depths:
        k = 30000
        xyzs.append(rng.integers(0, 1 << d, size=(k, 3)).astype(np.int32))
        vals.append(rng.integers(0, 50, size=k).astype(np.uint16))  # 0: a clear, ignored by the apply
        offs.append(offs[-1] + k)
    st = f.apply_edits(offs, api.to_device(np.concatenate(xyzs)), api.to_device(np.concatenate(vals)))
    assert (st == 1).all()
    for t, xyz, v in zip(trees_o, xyzs, vals):
        b = O.Batch(t.depth)
        b.set_many(xyz, v.astype(np.uint64))
        assert t.apply_batch(it_o, b)
    assert_same(it, f, it_o, trees_o, full_reads=True, what="large edit lists")


def test_unchanged_full_bricks_are_untouched(vx):
    """A completely written brick is rebuilt from its values alone (its old children are
    not read); when no value changed its probe finds the old node, and the level above
    must count it as untouched, as the reference leaves an unchanged brick unpushed
    (voxtree.rs:865-868): re-applying a full batch returns false, and a full batch that
    changes a few bricks only changes exactly what the oracle changes."""
    api, W = vx
    from gpu_helpers import assert_same, oracle_apply_dense
    depth, n = 4, 6
    side = 1 << depth
    rng = np.random.default_rng(31)
    dense = rng.integers(1, 5, size=(n, side, side, side)).astype(np.uint16)
    dense[1] = dense[0]  # two trees share every node
    it = api.VoxInterner(2, 64 * MB)
    f = api.Forest(it, [depth] * n)
    it_o = O.Interner(2, 64 * MB)
    trees_o = [O.Tree(depth) for _ in range(n)]

    def step(d, want):
        mv = [W.dense_to_batch(d[t], depth) for t in range(n)]
        st = f.apply([api.to_device(m) for m, _ in mv], [api.to_device(v) for _, v in mv])
        oks = [int(oracle_apply_dense(it_o, t, m, v)[0]) for t, (m, v) in zip(trees_o, mv)]
        assert list(st) == oks == want, (list(st), oks, want)
        assert_same(it, f, it_o, trees_o, full_reads=True, what=f"want {want}")
        roots = f.roots_host()
        for i, t in enumerate(trees_o):
            assert it.get_ref(roots[i]) == it_o.get_ref(t.root), i

    step(dense, [1] * n)
    step(dense, [0] * n)                    # nothing changed: every tree returns false
    d2 = dense.copy()
    d2[2, 0, 0, 0] = 7                       # one brick of tree 2
    d2[3, 4:6, 8:10, 12:14] = d2[3, 4:6, 8:10, 12:14] % 4 + 1  # one brick of tree 3, all its voxels
    step(d2, [0, 0, 1, 1, 0, 0])
    step(d2, [0] * n)
    it.close()


def test_reapplied_all_same_brick_is_pushed(vx):
    """The reference pushes an all-same brick whether or not it changed (voxtree.rs:887-895),
    so re-applying a batch whose only unchanged-but-pushed bricks are all-same ones rebuilds
    the same root and panics on assert_ne! (voxtree.rs:311, VX_E_SAME_ROOT here), while a
    tree whose bricks are all mixed and unchanged returns false."""
    api, W = vx
    depth, n = 4, 3
    side = 1 << depth
    rng = np.random.default_rng(41)
    dense = rng.integers(1, 5, size=(n, side, side, side)).astype(np.uint16)
    # no accidental all-same brick in tree 0; one all-same brick in trees 1 and 2
    for t in range(n):
        for y in range(0, side, 2):
            for z in range(0, side, 2):
                for x in range(0, side, 2):
                    b = dense[t, y:y + 2, z:z + 2, x:x + 2]
                    if (b == b.flat[0]).all():
                        b[0, 0, 0] = b.flat[0] % 4 + 1
    dense[1, 0:2, 0:2, 0:2] = 5
    dense[2, 6:8, 2:4, 10:12] = 3
    it = api.VoxInterner(2, 64 * MB)
    f = api.Forest(it, [depth] * n)
    it_o = O.Interner(2, 64 * MB)
    trees_o = [O.Tree(depth) for _ in range(n)]
    mv = [W.dense_to_batch(dense[t], depth) for t in range(n)]
    views = ([api.to_device(m) for m, _ in mv], [api.to_device(v) for _, v in mv])
    for step, want in ((0, [1, 1, 1]), (1, [0, -6, -6])):
        st = f.apply(*views)
        got_o = []
        for t, (m, v) in zip(trees_o, mv):
            b = O.Batch(depth)
            b.load_dense(m, v)
            try:
                got_o.append(int(t.apply_batch_atomic(it_o, b)))
            except O.OraclePanic as e:
                got_o.append({"SAME_ROOT": -6}[e.kind])
        assert list(st) == got_o == want, (step, list(st), got_o)
    from gpu_helpers import assert_same
    assert_same(it, f, it_o, trees_o, full_reads=True, what="after the re-apply")
    it.close()
