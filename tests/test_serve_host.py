"""Host-side pieces of the serving loop (CPU): sphere edit lists, the hard bound of a
request's new nodes, and the predictor's features.

The hard bound is what admission reserves without a predictor, so it must never be
below the nodes an apply creates: checked against the oracle's get_or_create misses
on random spheres over non-empty trees, and against brute-force position counts."""
import numpy as np

import oracle_lib as O
from voxelis_amd import serve


def _fake_world(shape=(2, 2, 2), depth=5):
    w = object.__new__(serve.World)
    w.shape = shape
    w.depth = depth
    w.side = 1 << depth
    w.n = int(np.prod(shape))
    w.kind = np.zeros(w.n, dtype=np.int8)
    w.history = np.zeros(w.n, dtype=np.int64)
    return w


def test_touched_positions_brute_force():
    rng = np.random.default_rng(0)
    for _ in range(20):
        k = int(rng.integers(1, 200))
        lo = rng.integers(0, 20, 3)
        pts = (lo + rng.integers(0, 12, (k, 3))).astype(np.int32)
        b = serve.touched_positions(pts, 5)
        for h in range(1, 6):
            exact = len({tuple(p >> h) for p in pts})
            assert b[h - 1] >= exact
    # a full box: the bound is exact
    g = np.stack(np.meshgrid(np.arange(4, 12), np.arange(0, 8), np.arange(8, 16), indexing="ij"), -1).reshape(-1, 3)
    b = serve.touched_positions(g.astype(np.int32), 5)
    for h in range(1, 6):
        assert b[h - 1] == len({tuple(p >> h) for p in g})


def test_covered_bricks_brute_force():
    rng = np.random.default_rng(1)
    pts = np.unique(rng.integers(0, 10, (600, 3)), axis=0).astype(np.int32)
    touched, covered = serve.covered_bricks(pts)
    bricks = {}
    for p in pts:
        bricks.setdefault(tuple(p >> 1), set()).add(tuple(p))
    assert touched == len(bricks)
    assert covered == sum(1 for v in bricks.values() if len(v) == 8)


def test_sphere_edits_partition_the_ball():
    w = _fake_world()
    c = np.array([30, 33, 20])
    per_tree = serve.sphere_edits(w, c, 9.5, 7)
    pts = []
    for t, (local, vals) in per_tree.items():
        tz, rem = divmod(t, 4)
        ty, tx = divmod(rem, 2)
        pts.append(local + np.array([tx, ty, tz]) * 32)
        assert (vals == 7).all()
    pts = np.concatenate(pts)
    d2 = ((pts - c) ** 2).sum(axis=1)
    assert (d2 <= 9.5 ** 2).all()
    assert len(pts) == len(np.unique(pts, axis=0))
    assert len(pts) == sum(1 for x in range(20, 41) for y in range(23, 44) for z in range(10, 31)
                           if (x - 30) ** 2 + (y - 33) ** 2 + (z - 20) ** 2 <= 90.25)


def test_hard_bound_covers_oracle_growth():
    w = _fake_world()
    it_o = O.Interner(2, 256 << 20)
    trees = [O.Tree(5) for _ in range(w.n)]
    rng = np.random.default_rng(2)
    # some base content
    for t in trees[:4]:
        b = t.create_batch()
        b.just_fill(3)
        t.apply_batch(it_o, b)
    for i in range(25):
        c = rng.integers(0, 64, 3)
        val = 0 if i % 3 == 0 else int(rng.integers(1, 60000))
        r = serve.Request(i, "sphere", 0.0)
        r.payload = serve.sphere_edits(w, c, float(rng.uniform(2, 14)), val)
        if not r.payload:
            continue
        r.trees = np.array(sorted(r.payload), dtype=np.int64)
        r.edits = sum(len(v) for _, v in r.payload.values())
        serve.describe(w, r)
        s0 = it_o.stats()
        for t, (local, vals) in r.payload.items():
            b = O.Batch(5)
            b.set_many(local, vals.astype(np.uint64))
            try:
                trees[t].apply_batch(it_o, b)
            except O.OraclePanic:
                pass
        s1 = it_o.stats()
        grown = (s1["leaf_misses"] - s0["leaf_misses"]) + (s1["branch_misses"] - s0["branch_misses"])
        assert grown <= r.bound, (i, grown, r.bound)


def test_features_shape_and_flags():
    from voxelis_amd import predictor
    w = _fake_world()
    r = serve.Request(0, "sphere", 0.0)
    r.payload = serve.sphere_edits(w, np.array([40, 10, 40]), 6.0, 0)
    r.trees = np.array(sorted(r.payload), dtype=np.int64)
    r.edits = sum(len(v) for _, v in r.payload.values())
    serve.describe(w, r)
    f = predictor.features(w, r)
    assert f.shape == (16,) and f.dtype == np.float32
    assert f[12] == 1.0 and f[14] == 0.0 and f[9] == 1.0  # carving, not a fill, empty trees
    assert np.isclose(f[15], np.log1p(r.bound))
