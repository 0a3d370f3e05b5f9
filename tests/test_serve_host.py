"""Host-side pieces of the serving loop (CPU): sphere edit lists, the hard bound of a
request's new nodes, the predictor's features, and the scheduler itself on a fake
device (every request completes, rolled-back applies are retried, each tree sees its
edits in arrival order).

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
    w.version = np.zeros(w.n, dtype=np.int64)
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
    k = len(r.payload)
    assert f.shape == (k, 16) and f.dtype == np.float32
    assert (f[:, 12] == 1).all()  # carving
    assert (f[:, 9] == 0).all() and (f[:, 10] == 0).all() and (f[:, 14] == 0).all()  # not sampled yet
    assert np.isclose(np.expm1(f[:, 15].astype(np.float64)).sum() + 1, r.bound, rtol=1e-5)
    assert np.allclose(np.expm1(f[:, 0].astype(np.float64)), [len(v) for _, (_, v) in sorted(r.payload.items())],
                       rtol=1e-5)
    r.occ = (np.full(k, 0.25), np.full(k, 0.5), np.ones(k), np.full(k, 0.125))
    f = predictor.features(w, r)
    assert (f[:, 9] == 0.25).all() and (f[:, 10] == 0.5).all() and (f[:, 14] == 1).all()
    assert (f[:, 13] == 0.125).all()
    seg = r.sample[3]
    assert len(seg) == k + 1 and (np.diff(seg) <= serve.SAMPLES_PER_TREE).all() and r.sample[2] == 0
    # one voxel of every partly written brick (a sphere's boundary), at even coordinates
    pseg = r.sample[6]
    assert len(pseg) == k + 1 and (r.sample[5] % 2 == 0).all()
    for i, (t, (local, _)) in enumerate(sorted(r.payload.items())):
        b, cnt = np.unique(local // 2, axis=0, return_counts=True)
        exp = {tuple(x * 2) for x in b[cnt < 8]}
        assert {tuple(x) for x in r.sample[5][pseg[i]:pseg[i + 1]]} == exp
    fill = serve.Request(1, "fill", 0.0)
    fill.payload, fill.trees = 3, np.array([2])
    serve.describe(w, fill)
    assert fill.bound == 1 and predictor.features(w, fill).shape == (0, 16)


# ---------------------------------------------------------------- scheduler on a fake device

class _FakeDevice:
    """Stands in for the GPU: an apply's exact growth is one node per 50 edits of each
    tree (at least one); an apply beyond its node_limit is rolled back (VX_E_BUDGET)."""

    def __init__(self, capacity):
        self.cap = capacity
        self.alive = 100

    @staticmethod
    def per_tree(trees, offs):
        return [max(1, (offs[i + 1] - offs[i]) // 50) for i in range(len(trees))]


def _fake_device_world(monkeypatch, capacity):
    import contextlib
    import torch

    class Ev:
        def record(self, s=None):
            pass

        def query(self):
            return True

        def synchronize(self):
            pass

    class St:
        cuda_stream = 0

        def wait_event(self, e):
            pass

    monkeypatch.setattr(torch.cuda, "Stream", lambda device=None: St())
    monkeypatch.setattr(torch.cuda, "Event", lambda *a, **k: Ev())
    monkeypatch.setattr(torch.cuda, "current_stream", lambda: St())
    monkeypatch.setattr(torch.cuda, "stream", lambda s: contextlib.nullcontext())
    monkeypatch.setattr(serve.api, "to_device", lambda a, device=None: a)
    w = _fake_world(shape=(4, 2, 4))
    dev = _FakeDevice(capacity)

    class Sub:
        def __init__(self, n, st, b, per):
            self.status = torch.tensor(st, dtype=torch.int32)
            self._c = (b, 0, np.asarray(per))

        def counts_host(self):
            return self._c

    def apply_edits(trees, offs, xyz, vals, dry_run=False, node_limit=0, counts=False):
        per = dev.per_tree(trees, offs)
        g = sum(per)
        if dry_run:
            return Sub(len(trees), [1] * len(trees), g, per)
        if node_limit and g > node_limit:
            return Sub(len(trees), [serve.VX_E_BUDGET] * len(trees), g, per)
        dev.alive += g
        return Sub(len(trees), [1] * len(trees), g, per)

    def fill(trees, vals):
        dev.alive += 1

    class It:
        value_size = 2
        np_dtype = np.uint16

        def capacity(self):
            return dev.cap

        def alive_nodes(self):
            return dev.alive

        def sync(self):
            pass

    class Forest:
        def get_device(self, chunk_t, xyz_t, out, pres, stream=None, stop=None):
            pass

    w.it = It()
    w.dev = "cpu"
    w.forest = Forest()
    w.subforest_apply_edits = apply_edits
    w.subforest_fill = fill
    return w, dev


class _Under:
    def predict(self, world, reqs):
        return [0.0] * len(reqs)


def test_scheduler_completes_and_keeps_per_tree_order(monkeypatch):
    for predictor, cap in ((None, 10 ** 9), (None, 2500), (_Under(), 10 ** 9)):
        w, dev = _fake_device_world(monkeypatch, cap)
        reqs = serve.make_requests(w, 150, rate=1e6, seed=4, mix=(0.0, 0.9, 0.1), radius=(2.0, 20.0))
        srv = serve.Server(w, predictor=predictor, margin=1.0)
        order = []
        orig = srv._launch

        def launch(cyc):
            order.append([r.rid for r in cyc.spheres])
            orig(cyc)
        srv._launch = launch
        srv.serve(reqs, realtime=False)
        assert all(r.done is not None for r in reqs if r.kind != "noop")
        if predictor is not None:
            assert srv.fallbacks > 0 and srv.recounts > 0
        # every tree saw its edit requests in arrival order
        applied = {}
        for cyc_rids in order:
            for rid in cyc_rids:
                r = reqs[rid]
                if np.asarray(r.result).size and (np.asarray(r.result) != serve.VX_E_BUDGET).all():
                    for t in r.trees:
                        applied.setdefault(int(t), []).append(rid)
        for t, rids in applied.items():
            assert rids == sorted(rids), (t, rids)
