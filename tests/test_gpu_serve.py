"""The serving loop (continuous batching) against the oracle executing the same
requests one at a time in arrival order: every read result, every status, every
tree's canonical hash and the interner's pattern counts must match.  Requests the
reference rejects with a panic (H1, H6) leave the tree and the interner unchanged on
both sides (the oracle restores a snapshot)."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_helpers import oracle_apply_dense

pytestmark = pytest.mark.gpu

MB = 1024 * 1024


@pytest.fixture(scope="module")
def mods():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    from voxelis_amd import api, serve, workloads
    return api, serve, workloads


def _base_world(api, serve, W, shape, budget):
    world = serve.World(2, budget, shape, depth=5)
    it_o = O.Interner(2, budget)
    trees_o = [O.Tree(5) for _ in range(world.n)]
    # base content: a flat-ish dune field (ids 1 below, 2 at the surface)
    sx, sy, sz = shape
    heights = W.terrain_heights(grid=max(sx, sz), chunk=32, seed=5)
    ms, vs = [], []
    for t in range(world.n):
        tx = t % sx
        ty = (t // sx) % sy
        tz = t // (sx * sy)
        dense = W.terrain_dense(heights, tx, tz).astype(np.uint16)
        if ty > 0:
            dense[:] = 0
        m, v = W.dense_to_batch(dense, 5)
        ms.append(m)
        vs.append(v)
    st = world.forest.apply([api.to_device(m) for m in ms], [api.to_device(v) for v in vs],
                            has_patches=[bool(m[:, 0].any()) for m in ms])
    for t in range(world.n):
        if ms[t][:, 0].any():
            oracle_apply_dense(it_o, trees_o[t], ms[t], vs[t])
    return world, it_o, trees_o


def _oracle_run(req, world, it_o, trees_o):
    if req.kind == "get":
        chunks, local = req.payload
        out = np.zeros(len(chunks), dtype=np.uint64)
        for t in np.unique(chunks):
            sel = chunks == t
            v, p = trees_o[int(t)].get_many(it_o, local[sel])
            out[sel] = np.where(p == 1, v, 0)
        return out
    if req.kind == "fill":
        trees_o[int(req.trees[0])].fill(it_o, req.payload)
        return 1
    st = []
    for t, (local, vals) in sorted(req.payload.items()):
        b = O.Batch(5)
        b.set_many(local, vals.astype(np.uint64))
        try:
            # a reference panic leaves the tree and the interner as they were (the device
            # path's policy; the reference aborts the process)
            st.append(1 if trees_o[t].apply_batch_atomic(it_o, b) else 0)
        except O.OraclePanic as e:
            st.append({"SAME_ROOT": -6, "LEAF_TYPES": -5}[e.kind])
    return np.array(st)


def _check_against_oracle(reqs, world, it_o, trees_o):
    for r in reqs:
        exp = _oracle_run(r, world, it_o, trees_o)
        if r.kind == "get":
            assert np.array_equal(r.result.astype(np.uint64), exp), r.rid
        elif r.kind == "sphere":
            assert np.array_equal(np.asarray(r.result), exp), (r.rid, r.result, exp)
    hashes = world.forest.canonical_hashes()
    for t in range(world.n):
        assert hashes[t] == trees_o[t].canonical_hash(it_o), t
    assert world.it.counts() == it_o.counts()


@pytest.mark.parametrize("carve", [False, True])
def test_served_equals_sequential_oracle(mods, carve):
    api, serve, W = mods
    shape = (4, 2, 4)
    world, it_o, trees_o = _base_world(api, serve, W, shape, 64 * MB)
    reqs = serve.make_requests(world, 80, rate=1e6, seed=7 + carve, max_gets=4096, carve=carve)
    srv = serve.Server(world)
    lat, wall = srv.serve(reqs, realtime=False)
    assert srv.cycles < len(reqs)  # requests were batched
    assert srv.fallbacks == 0  # hard-bound reservations never roll back
    _check_against_oracle(reqs, world, it_o, trees_o)


def test_realtime_overlapping_cycles(mods):
    """Poisson arrivals in real time, two cycles in flight, reads on their own stream."""
    api, serve, W = mods
    shape = (4, 2, 4)
    world, it_o, trees_o = _base_world(api, serve, W, shape, 64 * MB)
    reqs = serve.make_requests(world, 120, rate=4000.0, seed=9, max_gets=2048, carve=True)
    srv = serve.Server(world, max_inflight=2)
    lat, wall = srv.serve(reqs, realtime=True)
    assert len(lat) == len([r for r in reqs if r.kind != "noop"])
    _check_against_oracle(reqs, world, it_o, trees_o)


class _Under:
    """A predictor that under-predicts every request (forces rolled-back applies)."""

    def predict(self, world, reqs):
        return [0.0 for _ in reqs]


def test_under_prediction_rolls_back_and_recounts(mods):
    api, serve, W = mods
    shape = (4, 2, 4)
    world, it_o, trees_o = _base_world(api, serve, W, shape, 64 * MB)
    reqs = serve.make_requests(world, 60, rate=1e6, seed=13, max_gets=1024, mix=(0.4, 0.55, 0.05))
    srv = serve.Server(world, predictor=_Under(), margin=1.0)
    srv.serve(reqs, realtime=False)
    assert srv.fallbacks > 0 and srv.recounts > 0
    _check_against_oracle(reqs, world, it_o, trees_o)


def test_admission_by_budget_defers_but_completes(mods):
    """A tight interner budget: hard bounds do not fit, so edits wait and are counted
    exactly; every request still completes with the sequential result."""
    api, serve, W = mods
    shape = (2, 1, 2)
    world, _, _ = _base_world(api, serve, W, shape, 1 * MB)
    _, it_o, trees_o = _base_world(api, serve, W, shape, 64 * MB)
    # big spheres: their hard bounds (every touched node position) exceed the free
    # capacity, their exact growth does not
    reqs = serve.make_requests(world, 30, rate=1e6, seed=3, max_gets=512, mix=(0.5, 0.4, 0.1),
                               radius=(30.0, 40.0))
    srv = serve.Server(world)
    srv.serve(reqs, realtime=False)
    assert srv.recounts > 0
    for r in reqs:
        if r.kind == "sphere":
            assert (np.asarray(r.result) != serve.VX_E_OOM).all(), r.rid
    for r in reqs:
        exp = _oracle_run(r, world, it_o, trees_o)
        if r.kind == "get":
            assert np.array_equal(r.result.astype(np.uint64), exp), r.rid
    hashes = world.forest.canonical_hashes()
    for t in range(world.n):
        assert hashes[t] == trees_o[t].canonical_hash(it_o), t
