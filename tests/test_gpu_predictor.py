"""The predictor's HIP kernels against the float64 oracle, and what the trained network
gains in admission over reserving hard bounds.

Tolerances (FP32 kernels vs float64 oracle): forward |dy| <= 1e-5 * (1 + |y|);
gradient max |dg| <= 1e-4 * max |g| (sums over the batch in FP32); Adam step
|dp| <= 1e-6 * (1 + |p|).
"""
import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu

MB = 1024 * 1024


@pytest.fixture(scope="module")
def mods():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    from voxelis_amd import api, predictor, serve, workloads
    return torch, api, predictor, serve, workloads


def _case(seed, n):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.4, M.N_PARAMS).astype(np.float32)
    x = rng.normal(0, 1, (n, 16)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    return p, x, t


def test_forward_matches_oracle(mods):
    torch, api, P, _, _ = mods
    p, x, _ = _case(1, 1001)
    net = P.GrowthMLP(params=p)
    y = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    y_o, _ = M.forward(p.astype(np.float64), x.astype(np.float64))
    assert np.all(np.abs(y - y_o) <= 1e-5 * (1 + np.abs(y_o)))


@pytest.mark.parametrize("n", [1, 5, 777, 4096])
def test_gradient_matches_oracle(mods, n):
    torch, api, P, _, _ = mods
    p, x, t = _case(2 + n, n)
    net = P.GrowthMLP(params=p)
    g, y, loss = net.gradient(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    g = g.cpu().numpy()
    g_o, y_o, l_o = M.gradient(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert np.all(np.abs(y.cpu().numpy() - y_o) <= 1e-5 * (1 + np.abs(y_o)))
    assert abs(float(loss.item()) - l_o) <= 1e-5 * (1 + l_o)
    assert np.max(np.abs(g - g_o)) <= 1e-4 * np.max(np.abs(g_o)), np.max(np.abs(g - g_o))


def test_adam_matches_oracle(mods):
    torch, api, P, _, _ = mods
    p, x, t = _case(9, 300)
    net = P.GrowthMLP(params=p)
    xt, tt = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    po = p.astype(np.float64)
    m = np.zeros_like(po)
    v = np.zeros_like(po)
    for step in range(1, 4):
        g, _, _ = net.gradient(xt, tt)
        net.adam_step(g, 1e-3)
        go = g.cpu().numpy().astype(np.float64)  # same gradient on both sides
        po, m, v = M.adam(po, go, m, v, step, 1e-3)
    assert np.all(np.abs(net.p.cpu().numpy() - po) <= 1e-6 * (1 + np.abs(po)))


def test_training_reduces_loss(mods):
    torch, api, P, _, _ = mods
    rng = np.random.default_rng(11)
    x = rng.normal(0, 1, (2048, 16)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2]).astype(np.float32)
    net = P.GrowthMLP(seed=3)
    hist = net.train(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda(), epochs=60, batch=256, lr=3e-3)
    assert hist[-1] < 0.2 * hist[0], hist[::10]


def _world(api, serve, W, budget, shape=(8, 2, 8)):
    world = serve.World(2, budget, shape, depth=5)
    sx, sy, sz = shape
    heights = W.terrain_heights(grid=max(sx, sz), chunk=32, seed=5)
    ms, vs, hp = [], [], []
    for t in range(world.n):
        tx, ty, tz = t % sx, (t // sx) % sy, t // (sx * sy)
        dense = W.terrain_dense(heights, tx, tz).astype(np.uint16)
        if ty > 0:
            dense[:] = 0
        m, v = W.dense_to_batch(dense, 5)
        ms.append(api.to_device(m))
        vs.append(api.to_device(v))
        hp.append(bool(m[:, 0].any()))
        world.kind[t] = 2 if hp[-1] else 0
    world.forest.apply(ms, vs, has_patches=hp)
    return world


def _edits(serve, world, n, seed):
    return serve.make_requests(world, n, rate=1e6, seed=seed, mix=(0.0, 0.9, 0.1), carve=True,
                               radius=(2.0, 24.0))


def test_predictor_gain_over_hard_bounds(mods):
    """Train on logged (features, exact dry-run growth) pairs, calibrate the margin on
    held-out pairs, then serve the same edit stream under a tight interner budget with
    hard-bound reservations and with the network: same final state, fewer cycles and
    fewer exact counts with the network, few rolled-back applies."""
    torch, api, P, serve, W = mods
    world = _world(api, serve, W, 256 * MB)
    srv = serve.Server(world, log_exact=True)
    srv.serve(_edits(serve, world, 1500, seed=100), realtime=False)
    pred, margin, rep = P.Predictor.fit_from_log(srv.log, holdout=0.3, quantile=0.995)
    print(f"\npredictor: {rep}")
    assert rep["mae_nodes"] < 0.1 * rep["hard_bound_mae_nodes"]
    assert rep["heldout_under_rate"] <= 0.02
    base = _world(api, serve, W, 256 * MB).it.alive_nodes()
    budget = (base + 25000) * 80
    runs = {}
    for name, model in (("bound", None), ("mlp", pred)):
        w = _world(api, serve, W, budget)
        s = serve.Server(w, predictor=model, margin=margin)
        reqs = _edits(serve, w, 300, seed=200)
        s.serve(reqs, realtime=False)
        n_edit = sum(1 for r in reqs if r.kind in ("sphere", "fill"))
        assert all(r.done is not None for r in reqs if r.kind != "noop")
        runs[name] = (s.cycles, s.recounts, s.fallback_requests, n_edit, w.forest.canonical_hashes(),
                      w.it.counts())
        print(f"{name}: cycles {s.cycles}  exact recounts {s.recounts}  rolled-back requests "
              f"{s.fallback_requests} of {n_edit}")
    # both executions end in the same state (sequential semantics either way)
    assert runs["bound"][4] == runs["mlp"][4] and runs["bound"][5] == runs["mlp"][5]
    # the network admits more per cycle and needs fewer exact counts
    assert runs["mlp"][0] < runs["bound"][0]
    assert runs["mlp"][1] <= runs["bound"][1] + runs["mlp"][2]
    assert runs["mlp"][2] <= 0.05 * runs["mlp"][3]
