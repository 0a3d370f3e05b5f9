"""Learned node-growth predictor for admission by memory budget.

An FP32 MLP 16 -> 64 -> 64 -> 1 (ReLU) predicts log(1 + new interner nodes) of an
edit request from 16 features of its geometry and of the trees it touches.  It is
trained on the GPU by the library's own kernels (vx_mlp_gradient + vx_mlp_adam) on
(features, exact growth) pairs that the serving loop logs with dry runs
(Server(log_exact=True)).  The server reserves ceil(prediction * margin) + 2 nodes
per request (never more than the request's hard bound); an apply whose requests
need more is rolled back by the library and the exact dry-run count takes over.

Features (index: meaning):
  0   log1p(voxels edited)
  1   log1p(trees touched)
  2-7 log1p(node positions touched at height 1..6, summed over trees; 0 above depth)
  8   fraction of the touched 2x2x2 bricks written completely (uniform bricks)
  9   fraction of the touched trees that are empty
  10  fraction of the touched trees that are one leaf (filled)
  11  log1p(mean edit requests applied so far to the touched trees)
  12  1 if the request writes the default value (carving)
  13  tree depth / 8
  14  1 for a whole-tree fill
  15  log1p(hard bound of the request's new nodes)
"""
import ctypes as C
import math

import numpy as np
import torch

from ._lib import check, lib

N_FEATURES = 16


def features(world, r):
    f = np.zeros(N_FEATURES, dtype=np.float32)
    pos, cov = r.geom
    f[0] = math.log1p(r.edits)
    f[1] = math.log1p(len(r.trees))
    for h in range(min(len(pos), 6)):
        f[2 + h] = math.log1p(int(pos[h]))
    f[8] = cov
    kinds = world.kind[r.trees]
    f[9] = float(np.mean(kinds == 0))
    f[10] = float(np.mean(kinds == 1))
    f[11] = math.log1p(float(np.mean(world.history[r.trees])))
    if r.kind == "sphere":
        v = next(iter(r.payload.values()))[1]
        f[12] = 1.0 if len(v) and int(v[0]) == 0 else 0.0
    else:
        f[12] = 1.0 if r.payload == 0 else 0.0
    f[13] = world.depth / 8.0
    f[14] = 1.0 if r.kind == "fill" else 0.0
    f[15] = math.log1p(r.bound)
    return f


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class GrowthMLP:
    """The network's parameters and Adam state on the device, driven by the library."""

    def __init__(self, device=None, seed=0, params=None):
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.n_params = lib().vx_mlp_param_count()
        if params is None:
            params = self.init_params(seed)
        self.p = torch.from_numpy(np.asarray(params, dtype=np.float32).copy()).to(self.device)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.step = 0

    @staticmethod
    def init_params(seed):
        """He-uniform weights, zero biases."""
        rng = np.random.default_rng(seed)
        w1 = rng.uniform(-1, 1, (64, 16)) * math.sqrt(6.0 / 16)
        w2 = rng.uniform(-1, 1, (64, 64)) * math.sqrt(6.0 / 64)
        w3 = rng.uniform(-1, 1, 64) * math.sqrt(3.0 / 64)
        return np.concatenate([w1.ravel(), np.zeros(64), w2.ravel(), np.zeros(64), w3, np.zeros(1)]).astype(
            np.float32)

    def forward(self, x):
        """x: device float32 [n, 16] -> device float32 [n]."""
        x = x.contiguous()
        n = x.shape[0]
        y = torch.empty(n, dtype=torch.float32, device=self.device)
        check(lib().vx_mlp_forward(_stream(), self.p.data_ptr(), x.data_ptr(), n, y.data_ptr()), "mlp_forward")
        return y

    def gradient(self, x, t):
        """(grad [n_params], y [n], loss) on the device for L = mean((y - t)^2)."""
        x = x.contiguous()
        t = t.contiguous()
        n = x.shape[0]
        y = torch.empty(n, dtype=torch.float32, device=self.device)
        ws = torch.empty(n * 257, dtype=torch.float32, device=self.device)
        g = torch.empty(self.n_params, dtype=torch.float32, device=self.device)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        check(lib().vx_mlp_gradient(_stream(), self.p.data_ptr(), x.data_ptr(), t.data_ptr(), n, y.data_ptr(),
                                    ws.data_ptr(), g.data_ptr(), loss.data_ptr()), "mlp_gradient")
        return g, y, loss

    def adam_step(self, g, lr):
        self.step += 1
        check(lib().vx_mlp_adam(_stream(), self.p.data_ptr(), g.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
                                self.step, C.c_float(lr)), "mlp_adam")

    def train(self, x, t, epochs=200, batch=512, lr=3e-3, seed=0):
        """Minibatch Adam on device tensors x [n,16], t [n]; returns per-epoch mean loss."""
        n = x.shape[0]
        gen = torch.Generator(device="cpu").manual_seed(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen).to(self.device)
            tot = torch.zeros(1, dtype=torch.float32, device=self.device)
            nb = 0
            for s in range(0, n, batch):
                idx = perm[s:s + batch]
                g, _, loss = self.gradient(x[idx], t[idx])
                self.adam_step(g, lr)
                tot += loss
                nb += 1
            hist.append(float(tot.item()) / nb)
        return hist


class Predictor:
    """Features -> standardized -> network -> expm1: predicted new nodes per request."""

    def __init__(self, model, mu, sd):
        self.model = model
        self.mu = np.asarray(mu, dtype=np.float32)
        self.sd = np.asarray(sd, dtype=np.float32)

    def predict_features(self, X):
        z = (np.asarray(X, dtype=np.float32) - self.mu) / self.sd
        y = self.model.forward(torch.from_numpy(z).to(self.model.device))
        return np.maximum(np.expm1(y.cpu().numpy().astype(np.float64)), 0.0)

    def predict(self, world, reqs):
        return self.predict_features(np.stack([features(world, r) for r in reqs]))

    @classmethod
    def fit(cls, X, growth, epochs=300, batch=512, lr=3e-3, seed=0, device=None):
        X = np.asarray(X, dtype=np.float32)
        mu = X.mean(axis=0)
        sd = X.std(axis=0) + 1e-3
        model = GrowthMLP(device=device, seed=seed)
        xt = torch.from_numpy((X - mu) / sd).to(model.device)
        tt = torch.from_numpy(np.log1p(np.asarray(growth, dtype=np.float64)).astype(np.float32)).to(model.device)
        hist = model.train(xt, tt, epochs=epochs, batch=batch, lr=lr, seed=seed)
        return cls(model, mu, sd), hist

    def save(self, path):
        np.savez(path, params=self.model.p.cpu().numpy(), mu=self.mu, sd=self.sd)

    @classmethod
    def load(cls, path, device=None):
        d = np.load(path)
        return cls(GrowthMLP(device=device, params=d["params"]), d["mu"], d["sd"])
