"""Learned node-growth predictor for admission by memory budget.

An FP32 MLP 16 -> 64 -> 64 -> 1 (ReLU) predicts log(1 + new branch nodes) that an
edit request creates in one tree, from 16 features of the request's edits in that
tree and of the tree's current state; a request's prediction is the sum over its
trees plus its one new leaf.  It is trained on the GPU by the library's own kernels
(vx_mlp_gradient + vx_mlp_adam) on (features, exact growth) pairs that the serving
loop logs with dry runs (Server(log_exact=True)).  The server reserves
ceil(prediction * margin) + 2 nodes per request (never more than the request's hard
bound); an apply whose requests need more is rolled back by the library and the
exact dry-run count takes over.  The margin is calibrated on held-out requests.

Features of (request, tree) (index: meaning):
  0   log1p(voxels the request edits in the tree)
  1   log1p(trees the request touches)
  2-6 log1p(node positions touched at height 1..5)
  7   log1p(node positions touched at heights 6 and 7)
  8   fraction of the touched 2x2x2 bricks written completely (uniform bricks)
  9   fraction of sampled edited voxels that are non-empty now (batched get)
  10  fraction of sampled edited voxels that already hold the written value
  11  log1p(edit requests applied so far to the tree)
  12  1 if the request writes the default value (carving)
  13  fraction of the partly written bricks that lie inside a leaf above brick level
  14  1 if a partly written brick is a leaf at brick level now (the reference panics, H1)
  15  log1p(hard bound of the tree's new branch nodes)
Features 9, 10, 13 and 14 come from one batched read of the touched trees
(vx_trees_get_on with stop depths) at admission time.
"""
import ctypes as C
import math

import numpy as np
import torch

from ._lib import check, lib

N_FEATURES = 16


def features(world, r):
    """[trees touched, 16] feature rows of an edit request (sphere); fills have none."""
    if r.kind != "sphere":
        return np.zeros((0, N_FEATURES), dtype=np.float32)
    edits, pos, cov = r.geom
    k = len(edits)
    trees = np.asarray(sorted(r.payload), dtype=np.int64)
    f = np.zeros((k, N_FEATURES), dtype=np.float32)
    f[:, 0] = np.log1p(edits)
    f[:, 1] = math.log1p(k)
    d = pos.shape[1]
    f[:, 2:2 + min(d, 5)] = np.log1p(pos[:, :5])
    if d > 5:
        f[:, 7] = np.log1p(pos[:, 5:].sum(axis=1))
    f[:, 8] = cov
    if r.occ is not None:
        f[:, 9], f[:, 10], f[:, 14], f[:, 13] = r.occ
    f[:, 11] = np.log1p(world.history[trees])
    f[:, 12] = 1.0 if r.sample[2] == 0 else 0.0
    f[:, 15] = np.log1p(pos.sum(axis=1))
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

    def predict_features(self, X, bound=None):
        """Predicted new nodes; features outside the training range are clipped (at 8
        standard deviations), and a prediction never exceeds the hard bound (feature 15)."""
        X = np.asarray(X, dtype=np.float32)
        z = np.clip((X - self.mu) / self.sd, -8.0, 8.0).astype(np.float32)
        y = self.model.forward(torch.from_numpy(z).to(self.model.device)).cpu().numpy().astype(np.float64)
        cap = X[:, 15].astype(np.float64) if bound is None else np.log1p(np.asarray(bound, dtype=np.float64))
        return np.expm1(np.clip(y, 0.0, cap))

    def predict(self, world, reqs):
        """Predicted new nodes per request: per-tree branch predictions summed, plus the
        request's new leaf; a fill creates at most its leaf."""
        rows = [features(world, r) for r in reqs]
        n = [len(x) for x in rows]
        out = np.ones(len(reqs), dtype=np.float64)
        if sum(n):
            per = self.predict_features(np.concatenate(rows))
            o = 0
            for i, k in enumerate(n):
                out[i] += per[o:o + k].sum()
                o += k
        return np.minimum(out, [r.bound for r in reqs])

    @staticmethod
    def calibrate_margin(pred, exact, quantile=0.995, lo=1.0, hi=8.0):
        """Smallest margin m (in [lo, hi]) with exact <= ceil(pred * m) + 2 for `quantile`
        of the pairs."""
        pred = np.asarray(pred, dtype=np.float64)
        exact = np.asarray(exact, dtype=np.float64)
        need = np.where(exact <= 2, 0.0, (exact - 2) / np.maximum(pred, 1e-9))
        return float(np.clip(np.quantile(need, quantile), lo, hi))

    @classmethod
    def fit(cls, X, growth, epochs=300, batch=512, lr=3e-3, seed=0, device=None):
        X = np.asarray(X, dtype=np.float32)
        mu = X.mean(axis=0)
        sd = np.maximum(X.std(axis=0), 0.05)  # a feature constant in training stays near 0
        model = GrowthMLP(device=device, seed=seed)
        xt = torch.from_numpy((X - mu) / sd).to(model.device)
        tt = torch.from_numpy(np.log1p(np.asarray(growth, dtype=np.float64)).astype(np.float32)).to(model.device)
        hist = model.train(xt, tt, epochs=epochs, batch=batch, lr=lr, seed=seed)
        return cls(model, mu, sd), hist

    def predict_rows(self, rows_per_request):
        """Request-level predictions from per-request feature rows (sum + one leaf)."""
        n = [len(x) for x in rows_per_request]
        out = np.ones(len(n), dtype=np.float64)
        if sum(n):
            per = self.predict_features(np.concatenate(rows_per_request))
            o = 0
            for i, k in enumerate(n):
                out[i] += per[o:o + k].sum()
                o += k
        return out

    @classmethod
    def fit_from_log(cls, log, holdout=0.3, quantile=0.995, epochs=300, batch=512, lr=3e-3, seed=0):
        """Train on a serving log [(rows [k,16], per-tree growth [k], exact request growth)],
        holding out the last `holdout` of the requests to measure the error and to
        calibrate the margin.  Returns (predictor, margin, report)."""
        log = [e for e in log if len(e[0])]
        cut = int((1.0 - holdout) * len(log))
        train, test = log[:cut], log[cut:]
        X = np.concatenate([e[0] for e in train])
        y = np.concatenate([e[1] for e in train])
        pred, hist = cls.fit(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
        p = pred.predict_rows([e[0] for e in test])
        exact = np.array([e[2] for e in test], dtype=np.float64)
        bound = np.array([np.expm1(e[0][:, 15].astype(np.float64)).sum() + 1 for e in test])
        margin = cls.calibrate_margin(p, exact, quantile=quantile)
        report = {"pairs": int(len(y)), "requests": len(log), "heldout_requests": len(test),
                  "mae_nodes": float(np.mean(np.abs(p - exact))),
                  "hard_bound_mae_nodes": float(np.mean(np.abs(bound - exact))),
                  "mean_growth": float(exact.mean()) if len(exact) else 0.0, "margin": margin,
                  "heldout_under_rate": float(np.mean(np.ceil(p * margin) + 2 < exact)) if len(exact) else 0.0,
                  "train_loss": hist[-1] if hist else None}
        return pred, margin, report

    def save(self, path):
        np.savez(path, params=self.model.p.cpu().numpy(), mu=self.mu, sd=self.sd)

    @classmethod
    def load(cls, path, device=None):
        d = np.load(path)
        return cls(GrowthMLP(device=device, params=d["params"]), d["mu"], d["sd"])
