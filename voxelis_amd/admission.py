"""Admission of batched applies by memory budget, pre-sized by the learned growth predictor.

A batched apply of many trees creates new nodes that the interner must hold.  The exact
count needs a dry run of the whole apply (a second pass over every kernel); the hard
bound (every touched node position at every height is new) is safe but reserves about
ten times what an apply really creates.  Here the FP32 MLP of predictor.py (the same
16 -> 64 -> 64 -> 1 network, trained on the GPU by the library's own kernels) predicts
log(1 + new branch nodes) per tree from 16 features of the tree's batch and state; the
apply runs with node_limit = ceil(margin * sum of predictions) + the new-leaf allowance,
and when the library rolls it back with VX_E_BUDGET (an under-prediction) the exact
dry-run count is taken and the apply runs again under it (the fallback recount).

Features of (batch, tree) (index: meaning):
  0     log1p(voxels the batch sets)
  1     log1p(voxels the batch clears)
  2     MaxDepth
  3-8   log1p(node positions the batch touches at height 1..6)
  9     fraction of touched bricks written completely
  10    fraction of touched bricks written completely with one value (they collapse)
  11    log1p(distinct touched bricks: set mask, clear mask and eight values)
  12    1 if the tree is empty before the apply
  13    log1p(new branch nodes the tree's previous apply created), 0 before the first
  14    bytes per voxel value
  15    log1p(hard bound of the tree's new branch nodes: touched positions over all heights)
Features 0-11, 14 and 15 depend on the batch only; 12 and 13 on the tree's state.

Training pairs are (features, exact per-tree growth) from dry runs with
vx_apply_opts.counts, whose per-tree entry counts the new branch nodes first created
for that tree (a node shared by several trees of the apply counts once), so per-tree
targets are exact in sum over an apply.
"""
import math

import numpy as np

from .predictor import N_FEATURES, Predictor

VX_E_BUDGET = -8


def _popcount8(a):
    return np.unpackbits(np.asarray(a, dtype=np.uint8)[:, None], axis=1).sum(axis=1)


def batch_features(masks, values, depth, value_size):
    """The batch-only features (12 and 13 left 0) of one tree's dense Batch buffers
    (masks [P, 2] u8: set, clear bits per brick; values [P, 8])."""
    masks = np.asarray(masks)
    values = np.asarray(values)
    f = np.zeros(N_FEATURES, dtype=np.float32)
    sm, cm = masks[:, 0], masks[:, 1]
    touched = np.flatnonzero((sm | cm) != 0)  # brick positions in Morton order, ascending
    f[0] = math.log1p(int(_popcount8(sm).sum()))
    f[1] = math.log1p(int(_popcount8(cm).sum()))
    f[2] = depth
    bound = 0
    pos = touched.astype(np.int64)
    for h in range(1, depth + 1):
        if len(pos) == 0:
            break
        p = pos >> (3 * (h - 1))
        n = int(np.count_nonzero(np.diff(p))) + 1
        bound += n
        if h <= 6:
            f[2 + h] = math.log1p(n)
    if len(touched):
        full = sm[touched] == 0xFF
        vt = values[touched]
        uni = full & (vt == vt[:, :1]).all(axis=1)
        f[9] = float(full.mean())
        f[10] = float(uni.mean())
        rows = np.ascontiguousarray(np.concatenate([masks[touched], vt.view(np.uint8).reshape(len(touched), -1)],
                                                   axis=1))
        f[11] = math.log1p(len(np.unique(rows.view(np.dtype((np.void, rows.shape[1]))))))
    f[14] = value_size
    f[15] = math.log1p(bound)
    return f


def with_state(static, empty, last_growth):
    """Feature rows [T, 16]: the batch features with the trees' state filled in."""
    X = np.array(static, dtype=np.float32, copy=True)
    X[:, 12] = np.asarray(empty, dtype=np.float32)
    X[:, 13] = np.log1p(np.asarray(last_growth, dtype=np.float64)).astype(np.float32)
    return X


def leaf_allowance(value_size, distinct_values=None):
    """New leaves one apply can create: every value of a u8/u16 interner, else the
    batch's distinct values (value_size 0: none, for tests that force an under-prediction)."""
    if value_size == 0:
        return 0
    if value_size <= 2:
        return 1 << (8 * value_size)
    return int(distinct_values)


class BatchAdmission:
    """node_limit from the predictor; the exact dry-run count when the library reports
    an under-prediction (VX_E_BUDGET).  Tracks fallbacks and the predictions' error."""

    def __init__(self, predictor, margin, value_size):
        self.pred = predictor
        self.margin = float(margin)
        self.value_size = value_size
        self.applies = 0
        self.fallbacks = 0
        self.errors = []      # |predicted - exact| new branch nodes per apply
        self.reserved = []    # node_limit per apply
        self.used = []        # new nodes the apply created

    def predict(self, X):
        """Per-tree predicted new branch nodes (never above the hard bound)."""
        return self.pred.predict_features(X)

    def node_limit(self, X, distinct_values=None):
        p = self.predict(X)
        return int(math.ceil(self.margin * float(p.sum()))) + leaf_allowance(self.value_size, distinct_values), p

    def apply(self, forest, views, X, distinct_values=None):
        """One admitted apply of views on forest; returns per-tree new branch nodes."""
        import torch
        limit, p = self.node_limit(X, distinct_values)
        forest.apply_views(views, node_limit=limit, counts=True)
        forest.it.sync()
        nb, nl, per = forest.counts_host()
        if nb is None:
            st = forest.status.cpu().numpy()
            assert (st == VX_E_BUDGET).all(), np.unique(st)
            self.fallbacks += 1
            forest.apply_views(views, dry_run=True, counts=True)
            forest.it.sync()
            eb, el, _ = forest.counts_host()
            limit = eb + el
            forest.apply_views(views, node_limit=limit, counts=True)
            forest.it.sync()
            nb, nl, per = forest.counts_host()
            assert nb is not None, "the exact count must admit the apply"
        torch.cuda.current_stream().synchronize()
        self.applies += 1
        self.errors.append(abs(float(p.sum()) - nb))
        self.reserved.append(limit)
        self.used.append(nb + nl)
        self.last = (nb, nl)
        return per

    def report(self):
        used = np.maximum(np.asarray(self.used, dtype=np.float64), 1.0)
        return {"applies": self.applies, "fallbacks": self.fallbacks,
                "fallback_rate": self.fallbacks / max(self.applies, 1), "margin": self.margin,
                "mae_new_branches_per_apply": float(np.mean(self.errors)) if self.errors else None,
                "mean_new_nodes_per_apply": float(np.mean(self.used)) if self.used else None,
                "reserved_over_used": float(np.mean(np.asarray(self.reserved) / used)) if self.used else None}


def fit(X, y, groups, holdout=0.2, quantile=0.995, epochs=30, batch=8192, lr=3e-3, seed=0):
    """Train on (X [n,16], per-tree growth y [n]) logged in applies (groups [n]: the apply
    of each pair); the last `holdout` of the applies measure the error and calibrate the
    margin on apply totals.  Returns (predictor, margin, report)."""
    X = np.asarray(X, dtype=np.float32)
    y = np.asarray(y, dtype=np.float64)
    groups = np.asarray(groups)
    ug = np.unique(groups)
    cut = ug[int((1.0 - holdout) * len(ug))] if len(ug) > 1 else ug[-1] + 1
    tr, te = groups < cut, groups >= cut
    pred, hist = Predictor.fit(X[tr], y[tr], epochs=epochs, batch=batch, lr=lr, seed=seed)
    p = pred.predict_features(X[te])
    bound = np.expm1(X[te, 15].astype(np.float64))
    g = groups[te]
    gi = np.unique(g, return_inverse=True)[1]
    tot_p = np.bincount(gi, weights=p)
    tot_y = np.bincount(gi, weights=y[te])
    tot_b = np.bincount(gi, weights=bound)
    need = np.where(tot_y <= 0, 0.0, tot_y / np.maximum(tot_p, 1e-9))
    margin = float(np.clip(np.quantile(need, quantile) * 1.02, 1.0, 8.0)) if len(need) else 1.5
    report = {"pairs": int(len(y)), "train_pairs": int(tr.sum()), "heldout_pairs": int(te.sum()),
              "applies": int(len(ug)), "heldout_applies": int(len(tot_y)),
              "mae_new_branches_per_tree": float(np.mean(np.abs(p - y[te]))) if te.any() else None,
              "mean_new_branches_per_tree": float(np.mean(y[te])) if te.any() else None,
              "mae_new_branches_per_apply": float(np.mean(np.abs(tot_p - tot_y))) if len(tot_y) else None,
              "mean_new_branches_per_apply": float(np.mean(tot_y)) if len(tot_y) else None,
              "hard_bound_over_exact": float(np.sum(tot_b) / max(np.sum(tot_y), 1.0)),
              "prediction_over_exact": float(np.sum(tot_p) / max(np.sum(tot_y), 1.0)),
              "margin": margin,
              "heldout_under_rate": float(np.mean(np.ceil(tot_p * margin) < tot_y)) if len(tot_y) else None,
              "train_loss": hist[-1] if hist else None}
    return pred, margin, report
