"""Serving loop over a world of trees resident in HBM (continuous batching).

Requests arrive over time (Poisson arrivals): get batches, sphere edits and whole-
tree fills.  The server keeps admitting arrived requests into launch cycles while
earlier cycles are still on the GPU:

- reads run on their own stream, concurrently with applies of other trees;
- a cycle's edits (one vx_trees_apply_edits over every admitted sphere, plus one
  vx_trees_fill) run on the interner's stream after the cycle's reads;
- a request may overtake an earlier one that is still waiting only when the two
  commute (no tree that one edits and the other reads or edits), so every result
  equals executing the requests one at a time in arrival order;
- a request never depends on an edit that is not confirmed yet (its cycle has not
  finished), because an apply can be rolled back.

Admission by memory budget: every edit request holds a reservation of interner
nodes; a cycle is admitted only while the reservations fit the free capacity, and
its apply runs with node_limit = the sum of its reservations.  A reservation is
the request's hard bound (every node position its edits touch, plus its new
leaf), or, with a predictor, the predicted growth times a margin (capped by the
bound).  An apply that exceeds its limit is rolled back by the library (state
unchanged, status VX_E_BUDGET); its requests then get their exact growth from a
dry run (the exact two-pass count) and are admitted again with that reservation.
"""
import collections
import math
import time

import numpy as np
import torch

from . import api

VX_E_OOM = -2
VX_E_BUDGET = -8


class World:
    """tx * ty * tz trees of one depth sharing an interner, roots resident on the device."""

    def __init__(self, value_size, budget, shape, depth=5, load_factor=0.5):
        self.it = api.VoxInterner(value_size, budget, load_factor)
        self.shape = shape
        self.depth = depth
        self.side = 1 << depth
        self.n = int(np.prod(shape))
        self.forest = api.Forest(self.it, [depth] * self.n)
        self.dev = self.it.device
        # host-side facts about every tree (features of the growth predictor)
        self.kind = np.zeros(self.n, dtype=np.int8)   # 0 empty, 1 single leaf (filled), 2 other
        self.history = np.zeros(self.n, dtype=np.int64)  # edit requests applied so far
        self.version = np.zeros(self.n, dtype=np.int64)  # bumped whenever a tree changes

    def tree_index(self, tx, ty, tz):
        sx, sy, sz = self.shape
        return (tz * sy + ty) * sx + tx

    def split_world(self, xyz):
        """World voxel coordinates [n,3] -> (tree index, local coordinates)."""
        t = xyz // self.side
        local = (xyz % self.side).astype(np.int32)
        return self.tree_index(t[:, 0], t[:, 1], t[:, 2]).astype(np.int64), local

    def sub(self, trees):
        idx = torch.from_numpy(np.asarray(trees, dtype=np.int64)).to(self.dev)
        sub = api.Forest(self.it, [self.depth] * len(trees))
        sub.roots.copy_(self.forest.roots[idx])
        return sub, idx

    def subforest_apply_edits(self, trees, offsets, xyz_dev, val_dev, **opts):
        """vx_trees_apply_edits on a subset of trees (roots gathered and scattered back);
        the caller reads sub.status / sub.counts."""
        sub, idx = self.sub(trees)
        sub.apply_edits(offsets, xyz_dev, val_dev, sync=False, **opts)
        if not opts.get("dry_run"):
            self.forest.roots[idx] = sub.roots
        return sub

    def subforest_fill(self, trees, values):
        from ._lib import check, lib
        sub, idx = self.sub(trees)
        vals = torch.from_numpy(np.asarray(values, dtype=np.uint64).view(np.int64)).to(self.dev)
        check(lib().vx_trees_fill(self.it.h, len(trees), sub.roots.data_ptr(), vals.data_ptr()), "fill")
        self.forest.roots[idx] = sub.roots
        return sub


class Request:
    __slots__ = ("rid", "kind", "arrival", "trees", "payload", "edits", "done", "result", "bound",
                 "pred", "reserve", "exact", "growth", "geom", "feat", "occ", "sample", "_sel")

    def __init__(self, rid, kind, arrival):
        self.rid = rid
        self.kind = kind
        self.arrival = arrival
        self.trees = None
        self.payload = None
        self.edits = 0
        self.done = None
        self.result = None
        self.bound = 0      # hard upper bound of new nodes
        self.pred = None    # predicted new nodes
        self.reserve = 0
        self.exact = None   # exact new nodes (dry run), when measured
        self.growth = None  # nodes its cycle created, shared out over its requests
        self.geom = None    # per-level touched positions, covered-brick fraction
        self.feat = None    # version stamp of the trees when last predicted
        self._sel = None
        self.occ = None     # per touched tree: sampled non-empty / equal fractions, H1 flag, split fraction
        self.sample = None  # sampled edits and partly written bricks, with per-tree segments



def touched_positions(local, depth):
    """Node positions at heights 1..depth that contain an edit of one tree, bounded by
    the edits' bounding box (exact for boxes, an upper bound otherwise)."""
    lo = local.min(axis=0).astype(np.int64)
    hi = local.max(axis=0).astype(np.int64)
    out = np.zeros(depth, dtype=np.int64)
    for h in range(1, depth + 1):
        out[h - 1] = int(np.prod((hi >> h) - (lo >> h) + 1))
    return out


def covered_bricks(local, with_partial=False):
    """(bricks touched, bricks whose 8 voxels are all written) of one tree's edits, and
    with_partial: the first voxel of every partly written brick [m, 3]."""
    b = local.astype(np.int64) >> 1
    lo = b.min(axis=0)
    ext = b.max(axis=0) - lo + 1
    key = ((b[:, 0] - lo[0]) * ext[1] + (b[:, 1] - lo[1])) * ext[2] + (b[:, 2] - lo[2])
    cnt = np.bincount(key, minlength=int(np.prod(ext)))
    res = (int((cnt > 0).sum()), int((cnt == 8).sum()))
    if not with_partial:
        return res
    k = np.flatnonzero((cnt > 0) & (cnt < 8))
    bz = k % ext[2]
    by = (k // ext[2]) % ext[1]
    bx = k // (ext[2] * ext[1])
    partial = (np.stack([bx, by, bz], axis=1) + lo) * 2
    return res + (partial.astype(np.int32),)


def sphere_edits(world, center, radius, value):
    """Per-tree edit lists of a ball of `value`: {tree: (local xyz [k,3], values [k])}."""
    r = int(radius)
    lo = np.maximum(center - r, 0)
    hi = np.minimum(center + r + 1, np.array(world.shape) * world.side)
    if np.any(hi <= lo):
        return {}
    xs = np.arange(lo[0], hi[0])
    ys = np.arange(lo[1], hi[1])
    zs = np.arange(lo[2], hi[2])
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    inside = (X - center[0]) ** 2 + (Y - center[1]) ** 2 + (Z - center[2]) ** 2 <= radius * radius
    pts = np.stack([X[inside], Y[inside], Z[inside]], axis=1)
    if len(pts) == 0:
        return {}
    trees, local = world.split_world(pts)
    order = np.argsort(trees, kind="stable")
    trees, local = trees[order], local[order]
    out = {}
    cut = np.flatnonzero(np.diff(trees)) + 1
    for seg_t, seg_l in zip(np.split(trees, cut), np.split(local, cut)):
        out[int(seg_t[0])] = (seg_l, np.full(len(seg_l), value, dtype=np.uint16))
    return out


SAMPLES_PER_TREE = 64
PARTIAL_PER_TREE = 2048


def describe(world, r):
    """Fill r.bound and r.geom for an edit request (host geometry only).

    r.geom, per touched tree in sorted tree order: (edits [k], touched node positions
    [k, depth] at heights 1..depth, fraction of touched bricks written completely [k]);
    r.sample: up to SAMPLES_PER_TREE evenly strided edited positions of every tree."""
    if r.kind == "fill":
        r.bound = 1
        r.geom = None
        return
    items = sorted(r.payload.items())
    k = len(items)
    edits = np.zeros(k, dtype=np.int64)
    pos = np.zeros((k, world.depth), dtype=np.int64)
    cov = np.zeros(k, dtype=np.float64)
    s_tree, s_local, s_seg = [], [], [0]
    p_tree, p_local, p_seg = [], [], [0]
    for i, (t, (local, v)) in enumerate(items):
        edits[i] = len(v)
        pos[i] = touched_positions(local, world.depth)
        a, b, partial = covered_bricks(local, with_partial=True)
        cov[i] = b / max(a, 1)
        step = max(1, len(v) // SAMPLES_PER_TREE)
        pick = local[::step][:SAMPLES_PER_TREE]
        s_tree.append(np.full(len(pick), t, dtype=np.uint32))
        s_local.append(pick)
        s_seg.append(s_seg[-1] + len(pick))
        pstep = max(1, -(-len(partial) // PARTIAL_PER_TREE))
        partial = partial[::pstep]
        p_tree.append(np.full(len(partial), t, dtype=np.uint32))
        p_local.append(partial)
        p_seg.append(p_seg[-1] + len(partial))
    r.bound = int(pos.sum()) + 1  # + the request's one new leaf value
    r.geom = (edits, pos, cov)
    value = int(items[0][1][1][0])
    # sampled edited voxels (occupancy) and partly written bricks (one voxel each)
    r.sample = (np.concatenate(s_tree), np.concatenate(s_local).astype(np.int32), value, np.array(s_seg),
                np.concatenate(p_tree), np.concatenate(p_local).astype(np.int32), np.array(p_seg))


def make_requests(world, n, rate, seed, mix=(0.70, 0.25, 0.05), max_gets=65536, carve=False,
                  radius=(2.0, 40.0)):
    """Poisson arrivals at `rate` req/s; kinds get / sphere / fill per `mix`."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    world_size = np.array(world.shape) * world.side
    reqs = []
    for i in range(n):
        u = rng.random()
        if u < mix[0]:
            r = Request(i, "get", t[i])
            q = int(rng.integers(1, max_gets + 1))
            pts = (rng.random((q, 3)) * world_size).astype(np.int64)
            trees, local = world.split_world(pts)
            r.trees = np.unique(trees)
            r.payload = (trees.astype(np.uint32), local)
        elif u < mix[0] + mix[1]:
            r = Request(i, "sphere", t[i])
            center = (rng.random(3) * world_size).astype(np.int64)
            rad = float(rng.uniform(*radius))
            value = 0 if (carve and rng.random() < 0.5) else int(rng.integers(1, 65536))
            per_tree = sphere_edits(world, center, rad, value)
            r.trees = np.array(sorted(per_tree), dtype=np.int64)
            r.payload = per_tree
            r.edits = int(sum(len(v[1]) for v in per_tree.values()))
            if not per_tree:
                r.kind = "noop"
                r.trees = np.zeros(0, dtype=np.int64)
        else:
            r = Request(i, "fill", t[i])
            tree = int(rng.integers(0, world.n))
            value = 0 if rng.random() < 0.5 else int(rng.integers(1, 65536))
            r.trees = np.array([tree], dtype=np.int64)
            r.payload = value
            r.edits = world.side ** 3
        if r.kind != "get":
            describe(world, r)
        reqs.append(r)
    return reqs


def _selector(r, n):
    """The request's trees as an index array, or as a bool mask when it touches many
    (a mask tests and merges in one vector pass)."""
    if r._sel is None:
        r._sel = r.trees
        if len(r.trees) * 32 > n:
            m = np.zeros(n, dtype=bool)
            m[r.trees] = True
            r._sel = m
    return r._sel


def _hit(mask, sel):
    return bool(np.logical_and(mask, sel).any()) if sel.dtype == bool else bool(mask[sel].any())


def _mark(mask, sel):
    if sel.dtype == bool:
        np.logical_or(mask, sel, out=mask)
    else:
        mask[sel] = True


class Cycle:
    __slots__ = ("gets", "spheres", "fills", "reserve", "ev_r", "ev_e", "sub", "out", "pres", "edit_trees",
                 "counts")

    def __init__(self):
        self.gets, self.spheres, self.fills = [], [], []
        self.reserve = 0
        self.ev_r = self.ev_e = self.sub = self.out = self.pres = None
        self.edit_trees = None  # bool mask over the world's trees
        self.counts = None


class Server:
    """Continuous-batching executor of requests against one World.

    predictor: None (reserve the hard bound) or an object with
    predict(world, requests) -> predicted new nodes per request.
    """

    def __init__(self, world, predictor=None, margin=1.5, max_inflight=2, max_cycle_edits=1 << 21,
                 max_cycle_gets=1 << 20, log_exact=False, max_scan=1024):
        self.w = world
        self.predictor = predictor
        self.margin = margin
        self.max_inflight = max_inflight
        self.max_cycle_edits = max_cycle_edits
        self.max_cycle_gets = max_cycle_gets
        self.max_scan = max_scan  # requests examined per admission pass
        self.log_exact = log_exact  # dry-run every edit request at admission (training data)
        self.R = torch.cuda.Stream(device=world.dev)
        self.cycles = 0
        self.fallbacks = 0        # applies rolled back for exceeding their reservation
        self.fallback_requests = 0  # edit requests of those applies (counted exactly afterwards)
        self.recounts = 0         # exact dry-run counts taken
        self.admitted_edits = 0
        self.log = []             # (features, exact growth) of measured requests
        self.capacity = world.it.capacity()
        self.free = None

    # ---------------------------------------------------------------- budget

    def _refresh_free(self):
        self.free = self.capacity - self.w.it.alive_nodes()

    def _exact(self, r):
        """The exact two-pass count: a dry run of the request alone on the current state."""
        self.recounts += 1
        if r.kind == "fill":
            r.exact = 1
            return r.exact
        offs, xyz_t, val_t, trees = self._pack([r])
        sub = self.w.subforest_apply_edits(trees, offs, xyz_t, val_t, dry_run=True, counts=True)
        b, l, per_tree = sub.counts_host()
        r.exact = b + l
        if self.log_exact:
            from .predictor import features
            self._occupancy([r])
            # one pair per touched tree: its new branch nodes (a node shared by two trees
            # of the request counts for one of them)
            self.log.append((features(self.w, r), np.asarray(per_tree, dtype=np.float64), r.exact))
        return r.exact

    def _occupancy(self, reqs):
        """Sampled reads of the voxels each request edits (one batched get on the read
        stream): fraction non-empty and fraction already equal to the written value.
        Only called for requests whose trees no unconfirmed apply touches."""
        todo = [r for r in reqs if r.kind == "sphere" and r.sample is not None]
        for r in reqs:
            if r.kind != "sphere":
                r.occ = None
        if not todo:
            return
        chunks = np.concatenate([np.concatenate([r.sample[0], r.sample[4]]) for r in todo])
        xyz = np.concatenate([np.concatenate([r.sample[1], r.sample[5]]) for r in todo]).astype(np.int32)
        dev = self.w.dev
        with torch.cuda.stream(self.R):
            chunk_t = torch.from_numpy(chunks.view(np.int32)).to(dev)
            xyz_t = torch.from_numpy(xyz.reshape(-1)).to(dev)
            out = torch.empty(len(chunks) * self.w.it.value_size, dtype=torch.uint8, device=dev)
            pres = torch.empty(len(chunks), dtype=torch.uint8, device=dev)
            stop = torch.empty(len(chunks), dtype=torch.uint8, device=dev)
            self.w.forest.get_device(chunk_t, xyz_t, out, pres, stream=self.R, stop=stop)
            vals = out.cpu().numpy().view(self.w.it.np_dtype)
            p = pres.cpu().numpy()
            sd = stop.cpu().numpy()
        brick_level = self.w.depth - 1
        o = 0
        for r in todo:
            k = len(r.sample[0])
            v = np.where(p[o:o + k] == 1, vals[o:o + k], 0).astype(np.int64)
            seg = r.sample[3]
            nonempty = np.add.reduceat((v != 0).astype(np.float64), seg[:-1]) / np.maximum(np.diff(seg), 1)
            equal = np.add.reduceat((v == r.sample[2]).astype(np.float64), seg[:-1]) / np.maximum(np.diff(seg), 1)
            o += k
            m = len(r.sample[4])
            st = sd[o:o + m].astype(np.int64)
            o += m
            pseg = r.sample[6]
            at_leaf = (st & 0x80) != 0
            depth = st & 0x7F
            # a partly written brick whose node is a leaf at brick level: the apply of
            # that tree panics (H1); a leaf above brick level is split along the path
            h1_hits = (at_leaf & (depth == brick_level)).astype(np.float64)
            split = (at_leaf & (depth < brick_level)).astype(np.float64)
            n_p = np.maximum(np.diff(pseg), 1)
            if m:
                h1 = np.maximum.reduceat(h1_hits, np.minimum(pseg[:-1], m - 1)) * (np.diff(pseg) > 0)
                sp = np.add.reduceat(split, np.minimum(pseg[:-1], m - 1)) * (np.diff(pseg) > 0) / n_p
            else:
                h1 = np.zeros(len(pseg) - 1)
                sp = np.zeros(len(pseg) - 1)
            if r.sample[2] == 0:
                h1 = np.zeros_like(h1)  # writing the default value records no set bits
            r.occ = (nonempty, equal, h1, sp)  # per touched tree

    def _reservation(self, r):
        if r.exact is not None:
            return r.exact
        if r.pred is None:
            return r.bound
        return int(min(r.bound, math.ceil(r.pred * self.margin) + 2))

    # ---------------------------------------------------------------- admission

    def _admit(self, queue, now, unconfirmed):
        """One cycle from the arrived requests, in arrival order, skipping requests
        that do not commute with an earlier skipped one or depend on unconfirmed edits."""
        cyc = Cycle()
        n = self.w.n
        # tree masks: edited / read by requests skipped in this pass, edited in this cycle
        def_edit = np.zeros(n, dtype=bool)
        def_read = np.zeros(n, dtype=bool)
        cyc.edit_trees = np.zeros(n, dtype=bool)
        no_edit = unconfirmed | cyc.edit_trees  # kept equal to unconfirmed | def_edit | cycle edits
        free = self.free - sum(c.reserve for c in self._inflight)
        n_edits = n_gets = 0
        budget_blocked = False
        taken = []
        scan = 0
        for k, r in enumerate(queue):
            if r.arrival > now or scan >= self.max_scan:
                break
            scan += 1
            if r.kind == "noop":
                r.result = np.zeros(0, dtype=np.int32)
                taken.append(k)
                continue
            ts = _selector(r, n)
            if r.kind == "get":
                if _hit(no_edit, ts) or n_gets + len(r.payload[0]) > self.max_cycle_gets:
                    _mark(def_read, ts)
                    continue
                cyc.gets.append(r)
                n_gets += len(r.payload[0])
                taken.append(k)
                continue
            if _hit(no_edit, ts) or _hit(def_read, ts) or (n_edits and n_edits + r.edits > self.max_cycle_edits):
                _mark(def_edit, ts)
                _mark(no_edit, ts)
                continue
            if self.log_exact:
                self._exact(r)  # training pair: features and exact growth on the current state
            need = self._reservation(r)
            if need > free - cyc.reserve:
                _mark(def_edit, ts)
                _mark(no_edit, ts)
                budget_blocked = True
                continue
            r.reserve = need
            cyc.reserve += need
            (cyc.fills if r.kind == "fill" else cyc.spheres).append(r)
            _mark(cyc.edit_trees, ts)
            _mark(no_edit, ts)
            n_edits += r.edits
            taken.append(k)
        for k in reversed(taken):
            del queue[k]
        return cyc, budget_blocked

    def _pack(self, spheres):
        trees, offs, xyzs, vals = [], [0], [], []
        for r in spheres:
            for t, (local, v) in sorted(r.payload.items()):
                trees.append(t)
                xyzs.append(local)
                vals.append(v)
                offs.append(offs[-1] + len(v))
        xyz_t = api.to_device(np.concatenate(xyzs).astype(np.int32))
        val_t = api.to_device(np.concatenate(vals).astype(np.uint16))
        return offs, xyz_t, val_t, trees

    # ---------------------------------------------------------------- launch / retire

    def _launch(self, cyc):
        dev = self.w.dev
        if cyc.gets:
            chunks = np.concatenate([r.payload[0] for r in cyc.gets])
            xyz = np.concatenate([r.payload[1] for r in cyc.gets]).astype(np.int32)
            with torch.cuda.stream(self.R):
                chunk_t = torch.from_numpy(chunks.view(np.int32)).to(dev, non_blocking=False)
                xyz_t = torch.from_numpy(xyz.reshape(-1)).to(dev, non_blocking=False)
                cyc.out = torch.empty(len(chunks) * self.w.it.value_size, dtype=torch.uint8, device=dev)
                cyc.pres = torch.empty(len(chunks), dtype=torch.uint8, device=dev)
                self.w.forest.get_device(chunk_t, xyz_t, cyc.out, cyc.pres, stream=self.R)
                cyc.ev_r = torch.cuda.Event()
                cyc.ev_r.record(self.R)
        E = torch.cuda.current_stream()
        if cyc.spheres or cyc.fills:
            if cyc.ev_r is not None:
                E.wait_event(cyc.ev_r)  # this cycle's reads see the state before its edits
            elif self._last_r is not None:
                E.wait_event(self._last_r)  # earlier cycles' reads of these trees
            if cyc.spheres:
                offs, xyz_t, val_t, trees = self._pack(cyc.spheres)
                cyc.sub = self.w.subforest_apply_edits(trees, offs, xyz_t, val_t,
                                                       node_limit=max(cyc.reserve, 1), counts=True)
            if cyc.fills:
                self.w.subforest_fill([int(r.trees[0]) for r in cyc.fills], [r.payload for r in cyc.fills])
            cyc.ev_e = torch.cuda.Event()
            cyc.ev_e.record(E)
        if cyc.ev_r is not None:
            self._last_r = cyc.ev_r

    def _retire(self, cyc, t0, lat, requeue):
        if cyc.ev_r is not None:
            cyc.ev_r.synchronize()
            with torch.cuda.stream(self.R):
                vals = cyc.out.cpu().numpy().view(self.w.it.np_dtype)
                p = cyc.pres.cpu().numpy()
            o = 0
            for r in cyc.gets:
                q = len(r.payload[0])
                r.result = np.where(p[o:o + q] == 1, vals[o:o + q], 0)
                o += q
        rolled_back = False
        if cyc.ev_e is not None:
            cyc.ev_e.synchronize()
            if cyc.spheres:
                st = cyc.sub.status.cpu().numpy()
                if (st == VX_E_BUDGET).any() or (st == VX_E_OOM).any():
                    rolled_back = True
                    self.fallbacks += 1
                    self.fallback_requests += len(cyc.spheres)
                    for r in cyc.spheres:
                        r.exact = None
                        r.pred = None
                        requeue.append(r)
                else:
                    b, l, per_tree = cyc.sub.counts_host()
                    self.free -= b + l  # nodes freed by its GC are found at the next refresh
                    o = 0
                    for r in cyc.spheres:
                        k = len(r.payload)
                        r.result = st[o:o + k]
                        r.growth = int(per_tree[o:o + k].sum())
                        o += k
                        self.w.kind[r.trees] = 2
                        self.w.history[r.trees] += 1
                        self.w.version[r.trees] += 1
            self.free -= len(cyc.fills)
            for r in cyc.fills:
                r.result = 1
                self.w.kind[int(r.trees[0])] = 0 if r.payload == 0 else 1
                self.w.history[int(r.trees[0])] += 1
                self.w.version[int(r.trees[0])] += 1
        t_done = time.perf_counter() - t0
        done = cyc.gets + cyc.fills + ([] if rolled_back else cyc.spheres)
        for r in done:
            r.done = t_done
            lat.append(t_done - r.arrival)
        self.admitted_edits += 0 if rolled_back else len(cyc.spheres) + len(cyc.fills)
        return rolled_back

    def _drain(self, t0, lat, requeue):
        while self._inflight:
            self._retire(self._inflight.popleft(), t0, lat, requeue)

    def _ready(self, c):
        return (c.ev_e is None or c.ev_e.query()) and (c.ev_r is None or c.ev_r.query())

    def serve(self, reqs, realtime=True):
        """Run all requests; returns (latencies in s, wall time).  realtime=False treats
        every request as arrived at time 0."""
        queue = list(reqs)
        if not realtime:
            for r in queue:
                r.arrival = 0.0
        self._inflight = collections.deque()
        self._last_r = None
        lat = []
        requeue = []
        self._refresh_free()
        fresh = True  # self.free was read with nothing in flight since the last change
        t0 = time.perf_counter()
        while queue or self._inflight or requeue:
            while self._inflight and self._ready(self._inflight[0]):
                self._retire(self._inflight.popleft(), t0, lat, requeue)
                fresh = False
            if requeue:
                # rolled-back edits: drain, then give them exact counts (arrival order kept)
                self._drain(t0, lat, requeue)
                self._refresh_free()
                for r in requeue:
                    self._exact(r)
                queue = sorted(queue + requeue, key=lambda r: (r.arrival, r.rid))
                requeue = []
                fresh = True
            now = time.perf_counter() - t0 if realtime else 0.0
            budget_blocked = False
            if len(self._inflight) < self.max_inflight:
                unconfirmed = np.zeros(self.w.n, dtype=bool)
                for c in self._inflight:
                    if c.edit_trees is not None:
                        unconfirmed |= c.edit_trees
                if self.predictor is not None:
                    # (re)predict the edits next in line whose trees no unconfirmed apply
                    # touches, with the trees' current features
                    window = []
                    for r in queue[:self.max_scan]:
                        if r.arrival > now or len(window) >= 256:
                            break
                        if (r.kind in ("sphere", "fill") and r.exact is None and not unconfirmed[r.trees].any()
                                and r.feat != int(self.w.version[r.trees].sum())):
                            window.append(r)
                    for r in window:
                        r.feat = int(self.w.version[r.trees].sum())  # trees' state the prediction saw
                    if window:
                        self._occupancy(window)
                        for r, p in zip(window, self.predictor.predict(self.w, window)):
                            r.pred = float(p)
                cyc, budget_blocked = self._admit(queue, now, unconfirmed)
                if cyc.gets or cyc.spheres or cyc.fills:
                    self._launch(cyc)
                    self._inflight.append(cyc)
                    self.cycles += 1
                    fresh = False
                    continue
            if self._inflight:
                self._retire(self._inflight.popleft(), t0, lat, requeue)
                fresh = False
                continue
            if not queue:
                continue  # rolled-back edits may be waiting in requeue
            if queue[0].arrival > now:
                time.sleep(min(queue[0].arrival - now, 0.0005))
                continue
            if not fresh:
                self._refresh_free()  # GC of the applied cycles may have freed nodes
                fresh = True
                continue
            # nothing in flight and the free capacity is current: the first waiting edit
            # does not fit its reservation.  Count it exactly; if even that does not fit,
            # it cannot run (the reference would panic with out-of-memory).
            r = next((x for x in queue if x.arrival <= now and x.kind in ("sphere", "fill")), None)
            if r is None or not budget_blocked:
                raise RuntimeError("admission made no progress")
            if r.exact is None:
                self._exact(r)
                continue
            if r.exact > self.free:
                r.result = np.full(max(len(r.trees), 1), VX_E_OOM, dtype=np.int32)
                r.done = time.perf_counter() - t0
                lat.append(r.done - r.arrival)
                queue.remove(r)
                continue
            raise RuntimeError("admission made no progress")
        self.w.it.sync()
        return np.array(lat), time.perf_counter() - t0
