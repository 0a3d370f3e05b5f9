"""Python mirror of the reference API on the HIP path.

VoxInterner  <- interner/mod.rs VoxInterner<T>   (device node columns + pattern tables)
VoxTree      <- spatial/voxtree.rs VoxTree<T>     (max depth, root id, dirty flag)
Batch        <- core/batch.rs Batch<T>            (dense per-brick edit buffers, host)
Forest       -- many trees of one interner with device-resident roots: the batched
                form of apply_batch / get / fill (one launch sequence per call)

Every compute call goes through libvoxelis_hip.so; device memory is held in torch
tensors (plumbing only).
"""
import ctypes as C

import numpy as np
import torch

from . import morton
from ._lib import APPLY_DRY_RUN, STAT_NAMES, ApplyOpts, BatchView, Tuning, VxError, check, lib

EMPTY = 0
INVALID = (1 << 64) - 1
_DTYPES = {1: (np.uint8, torch.uint8), 2: (np.uint16, torch.uint16), 4: (np.uint32, torch.int32)}  # 4: i32/u32 bit patterns


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("the HIP path needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


_TORCH_SAFE = {np.dtype(np.uint8): np.uint8, np.dtype(np.uint16): np.int16,
               np.dtype(np.uint32): np.int32, np.dtype(np.uint64): np.int64}


def to_device(arr, device=None):
    """Upload a numpy array as raw bytes-compatible torch storage (uint16 goes as int16)."""
    arr = np.ascontiguousarray(arr)
    view = arr.view(_TORCH_SAFE.get(arr.dtype, arr.dtype))
    return torch.from_numpy(view).to(device or _dev())


def _i64(u):
    """uint64 BlockId -> int64 bit pattern for torch storage."""
    return u - (1 << 64) if u >= (1 << 63) else u


def _u64(i):
    return i + (1 << 64) if i < 0 else i


class VoxInterner:
    """VoxInterner::<T>::with_memory_budget (interner/mod.rs:45-155) on the device."""

    def __init__(self, value_size, budget_bytes, load_factor=0.5):
        self.value_size = value_size
        self.np_dtype, self.torch_dtype = _DTYPES[value_size]
        self.device = _dev()
        h = C.c_void_p()
        check(lib().vx_interner_create(value_size, budget_bytes, load_factor, C.byref(h)),
              "vx_interner_create")
        self.h = h
        # order library work with torch's work on the current stream
        check(lib().vx_interner_set_stream(self.h, C.c_void_p(torch.cuda.current_stream().cuda_stream)),
              "vx_interner_set_stream")

    @classmethod
    def with_memory_budget(cls, budget_bytes, value_size=1, load_factor=0.5):
        return cls(value_size, budget_bytes, load_factor)

    def close(self):
        if getattr(self, "h", None):
            lib().vx_interner_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def capacity(self):
        out = C.c_uint64(0)
        check(lib().vx_interner_capacity(self.h, C.byref(out)))
        return out.value

    def counts(self):
        """(leaf_patterns().len(), branch_patterns().len()) incl. the EMPTY pattern."""
        a, b, c = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        check(lib().vx_interner_counts(self.h, C.byref(a), C.byref(b), C.byref(c)), "counts")
        return a.value, b.value

    def alive_nodes(self):
        a, b, c = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        check(lib().vx_interner_counts(self.h, C.byref(a), C.byref(b), C.byref(c)), "counts")
        return c.value

    def get_ref(self, block_id):
        out = C.c_uint32(0)
        check(lib().vx_interner_get_ref(self.h, block_id, C.byref(out)), "get_ref")
        return out.value

    def get_value(self, block_id):
        out = C.c_uint64(0)
        check(lib().vx_interner_get_value(self.h, block_id, C.byref(out)), "get_value")
        return out.value

    def get_children(self, block_id):
        out = (C.c_uint64 * 8)()
        check(lib().vx_interner_get_children(self.h, block_id, out), "get_children")
        return list(out)

    def stats(self):
        out = (C.c_uint64 * len(STAT_NAMES))()
        check(lib().vx_interner_stats(self.h, out, len(STAT_NAMES)), "stats")
        return dict(zip(STAT_NAMES, list(out)))

    def sync(self):
        check(lib().vx_interner_sync(self.h), "sync")

    def set_tuning(self, tile=None, dedup_rounds=None, commit_grid=None, gc_grid=None,
                   probe_blocks_bricks=None, probe_blocks=None, brick_merge_rounds=None, brick_rounds=None):
        """Launch shape of the apply (vx_tuning); None keeps a field."""
        t = self.get_tuning()
        for k, v in (("tile", tile), ("dedup_rounds", dedup_rounds), ("commit_grid", commit_grid),
                     ("gc_grid", gc_grid), ("probe_blocks_bricks", probe_blocks_bricks),
                     ("probe_blocks", probe_blocks), ("brick_merge_rounds", brick_merge_rounds),
                     ("brick_rounds", brick_rounds)):
            if v is not None:
                t[k] = int(v)
        tt = Tuning(*(t[f] for f, _ in Tuning._fields_))
        check(lib().vx_interner_set_tuning(self.h, C.byref(tt)), "set_tuning")

    def get_tuning(self):
        tt = Tuning()
        check(lib().vx_interner_get_tuning(self.h, C.byref(tt)), "get_tuning")
        return {f: getattr(tt, f) for f, _ in Tuning._fields_}

    def rebuild_table(self):
        """Clear the pattern table and re-insert every live branch (synchronizes)."""
        out = C.c_uint64(0)
        check(lib().vx_interner_rebuild_table(self.h, C.byref(out)), "rebuild_table")
        return out.value

    def table_stats(self):
        """(slots, tombstones, rebuilds so far) of the branch pattern table."""
        a, b, c = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        check(lib().vx_interner_table_stats(self.h, C.byref(a), C.byref(b), C.byref(c)), "table_stats")
        return a.value, b.value, c.value

    def set_graphs(self, enable=True):
        """Replay repeated batched applies from captured HIP graphs."""
        check(lib().vx_interner_set_graphs(self.h, 1 if enable else 0), "set_graphs")

    def profile(self, enable=True):
        check(lib().vx_interner_profile(self.h, 1 if enable else 0), "profile")

    def profile_read(self):
        """{kernel name: (total ms, launches)} since the previous read (synchronizes)."""
        ms = (C.c_double * 64)()
        cnt = (C.c_uint64 * 64)()
        n = lib().vx_interner_profile_read(self.h, ms, cnt, 64)
        if n < 0:
            check(n, "profile_read")
        return {lib().vx_kernel_name(i).decode(): (ms[i], cnt[i]) for i in range(n) if cnt[i]}


class Batch:
    """core/batch.rs Batch<T>: (set, clear) masks and 8 values per 2x2x2 brick."""

    def __init__(self, max_depth, np_dtype=np.uint8):
        self.max_depth = max_depth
        self.masks, self.values = morton.dense_batch_arrays(max_depth, np_dtype)
        self.to_fill = None
        self.has_patches = False

    def just_set(self, x, y, z, v):
        self.set_many(np.array([[x, y, z]], dtype=np.int32), np.array([v]))
        return True

    set = just_set

    def set_many(self, xyz, values):
        xyz = np.asarray(xyz, dtype=np.int32).reshape(-1, 3)
        if len(xyz) == 0:
            return
        morton.apply_edits_to_dense(self.masks, self.values, xyz, np.asarray(values))
        self.has_patches = True

    def just_clear(self):
        self.masks[:] = 0
        self.values[:] = 0
        self.to_fill = None
        self.has_patches = False

    def just_fill(self, v):
        self.just_clear()
        self.to_fill = int(v)

    fill = just_fill

    def size(self):
        return int(np.count_nonzero((self.masks[:, 0] != 0) | (self.masks[:, 1] != 0)))

    def view_host(self):
        v = BatchView()
        v.masks = self.masks.ctypes.data if self.has_patches else None
        v.values = self.values.ctypes.data if self.has_patches else None
        v.fill_value = self.to_fill or 0
        v.depth = self.max_depth
        v.has_fill = 1 if self.to_fill is not None else 0
        v.has_patches = 1 if self.has_patches else 0
        return v


class VoxTree:
    """spatial/voxtree.rs VoxTree<T>: the root id lives on the host, nodes in HBM."""

    def __init__(self, max_depth):
        self.max_depth = max_depth
        self.root = EMPTY
        self.dirty = False

    def create_batch(self, interner):
        return Batch(self.max_depth, interner.np_dtype)

    def voxels_per_axis(self):
        return 1 << self.max_depth

    def is_empty(self):
        return self.root == EMPTY

    def is_leaf(self):
        return (self.root >> 63) == 1

    def apply_batch(self, interner, batch):
        """VoxTree::apply_batch (voxtree.rs:303-328); reference panics raise VxError."""
        r = C.c_uint64(self.root)
        changed = C.c_int(0)
        v = batch.view_host()
        check(lib().vx_tree_apply_batch_host(interner.h, C.byref(r), C.byref(v), C.byref(changed)),
              "apply_batch")
        self.root = r.value
        if changed.value:
            self.dirty = True
        return bool(changed.value)

    def get_many(self, interner, xyz):
        f = Forest(interner, [self.max_depth])
        f.set_roots([self.root])
        vals, pres = f.get(np.zeros(len(xyz), dtype=np.uint32), xyz)
        return vals, pres

    def get(self, interner, x, y, z):
        vals, pres = self.get_many(interner, np.array([[x, y, z]], dtype=np.int32))
        if pres[0] == 0xFF:
            raise VxError(-4, "get")
        return int(vals[0]) if pres[0] else None

    def fill(self, interner, v):
        f = Forest(interner, [self.max_depth])
        f.set_roots([self.root])
        f.fill(np.array([v], dtype=np.uint64))
        was = self.root
        self.root = f.roots_host()[0]
        if v != 0 or was != EMPTY:
            self.dirty = True

    def clear(self, interner):
        self.fill(interner, 0)

    def canonical_hash(self, interner):
        f = Forest(interner, [self.max_depth])
        f.set_roots([self.root])
        return f.canonical_hashes()[0]


class Forest:
    """Many trees sharing one interner, roots resident in HBM (the batched form)."""

    def __init__(self, interner, depths):
        self.it = interner
        self.n = len(depths)
        self.depths_host = np.asarray(depths, dtype=np.uint8)
        self.depths = torch.from_numpy(self.depths_host.copy()).to(interner.device)
        self.roots = torch.zeros(self.n, dtype=torch.int64, device=interner.device)
        self.status = torch.zeros(self.n, dtype=torch.int32, device=interner.device)

    def set_roots(self, roots):
        self.roots.copy_(torch.tensor([_i64(int(r)) for r in roots], dtype=torch.int64))

    def roots_host(self):
        return [_u64(int(r)) for r in self.roots.cpu().tolist()]

    def make_views(self, masks, values, fills=None, has_patches=None):
        """Device batch views over per-tree tensors (masks [P,2] u8, values [P,8]).  The views
        hold the tensors' raw device pointers: the caller keeps the tensors alive while the
        views are used."""
        views = (BatchView * self.n)()
        for i in range(self.n):
            v = views[i]
            hp = True if has_patches is None else bool(has_patches[i])
            v.masks = masks[i].data_ptr() if (hp and masks[i] is not None) else None
            v.values = values[i].data_ptr() if (hp and values[i] is not None) else None
            f = None if fills is None else fills[i]
            v.fill_value = 0 if f is None else int(f)
            v.has_fill = 0 if f is None else 1
            v.has_patches = 1 if hp else 0
            v.depth = int(self.depths_host[i])
        return views

    def _opts(self, dry_run, node_limit, counts):
        """vx_apply_opts; counts=True records new-node counts in self.counts (device)."""
        if not dry_run and not node_limit and not counts:
            return None
        o = ApplyOpts()
        o.flags = APPLY_DRY_RUN if dry_run else 0
        o.node_limit = int(node_limit)
        if counts:
            self.counts = torch.zeros(2 + self.n, dtype=torch.int32, device=self.it.device)
            o.counts = self.counts.data_ptr()
        return o

    def counts_host(self):
        """(new branches, new leaves, per-tree new branches) of the last apply with counts=True;
        new branches/leaves are None when the apply was rolled back."""
        c = self.counts.cpu().numpy().view(np.uint32).astype(np.int64)
        if c[0] == 0xFFFFFFFF:
            return None, None, c[2:]
        return int(c[0]), int(c[1]), c[2:]

    def apply_views(self, views, dry_run=False, node_limit=0, counts=False):
        o = self._opts(dry_run, node_limit, counts)
        if o is None:
            check(lib().vx_trees_apply_batches(self.it.h, self.n, self.roots.data_ptr(), views,
                                               self.status.data_ptr()), "apply_batches")
        else:
            check(lib().vx_trees_apply_batches_ex(self.it.h, self.n, self.roots.data_ptr(), views,
                                                  self.status.data_ptr(), C.byref(o)), "apply_batches")

    def apply(self, masks, values, fills=None, has_patches=None, sync=True, **opts):
        """opts: dry_run, node_limit, counts (vx_apply_opts)."""
        views = self.make_views(masks, values, fills, has_patches)
        self._views = views
        self._keep = (masks, values)
        self.apply_views(views, **opts)
        if sync:
            self.it.sync()
            return self.status.cpu().numpy()
        return None

    def apply_edits(self, offsets, xyz_dev, values_dev, sync=True, dry_run=False, node_limit=0,
                    counts=False):
        """vx_trees_apply_edits: tree t's edits are [offsets[t], offsets[t+1]) of the device
        arrays xyz_dev (int32 [E,3]) and values_dev (value dtype [E]), in submission order."""
        offs = np.ascontiguousarray(offsets, dtype=np.uint64)
        assert len(offs) == self.n + 1
        o = self._opts(dry_run, node_limit, counts)
        args = [self.it.h, self.n, self.roots.data_ptr(), self.depths_host.ctypes.data, offs.ctypes.data,
                xyz_dev.data_ptr() if xyz_dev is not None else None,
                values_dev.data_ptr() if values_dev is not None else None, self.status.data_ptr()]
        if o is None:
            check(lib().vx_trees_apply_edits(*args), "apply_edits")
        else:
            check(lib().vx_trees_apply_edits_ex(*args, C.byref(o)), "apply_edits")
        if sync:
            self.it.sync()
            return self.status.cpu().numpy()
        return None

    def set_voxels(self, xyz_dev, values_dev, sync=True):
        """vx_trees_set: VoxOpsWrite::set of one voxel per tree (tree t: xyz_dev[t], values_dev[t];
        int32 [n,3] and value dtype [n] on the device); returns the per-tree status."""
        check(lib().vx_trees_set(self.it.h, self.n, self.roots.data_ptr(), self.depths_host.ctypes.data,
                                 xyz_dev.data_ptr(), values_dev.data_ptr(), self.status.data_ptr()), "set")
        if sync:
            self.it.sync()
            return self.status.cpu().numpy()
        return None

    def serialize(self, positions):
        """vx_model_serialize: the VTM data section of these trees at chunk positions
        (int32 [n, 3]), as bytes."""
        pos = np.ascontiguousarray(positions, dtype=np.int32).reshape(-1)
        assert len(pos) == 3 * self.n
        # an upper bound of the section from the pattern counts (5-byte varints), so one
        # call writes it
        leaves, branches = self.it.counts()
        s = self.it.value_size
        cap = 12 + leaves * (5 + s) + branches * (5 + 1 + 40 + s) + self.n * (24 + 5)
        buf = (C.c_uint8 * cap)()
        size = C.c_size_t(0)
        rc = lib().vx_model_serialize(self.it.h, self.n, self.roots.data_ptr(), pos.ctypes.data, buf, cap,
                                      C.byref(size))
        if rc == -4:  # not expected: the bound holds for every live set
            buf = (C.c_uint8 * size.value)()
            rc = lib().vx_model_serialize(self.it.h, self.n, self.roots.data_ptr(), pos.ctypes.data, buf,
                                          size.value, C.byref(size))
        check(rc, "serialize")
        return bytes(memoryview(buf)[:size.value])

    @classmethod
    def deserialize(cls, interner, data, depth):
        """vx_model_deserialize into a fresh interner: (forest of depth `depth`, positions [n, 3])."""
        raw = (C.c_uint8 * max(len(data), 1)).from_buffer_copy(data if data else b"\0")
        n = C.c_size_t(0)
        ncap = len(data) // 25 + 1  # a chunk record takes at least 25 bytes
        roots = torch.zeros(ncap, dtype=torch.int64, device=interner.device)
        pos = np.zeros(3 * ncap, dtype=np.int32)
        check(lib().vx_model_deserialize(interner.h, raw, len(data), ncap, roots.data_ptr(), pos.ctypes.data,
                                         C.byref(n)), "deserialize")
        f = cls(interner, [depth] * n.value)
        f.roots.copy_(roots[:n.value])
        return f, pos[:3 * n.value].reshape(-1, 3)

    def get(self, chunk_of, xyz, sync=True):
        nq = len(chunk_of)
        dev = self.it.device
        chunk_t = torch.from_numpy(np.ascontiguousarray(chunk_of, dtype=np.uint32).view(np.int32)).to(dev)
        xyz_t = torch.from_numpy(np.ascontiguousarray(xyz, dtype=np.int32).reshape(-1)).to(dev)
        out = torch.zeros(nq * self.it.value_size, dtype=torch.uint8, device=dev)
        pres = torch.zeros(nq, dtype=torch.uint8, device=dev)
        check(lib().vx_trees_get(self.it.h, self.n, self.roots.data_ptr(), self.depths.data_ptr(), nq,
                                 chunk_t.data_ptr(), xyz_t.data_ptr(), out.data_ptr(),
                                 pres.data_ptr()), "get")
        torch.cuda.current_stream().synchronize()
        rc = lib().vx_interner_sync(self.it.h)
        p = pres.cpu().numpy()
        if rc != 0 and not (rc == -4 and (p == 0xFF).any()):
            check(rc, "get")
        return out.cpu().numpy().view(self.it.np_dtype).astype(np.uint64), p

    def get_device(self, chunk_t, xyz_t, out, pres, stream=None, stop=None):
        """Device-to-device get; stream (a torch.cuda.Stream) runs it off the interner's stream;
        stop (uint8 [nq], needs stream) receives the traversal's stop depth (+0x80 at a leaf)."""
        args = (self.n, self.roots.data_ptr(), self.depths.data_ptr(), chunk_t.numel(), chunk_t.data_ptr(),
                xyz_t.data_ptr(), out.data_ptr(), pres.data_ptr() if pres is not None else None)
        if stream is None and stop is None:
            check(lib().vx_trees_get(self.it.h, *args), "get")
        else:
            sp = C.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)
            check(lib().vx_trees_get_on(self.it.h, sp, *args, stop.data_ptr() if stop is not None else None), "get")

    def get_grouped(self, offsets, xyz):
        """get of queries grouped by tree: tree t's voxels are xyz[offsets[t]:offsets[t+1]]
        (offsets: n+1 non-decreasing ints).  Returns (values u64, present u8) like get."""
        dev = self.it.device
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        nq = int(len(xyz))
        off_t = torch.from_numpy(off.view(np.int64)).to(dev)
        xyz_t = torch.from_numpy(np.ascontiguousarray(xyz, dtype=np.int32).reshape(-1)).to(dev)
        out = torch.zeros(max(nq, 1) * self.it.value_size, dtype=torch.uint8, device=dev)
        pres = torch.zeros(max(nq, 1), dtype=torch.uint8, device=dev)
        self.get_grouped_device(off_t, xyz_t, out, pres, nq=nq)
        torch.cuda.current_stream().synchronize()
        rc = lib().vx_interner_sync(self.it.h)
        p = pres.cpu().numpy()[:nq]
        if rc != 0 and not (rc == -4 and (p == 0xFF).any()):
            check(rc, "get_grouped")
        return out.cpu().numpy().view(self.it.np_dtype)[:nq].astype(np.uint64), p

    def get_grouped_device(self, off_t, xyz_t, out, pres, nq=None, stream=None, stop=None):
        """Device form of get_grouped (off_t int64 [n+1], xyz_t int32 [3 nq])."""
        nq = xyz_t.numel() // 3 if nq is None else nq
        sp = C.c_void_p(stream.cuda_stream) if stream is not None else None
        check(lib().vx_trees_get_grouped(self.it.h, sp, self.n, self.roots.data_ptr(), self.depths.data_ptr(),
                                         off_t.data_ptr(), nq, xyz_t.data_ptr(), out.data_ptr(),
                                         pres.data_ptr() if pres is not None else None,
                                         stop.data_ptr() if stop is not None else None), "get_grouped")

    def fill(self, values):
        vals = torch.from_numpy(np.asarray(values, dtype=np.uint64).view(np.int64)).to(self.it.device)
        check(lib().vx_trees_fill(self.it.h, self.n, self.roots.data_ptr(), vals.data_ptr()), "fill")
        self.it.sync()

    def to_vec(self, lod=0):
        """to_vec of every tree at LOD `lod`: a list of numpy arrays [N, N, N] indexed
        [y, z, x] (N = 2^(depth - lod))."""
        d = np.maximum(self.depths_host.astype(np.int64) - lod, 0)
        stride = int((1 << (3 * int(d.max()))) if self.n else 1)
        out = torch.zeros(self.n * stride * self.it.value_size, dtype=torch.uint8, device=self.it.device)
        check(lib().vx_trees_to_vec(self.it.h, self.n, self.roots.data_ptr(), self.depths.data_ptr(), int(lod),
                                    stride, out.data_ptr()), "to_vec")
        self.it.sync()
        flat = out.cpu().numpy().view(self.it.np_dtype).reshape(self.n, stride)
        res = []
        for t in range(self.n):
            N = 1 << int(d[t])
            res.append(flat[t, :N ** 3].reshape(N, N, N))
        return res

    def canonical_hashes(self):
        out = torch.zeros(2 * self.n, dtype=torch.int64, device=self.it.device)
        check(lib().vx_trees_canonical_hash(self.it.h, self.n, self.roots.data_ptr(), out.data_ptr()),
              "canonical_hash")
        self.it.sync()
        o = out.cpu().numpy().view(np.uint64).reshape(-1, 2)
        return [(int(a), int(b)) for a, b in o]
