"""ctypes view of the CPU oracle (oracle/liboracle.so) for the test suite.

The oracle is test infrastructure: a literal C restatement of the reference
path.  Tests use it only as the checker the HIP path is compared against.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

EMPTY = 0
INVALID = (1 << 64) - 1

PANIC = {
    -1: "OOM", -2: "DEPTH", -3: "RANGE", -4: "LEAF_TYPES", -5: "SAME_ROOT",
    -6: "EXPIRED", -7: "BUDGET", -8: "GENERATION", -9: "INVALID", -10: "REMOVE",
}
ST_NAMES = ["leaf_hits", "leaf_misses", "branch_hits", "branch_misses", "collapses",
            "recycled", "h2_uncollapsed", "h3_leaks", "h4_undercount",
            "h5_collapse_cloned", "transient_free"]


class OraclePanic(RuntimeError):
    def __init__(self, code):
        super().__init__(f"reference panic: {PANIC.get(code, code)}")
        self.code = code
        self.kind = PANIC.get(code, str(code))


def _build_if_missing():
    if not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    _build_if_missing()
    L = C.CDLL(LIB_PATH)
    u8, u16, u32, u64, i32, sz, vp = (C.c_uint8, C.c_uint16, C.c_uint32, C.c_uint64,
                                      C.c_int32, C.c_size_t, C.c_void_p)
    P64 = C.POINTER(C.c_uint64)
    sig = {
        "vxo_bid_new_leaf": (u64, [u32, u16]),
        "vxo_bid_new_branch": (u64, [u32, u16, u8, u8]),
        "vxo_bid_path_mask": (u32, [u32, u32]),
        "vxo_encode_child_index_path": (u32, [i32, i32, i32]),
        "vxo_calc_average": (u64, [P64, sz]),
        "vxo_interner_new": (vp, [u32, u64, C.POINTER(C.c_int)]),
        "vxo_interner_free": (None, [vp]),
        "vxo_set_extended_depth": (None, [vp, C.c_int]),
        "vxo_capacity": (u64, [vp]),
        "vxo_node_size": (u64, [u32]),
        "vxo_counts": (None, [vp, P64, P64]),
        "vxo_get_ref": (u32, [vp, u64]),
        "vxo_get_value": (u64, [vp, u64]),
        "vxo_get_children": (None, [vp, u64, P64]),
        "vxo_stats": (None, [vp, P64]),
        "vxo_get_or_create_leaf": (C.c_int, [vp, u64, P64]),
        "vxo_get_or_create_branch": (C.c_int, [vp, P64, u8, u8, P64]),
        "vxo_dec_ref_recursive": (C.c_int, [vp, u64]),
        "vxo_alive_nodes": (u64, [vp]),
        "vxo_batch_new": (vp, [u8]),
        "vxo_batch_free": (None, [vp]),
        "vxo_batch_just_set": (None, [vp, i32, i32, i32, u64]),
        "vxo_batch_set_many": (None, [vp, sz, vp, vp]),
        "vxo_batch_just_fill": (None, [vp, u64]),
        "vxo_batch_just_clear": (None, [vp]),
        "vxo_batch_size": (sz, [vp]),
        "vxo_batch_has_patches": (C.c_int, [vp]),
        "vxo_batch_load": (None, [vp, vp, vp, u32, C.c_int]),
        "vxo_batch_bricks": (sz, [vp]),
        "vxo_batch_masks": (vp, [vp]),
        "vxo_batch_values": (vp, [vp]),
        "vxo_tree_get": (C.c_int, [vp, u64, u8, i32, i32, i32, P64, C.POINTER(C.c_int)]),
        "vxo_tree_get_many": (C.c_int, [vp, u64, u8, sz, vp, vp, vp]),
        "vxo_tree_set": (C.c_int, [vp, P64, u8, i32, i32, i32, u64, C.POINTER(C.c_int)]),
        "vxo_trees_set_round": (C.c_int, [vp, sz, vp, vp, vp, vp, vp]),
        "vxo_tree_fill": (C.c_int, [vp, P64, u8, u64]),
        "vxo_tree_clear": (C.c_int, [vp, P64]),
        "vxo_tree_apply_batch": (C.c_int, [vp, P64, u8, vp, C.POINTER(C.c_int)]),
        "vxo_snapshot_take": (vp, [vp]),
        "vxo_snapshot_restore": (None, [vp, vp]),
        "vxo_snapshot_free": (None, [vp]),
        "vxo_tree_to_vec": (C.c_int, [vp, u64, u8, vp]),
        "vxo_tree_canonical_hash": (C.c_int, [vp, u64, u8, P64]),
        "vxo_live_patterns": (C.c_int, [vp, vp, vp, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise OraclePanic(rc)


def all_coords(depth):
    """All voxel coordinates of a 2^depth cube in (y, z, x) loop order, as int32 [n, 3] xyz."""
    n = 1 << depth
    y, z, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return np.ascontiguousarray(np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32))


class Batch:
    """Mirror of core/batch.rs Batch<T> held by the oracle."""

    def __init__(self, depth):
        self.depth = depth
        self.h = lib().vxo_batch_new(depth)

    def __del__(self):
        if getattr(self, "h", None):
            lib().vxo_batch_free(self.h)
            self.h = None

    def just_set(self, x, y, z, v):
        lib().vxo_batch_just_set(self.h, x, y, z, v)
        return True

    set = just_set

    def set_many(self, xyz, values):
        xyz = np.ascontiguousarray(xyz, dtype=np.int32)
        vals = np.ascontiguousarray(values, dtype=np.uint64)
        lib().vxo_batch_set_many(self.h, len(vals), xyz.ctypes.data, vals.ctypes.data)

    def just_fill(self, v):
        lib().vxo_batch_just_fill(self.h, v)

    fill = just_fill

    def just_clear(self):
        lib().vxo_batch_just_clear(self.h)

    def load_dense(self, masks, values):
        masks = np.ascontiguousarray(masks, dtype=np.uint8)
        values = np.ascontiguousarray(values)
        lib().vxo_batch_load(self.h, masks.ctypes.data, values.ctypes.data, values.itemsize,
                             1 if masks[:, 0].any() or masks[:, 1].any() else 0)

    def size(self):
        return lib().vxo_batch_size(self.h)

    def has_patches(self):
        return bool(lib().vxo_batch_has_patches(self.h))

    def bricks(self):
        return lib().vxo_batch_bricks(self.h)

    def masks(self):
        n = self.bricks()
        p = C.cast(lib().vxo_batch_masks(self.h), C.POINTER(C.c_uint8 * (2 * n)))
        return np.frombuffer(p.contents, dtype=np.uint8).reshape(n, 2).copy()

    def values(self):
        n = self.bricks()
        p = C.cast(lib().vxo_batch_values(self.h), C.POINTER(C.c_uint64 * (8 * n)))
        return np.frombuffer(p.contents, dtype=np.uint64).reshape(n, 8).copy()


def set_round(it, trees, xyz, values):
    """One VoxOpsWrite::set on each tree (tree i: xyz[i], values[i]) inside one C call;
    returns the per-tree status (1 changed, 0 unchanged, < 0 the panic's code)."""
    n = len(trees)
    roots = np.array([t.root for t in trees], dtype=np.uint64)
    depths = np.array([t.depth for t in trees], dtype=np.uint8)
    xyz = np.ascontiguousarray(xyz, dtype=np.int32).reshape(-1)
    vals = np.ascontiguousarray(values, dtype=np.uint64)
    st = np.zeros(n, dtype=np.int32)
    lib().vxo_trees_set_round(it.h, n, roots.ctypes.data, depths.ctypes.data, xyz.ctypes.data, vals.ctypes.data,
                              st.ctypes.data)
    for i, t in enumerate(trees):
        if st[i] == 1:
            t.dirty = True
        t.root = int(roots[i])
    return st


class Interner:
    """Mirror of interner/mod.rs VoxInterner<T> held by the oracle."""

    def __init__(self, value_size, budget, extended_depth=False):
        err = C.c_int(0)
        self.h = lib().vxo_interner_new(value_size, budget, C.byref(err))
        if not self.h:
            raise OraclePanic(err.value)
        self.value_size = value_size
        if extended_depth:
            lib().vxo_set_extended_depth(self.h, 1)

    def __del__(self):
        if getattr(self, "h", None):
            lib().vxo_interner_free(self.h)
            self.h = None

    def capacity(self):
        return lib().vxo_capacity(self.h)

    def counts(self):
        a, b = C.c_uint64(0), C.c_uint64(0)
        lib().vxo_counts(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def get_ref(self, bid):
        return lib().vxo_get_ref(self.h, bid)

    def get_value(self, bid):
        return lib().vxo_get_value(self.h, bid)

    def children(self, bid):
        out = (C.c_uint64 * 8)()
        lib().vxo_get_children(self.h, bid, out)
        return list(out)

    def stats(self):
        out = (C.c_uint64 * len(ST_NAMES))()
        lib().vxo_stats(self.h, out)
        return dict(zip(ST_NAMES, list(out)))

    def alive_nodes(self):
        return lib().vxo_alive_nodes(self.h)

    def live_patterns(self):
        """(leaf values, branch canonical hash pairs) of every live pattern (EMPTY excluded)."""
        nl, nb = C.c_size_t(0), C.c_size_t(0)
        lib().vxo_live_patterns(self.h, None, None, 0, C.byref(nl), C.byref(nb))
        cap = max(nl.value, nb.value, 1)
        lv = np.zeros(cap, dtype=np.uint64)
        bc = np.zeros((cap, 2), dtype=np.uint64)
        _check(lib().vxo_live_patterns(self.h, lv.ctypes.data, bc.ctypes.data, cap, C.byref(nl), C.byref(nb)))
        return lv[: nl.value], bc[: nb.value]

    def get_or_create_leaf(self, v):
        out = C.c_uint64(0)
        _check(lib().vxo_get_or_create_leaf(self.h, v, C.byref(out)))
        return out.value

    def get_or_create_branch(self, children, types, mask):
        ch = (C.c_uint64 * 8)(*children)
        out = C.c_uint64(0)
        _check(lib().vxo_get_or_create_branch(self.h, ch, types, mask, C.byref(out)))
        return out.value

    def dec_ref_recursive(self, bid):
        _check(lib().vxo_dec_ref_recursive(self.h, bid))


class Tree:
    """Mirror of spatial/voxtree.rs VoxTree<T>: max depth, root id, dirty flag."""

    def __init__(self, depth):
        self.depth = depth
        self.root = EMPTY
        self.dirty = False

    def create_batch(self):
        return Batch(self.depth)

    def voxels_per_axis(self):
        return 1 << self.depth

    def is_empty(self):
        return self.root == EMPTY

    def is_leaf(self):
        return (self.root >> 63) == 1

    def get(self, it, x, y, z):
        v, p = C.c_uint64(0), C.c_int(0)
        _check(lib().vxo_tree_get(it.h, self.root, self.depth, x, y, z, C.byref(v), C.byref(p)))
        return v.value if p.value else None

    def get_many(self, it, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.int32)
        n = len(xyz)
        vals = np.zeros(n, dtype=np.uint64)
        pres = np.zeros(n, dtype=np.uint8)
        _check(lib().vxo_tree_get_many(it.h, self.root, self.depth, n, xyz.ctypes.data,
                                       vals.ctypes.data, pres.ctypes.data))
        return vals, pres

    def set(self, it, x, y, z, v):
        r = C.c_uint64(self.root)
        ch = C.c_int(0)
        _check(lib().vxo_tree_set(it.h, C.byref(r), self.depth, x, y, z, v, C.byref(ch)))
        self.root = r.value
        if ch.value:
            self.dirty = True
        return bool(ch.value)

    def fill(self, it, v):
        r = C.c_uint64(self.root)
        _check(lib().vxo_tree_fill(it.h, C.byref(r), self.depth, v))
        if v != 0 or self.root != EMPTY:
            self.dirty = True
        self.root = r.value

    def clear(self, it):
        r = C.c_uint64(self.root)
        was = self.root
        _check(lib().vxo_tree_clear(it.h, C.byref(r)))
        if was != EMPTY:
            self.dirty = True
        self.root = r.value

    def apply_batch(self, it, batch):
        r = C.c_uint64(self.root)
        ch = C.c_int(0)
        _check(lib().vxo_tree_apply_batch(it.h, C.byref(r), self.depth, batch.h, C.byref(ch)))
        self.root = r.value
        if ch.value:
            self.dirty = True
        return bool(ch.value)

    def apply_batch_atomic(self, it, batch):
        """apply_batch with the device path's policy for a reference panic: the tree and
        the interner are left exactly as before the call (the panic is re-raised)."""
        snap = lib().vxo_snapshot_take(it.h)
        try:
            return self.apply_batch(it, batch)
        except OraclePanic:
            lib().vxo_snapshot_restore(it.h, snap)
            raise
        finally:
            lib().vxo_snapshot_free(snap)

    def to_vec(self, it, lod=0):
        """utils/common.rs to_vec with MaxDepth::for_lod(lod) (saturating)."""
        d = max(self.depth - lod, 0)
        n = 1 << d
        out = np.zeros(n * n * n, dtype=np.uint64)
        _check(lib().vxo_tree_to_vec(it.h, self.root, d, out.ctypes.data))
        return out

    def canonical_hash(self, it):
        out = (C.c_uint64 * 2)()
        _check(lib().vxo_tree_canonical_hash(it.h, self.root, self.depth, out))
        return (out[0], out[1])


def bid_new_leaf(index, gen):
    return lib().vxo_bid_new_leaf(index, gen)


def bid_new_branch(index, gen, types, mask):
    return lib().vxo_bid_new_branch(index, gen, types, mask)


def path_mask(max_depth, level):
    return lib().vxo_bid_path_mask(max_depth, level)


def encode_path(x, y, z):
    return lib().vxo_encode_child_index_path(x, y, z)


def calc_average(values):
    arr = (C.c_uint64 * len(values))(*values)
    return lib().vxo_calc_average(arr, len(values))
