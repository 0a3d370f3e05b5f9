"""ctypes binding of libvoxelis_hip.so (the C ABI in include/voxelis_hip.h).

The HIP library is the product path; there is no fallback.  Loading fails with
an ImportError naming the missing file when the library has not been built.
"""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libvoxelis_hip.so")

VX_OK = 0
STATUS_NAMES = {
    0: "OK", -1: "INVALID_ARG", -2: "OOM", -3: "DEPTH", -4: "RANGE", -5: "LEAF_BRICK",
    -6: "SAME_ROOT", -7: "HIP", -8: "BUDGET", -9: "UNSUPPORTED", -10: "TABLE_FULL",
    -11: "POISONED", -12: "SCRATCH",
}
STAT_NAMES = ["branch_hits", "branch_misses", "branch_dups", "leaf_misses", "collapses",
              "recycled", "h2_uncollapsed", "h3_leaks", "h4_undercount", "h5_collapse_cloned",
              "candidates", "probes", "claims", "claims_lost", "merged", "brick_keys", "brick_merged",
              "brick_hits", "brick_misses", "brick_dups"]

# Every symbol include/voxelis_hip.h declares, with its ctypes signature.
_u8, _u32, _u64, _sz, _vp, _int = (C.c_uint8, C.c_uint32, C.c_uint64, C.c_size_t, C.c_void_p,
                                   C.c_int)


class BatchView(C.Structure):
    _fields_ = [("masks", C.c_void_p), ("values", C.c_void_p), ("fill_value", C.c_uint64),
                ("depth", C.c_uint8), ("has_fill", C.c_uint8), ("has_patches", C.c_uint8),
                ("reserved", C.c_uint8)]


APPLY_DRY_RUN = 1


class Tuning(C.Structure):
    _fields_ = [("tile", C.c_uint32), ("dedup_rounds", C.c_uint32), ("commit_grid", C.c_uint32),
                ("gc_grid", C.c_uint32), ("probe_blocks_bricks", C.c_uint32), ("probe_blocks", C.c_uint32),
                ("brick_merge_rounds", C.c_uint32), ("brick_rounds", C.c_uint32)]


class ApplyOpts(C.Structure):
    _fields_ = [("flags", C.c_uint32), ("reserved", C.c_uint32), ("node_limit", C.c_uint64),
                ("counts", C.c_void_p)]


SIGNATURES = {
    "vx_interner_create": (_int, [_u32, _u64, C.c_double, C.POINTER(_vp)]),
    "vx_interner_destroy": (_int, [_vp]),
    "vx_interner_capacity": (_int, [_vp, C.POINTER(_u64)]),
    "vx_interner_counts": (_int, [_vp, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)]),
    "vx_interner_get_ref": (_int, [_vp, _u64, C.POINTER(_u32)]),
    "vx_interner_get_value": (_int, [_vp, _u64, C.POINTER(_u64)]),
    "vx_interner_get_children": (_int, [_vp, _u64, C.POINTER(_u64)]),
    "vx_interner_stats": (_int, [_vp, C.POINTER(_u64), _u32]),
    "vx_interner_set_stream": (_int, [_vp, _vp]),
    "vx_interner_sync": (_int, [_vp]),
    "vx_interner_set_graphs": (_int, [_vp, _int]),
    "vx_interner_rebuild_table": (_int, [_vp, C.POINTER(_u64)]),
    "vx_interner_table_stats": (_int, [_vp, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)]),
    "vx_interner_set_tuning": (_int, [_vp, C.POINTER(Tuning)]),
    "vx_interner_get_tuning": (_int, [_vp, C.POINTER(Tuning)]),
    "vx_trees_apply_batches": (_int, [_vp, _sz, _vp, C.POINTER(BatchView), _vp]),
    "vx_tree_apply_batch_host": (_int, [_vp, C.POINTER(_u64), C.POINTER(BatchView),
                                        C.POINTER(_int)]),
    "vx_trees_apply_edits": (_int, [_vp, _sz, _vp, _vp, _vp, _vp, _vp, _vp]),
    "vx_trees_apply_batches_ex": (_int, [_vp, _sz, _vp, C.POINTER(BatchView), _vp,
                                         C.POINTER(ApplyOpts)]),
    "vx_trees_apply_edits_ex": (_int, [_vp, _sz, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(ApplyOpts)]),
    "vx_trees_get": (_int, [_vp, _sz, _vp, _vp, _sz, _vp, _vp, _vp, _vp]),
    "vx_trees_get_on": (_int, [_vp, _vp, _sz, _vp, _vp, _sz, _vp, _vp, _vp, _vp, _vp]),
    "vx_trees_get_grouped": (_int, [_vp, _vp, _sz, _vp, _vp, _vp, _sz, _vp, _vp, _vp, _vp]),
    "vx_trees_fill": (_int, [_vp, _sz, _vp, _vp]),
    "vx_trees_set": (_int, [_vp, _sz, _vp, _vp, _vp, _vp, _vp]),
    "vx_model_serialize": (_int, [_vp, _sz, _vp, _vp, _vp, _sz, C.POINTER(C.c_size_t)]),
    "vx_model_serialize_noalloc": (_int, [_vp, _sz, _vp, _vp, _vp, _sz, C.POINTER(C.c_size_t)]),
    "vx_model_deserialize": (_int, [_vp, _vp, _sz, _sz, _vp, _vp, C.POINTER(C.c_size_t)]),
    "vx_trees_canonical_hash": (_int, [_vp, _sz, _vp, _vp]),
    "vx_trees_to_vec": (_int, [_vp, _sz, _vp, _vp, _u32, _u64, _vp]),
    "vx_interner_profile": (_int, [_vp, _int]),
    "vx_interner_profile_read": (_int, [_vp, C.POINTER(C.c_double), C.POINTER(_u64), _u32]),
    "vx_kernel_name": (C.c_char_p, [_u32]),
    "vx_bench_atomics": (_int, [_vp, _int, _vp, _u64, _u64, C.POINTER(C.c_float)]),
    "vx_bench_copy": (_int, [_vp, _vp, _vp, _u64, C.POINTER(C.c_float)]),
    "vx_mlp_param_count": (_int, []),
    "vx_mlp_forward": (_int, [_vp, _vp, _vp, _u32, _vp]),
    "vx_mlp_gradient": (_int, [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp]),
    "vx_mlp_adam": (_int, [_vp, _vp, _vp, _vp, _vp, _u32, C.c_float]),
    "vx_device_alloc": (_int, [_sz, C.POINTER(_vp)]),
    "vx_device_free": (_int, [_vp]),
    "vx_memcpy_h2d": (_int, [_vp, _vp, _sz]),
    "vx_memcpy_d2h": (_int, [_vp, _vp, _sz]),
    "vx_status_string": (C.c_char_p, [_int]),
}


class VxError(RuntimeError):
    def __init__(self, code, where=""):
        name = STATUS_NAMES.get(code, str(code))
        super().__init__(f"{where}: {name} ({code})" if where else f"{name} ({code})")
        self.code = code
        self.kind = name


_lib = None


def lib():
    """Load the HIP library (no fallback: raises ImportError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          f"(hipcc --offload-arch=gfx950)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc, where=""):
    if rc != VX_OK:
        raise VxError(rc, where)
    return rc
