"""Morton paths and dense Batch buffers (host side, numpy).

encode_path follows utils/common.rs:24-55: x -> bit 3n, y -> 3n+1, z -> 3n+2.
build_dense_batch reproduces core/batch.rs:145-175 just_set over many edits
with the reference's last-write-wins order.
"""
import numpy as np


def _spread3(v):
    v = v.astype(np.uint32) & np.uint32(0x3FF)
    v = (v | (v << 16)) & np.uint32(0x30000FF)
    v = (v | (v << 8)) & np.uint32(0x300F00F)
    v = (v | (v << 4)) & np.uint32(0x30C30C3)
    v = (v | (v << 2)) & np.uint32(0x9249249)
    return v


def encode_path(xyz):
    """Morton path of int32 [n, 3] coordinates (x, y, z) -> uint32 [n]."""
    xyz = np.asarray(xyz)
    return _spread3(xyz[:, 0]) | (_spread3(xyz[:, 1]) << 1) | (_spread3(xyz[:, 2]) << 2)


def bricks_for_depth(depth):
    """Batch::new size (batch.rs:67-72): 8^(depth-1) bricks, 1 for depth 0."""
    return 1 << (3 * max(depth - 1, 0))


def dense_batch_arrays(depth, dtype):
    n = bricks_for_depth(depth)
    return np.zeros((n, 2), dtype=np.uint8), np.zeros((n, 8), dtype=dtype)


def last_write_indices(keys):
    """Indices of the last occurrence of every distinct key, ascending by index."""
    keys = np.asarray(keys)
    if len(keys) == 0:
        return np.zeros(0, dtype=np.int64)
    rev = keys[::-1]
    _, first_rev = np.unique(rev, return_index=True)
    return np.sort(len(keys) - 1 - first_rev)


def apply_edits_to_dense(masks, values, xyz, vals):
    """just_set for each (xyz, value) in order: set bit for non-default values, clear
    bit for the default, value stored regardless (batch.rs:160-170)."""
    full = encode_path(xyz)
    keep = last_write_indices(full)
    full = full[keep]
    vals = np.asarray(vals)[keep].astype(values.dtype)
    p = (full >> 3).astype(np.int64)
    slot = (full & 7).astype(np.int64)
    bit = (np.uint8(1) << slot.astype(np.uint8)).astype(np.uint8)
    nz = vals != 0
    np.bitwise_or.at(masks[:, 0], p[nz], bit[nz])
    np.bitwise_and.at(masks[:, 1], p[nz], ~bit[nz])
    np.bitwise_and.at(masks[:, 0], p[~nz], ~bit[~nz])
    np.bitwise_or.at(masks[:, 1], p[~nz], bit[~nz])
    values[p, slot] = vals
    return masks, values
