"""VTM files of a resident world: export_model_to_vtm / import_model_from_vtm
(io/export.rs:90-151, io/import.rs:14-98).

Layout (big-endian): magic "VoxTreeModel", version u16 0x0100, flags u16, max depth
u8, chunk world size f32, two reserved u32, world bounds 3 x i32, name length u8 and
name, MD5 of the (uncompressed) data section, data length u32, data.  The data
section is VoxModel::serialize (vx_model_serialize, computed on the device).

Flags: COMPRESSED (bit 0) means the data is a zstd stream.  No zstd codec is
available to this package, so files are written uncompressed (flags 0, which the
reference's importer reads as they are) and a compressed file is refused with a
clear error."""
import hashlib
import struct

import numpy as np

from . import api

VTM_MAGIC = b"VoxTreeModel"
VTM_VERSION = 0x0100
FLAG_COMPRESSED = 0x1


def _header(name, max_depth, chunk_world_size, world_bounds, flags, md5, data_len):
    nb = name.encode()
    if len(nb) > 255:
        raise ValueError("name longer than 255 bytes")
    return (VTM_MAGIC + struct.pack(">HHBfII", VTM_VERSION, flags, max_depth, chunk_world_size, 0, 0) +
            struct.pack(">iii", *[int(b) for b in world_bounds]) + struct.pack(">B", len(nb)) + nb + md5 +
            struct.pack(">I", data_len))


def export_model_to_vtm(name, path, forest, positions, chunk_world_size=1.0, world_bounds=(0, 0, 0)):
    """Write the trees of `forest` (one max depth) at chunk `positions` (int32 [n, 3])."""
    depths = set(int(d) for d in forest.depths_host)
    if len(depths) > 1:
        raise ValueError("a VTM model has one max depth for every chunk")
    max_depth = depths.pop() if depths else 0
    data = forest.serialize(positions)
    md5 = hashlib.md5(data).digest()
    with open(path, "wb") as fh:
        fh.write(_header(name, max_depth, chunk_world_size, world_bounds, 0, md5, len(data)))
        fh.write(data)
    return len(data)


def read_vtm(path):
    """(header fields, data section) of a VTM file; checks magic, version and MD5."""
    with open(path, "rb") as fh:
        raw = fh.read()
    if raw[:12] != VTM_MAGIC:
        raise ValueError("not a VTM file")
    version, flags, max_depth, chunk_world_size, _r1, _r2 = struct.unpack(">HHBfII", raw[12:29])
    if version != VTM_VERSION:
        raise ValueError(f"VTM version {version:#x}")
    bounds = struct.unpack(">iii", raw[29:41])
    nlen = raw[41]
    name = raw[42:42 + nlen].decode()
    at = 42 + nlen
    md5 = raw[at:at + 16]
    (size,) = struct.unpack(">I", raw[at + 16:at + 20])
    data = raw[at + 20:at + 20 + size]
    if flags & FLAG_COMPRESSED:
        raise NotImplementedError("compressed (zstd) VTM data: no zstd codec is available here")
    if hashlib.md5(data).digest() != md5:
        raise ValueError("VTM data MD5 mismatch")
    return {"name": name, "max_depth": max_depth, "chunk_world_size": chunk_world_size,
            "world_bounds": bounds, "flags": flags}, data


def import_model_from_vtm(path, memory_budget, target_chunk_world_size=None, value_size=1, load_factor=0.5):
    """A fresh interner holding the file's world: (meta, interner, forest, positions)."""
    meta, data = read_vtm(path)
    if target_chunk_world_size is not None:
        meta["chunk_world_size"] = target_chunk_world_size
    it = api.VoxInterner(value_size, memory_budget, load_factor)
    forest, positions = api.Forest.deserialize(it, data, meta["max_depth"])
    return meta, it, forest, np.asarray(positions)
