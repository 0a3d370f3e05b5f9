"""CPU restatement of the VTM data section, for tests only (the checker, never the
product): VoxModel::serialize / deserialize (voxelis/src/world/voxmodel.rs:177-408),
the chunk records (world/voxchunk.rs:381-440) and the varint (io/varint.rs).

Pure Python over small worlds.  Node access goes through callables so the module
does not depend on a binding: children(bid) -> 8 BlockIds, value(bid) -> int.

Pinned by tests/test_vtm_oracle.py: LEB128 textbook values, the section's length
closed form for the plain checker tree, and serialize -> deserialize -> rebuild
giving back every tree's canonical hash.
"""
import struct

VTC_MAGIC = b"VoxTreeChunk"


def _is_leaf(bid):
    return (bid >> 63) & 1 == 1


def _index(bid):
    return bid & 0xFFFFFFFF


def _mask(bid):
    return (bid >> 47) & 0xFF


# io/varint.rs:5-15 (encode) and 63-84 (decode, at most 32 bits)
def varint_encode(v):
    out = bytearray()
    while v >= 0x80:
        out.append((v & 0x7F) | 0x80)
        v >>= 7
    out.append(v)
    return bytes(out)


def varint_decode(buf, at):
    """(value, next offset), or (None, at) for a value beyond 32 bits or a cut stream."""
    result, shift = 0, 0
    while True:
        if at >= len(buf):
            return None, at
        b = buf[at]
        at += 1
        result |= (b & 0x7F) << shift
        if not b & 0x80:
            return result, at
        shift += 7
        if shift >= 32:
            return None, at


def live_nodes(roots, children):
    """Leaf and branch ids reachable from the roots (the interner's live patterns when
    no node leaked); the EMPTY branch is left out."""
    leaves, branches, stack = set(), set(), [r for r in roots if r != 0]
    while stack:
        b = stack.pop()
        if _is_leaf(b):
            leaves.add(b)
        elif b not in branches:
            branches.add(b)
            stack.extend(c for c in children(b) if c != 0)
    return leaves, branches


# values are big-endian T::to_be_bytes (u8, u16, i32/u32 as their bit pattern)
_FMT = {1: ">B", 2: ">H", 4: ">I"}


def serialize(leaves, branches, roots, positions, children, value, value_size):
    """voxmodel.rs:177-283: ids 1.. for leaves then branches, each sorted by node index."""
    leaves = sorted(leaves, key=_index)
    branches = sorted(branches, key=_index)
    id_map = {0: 0}
    nxt = 1
    for b in leaves:
        id_map[_index(b)] = nxt
        nxt += 1
    for b in branches:
        id_map[_index(b)] = nxt
        nxt += 1
    fmt = _FMT[value_size]
    out = bytearray(struct.pack(">I", len(leaves)))
    for b in leaves:
        out += varint_encode(id_map[_index(b)]) + struct.pack(fmt, value(b))
    out += struct.pack(">I", len(branches))  # branch_patterns().len() - 1: the EMPTY branch is not written
    for b in branches:
        out += varint_encode(id_map[_index(b)]) + struct.pack(">B", _mask(b))
        for c in children(b):
            if c != 0:
                out += varint_encode(id_map[_index(c)])
        out += struct.pack(fmt, value(b))
    out += struct.pack(">I", len(roots))
    for r, p in zip(roots, positions):
        out += VTC_MAGIC + struct.pack(">iii", *[int(x) for x in p]) + varint_encode(id_map[_index(r)] if r else 0)
    return bytes(out)


def deserialize(data, value_size):
    """voxmodel.rs:285-408 parsing: ({leaf id: value}, {branch id: (mask, [8 child ids], lod)},
    [(position, root id)]); ids are checked to run 1.. as the importer asserts."""
    size = value_size
    fmt = _FMT[value_size]
    at = 0
    (nl,) = struct.unpack_from(">I", data, at)
    at += 4
    leaves = {}
    for i in range(nl):
        fid, at = varint_decode(data, at)
        assert fid == i + 1, "leaf ids run 1.."
        (leaves[fid],) = struct.unpack_from(fmt, data, at)
        at += size
    (nb,) = struct.unpack_from(">I", data, at)
    at += 4
    branches = {}
    for j in range(nb):
        fid, at = varint_decode(data, at)
        assert fid == nl + 1 + j, "branch ids follow the leaves"
        mask = data[at]
        at += 1
        assert mask != 0
        kids = [0] * 8
        for k in range(8):
            if mask & (1 << k):
                kids[k], at = varint_decode(data, at)
        (lod,) = struct.unpack_from(fmt, data, at)
        at += size
        branches[fid] = (mask, kids, lod)
    (nc,) = struct.unpack_from(">I", data, at)
    at += 4
    chunks = []
    for _ in range(nc):
        assert data[at:at + 12] == VTC_MAGIC
        pos = struct.unpack_from(">iii", data, at + 12)
        at += 24
        rid, at = varint_decode(data, at)
        chunks.append((pos, rid))
    assert at == len(data)
    return leaves, branches, chunks


def rebuild(parsed, leaf_of, branch_of):
    """Root ids of the parsed chunks, rebuilt bottom-up with the interner's own
    get_or_create_leaf(v) / get_or_create_branch(children, types, mask) (leaf_of,
    branch_of); children are resolved before their parents whatever the id order."""
    leaves, branches, chunks = parsed
    made = {0: 0}

    def node(fid):
        if fid in made:
            return made[fid]
        if fid in leaves:
            made[fid] = leaf_of(leaves[fid])
            return made[fid]
        mask, kids, _lod = branches[fid]
        ch = [node(k) if mask & (1 << i) else 0 for i, k in enumerate(kids)]
        types = sum(1 << i for i, c in enumerate(ch) if c != 0 and _is_leaf(c))
        made[fid] = branch_of(ch, types, mask)
        return made[fid]

    return [node(rid) for _, rid in chunks], [pos for pos, _ in chunks]
