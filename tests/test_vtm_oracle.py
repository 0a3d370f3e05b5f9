"""Pins the CPU restatement of the VTM data section (oracle/vtm_oracle.py): the
varint against LEB128 textbook values, the section's exact length for the plain
checker tree (closed form: 2 leaves, 5 branches), and serialize -> deserialize ->
rebuild giving back every tree's canonical hash on multi-tree worlds."""
import struct

import numpy as np

import oracle_lib as O
from gpu_helpers import oracle_apply_dense
from oracle import vtm_oracle as V
from voxelis_amd import workloads as W

MB = 1024 * 1024


def test_varint_textbook_values():
    cases = {0: b"\x00", 1: b"\x01", 127: b"\x7f", 128: b"\x80\x01", 300: b"\xac\x02", 16383: b"\xff\x7f",
             16384: b"\x80\x80\x01", 0xFFFFFFFF: b"\xff\xff\xff\xff\x0f"}
    for v, enc in cases.items():
        assert V.varint_encode(v) == enc
        assert V.varint_decode(enc, 0) == (v, len(enc))
    # a stream that does not end within 32 bits is rejected, as io/varint.rs:63-84
    assert V.varint_decode(b"\x80\x80\x80\x80\x80\x01", 0)[0] is None
    assert V.varint_decode(b"\x80", 0)[0] is None


def _oracle_world(trees, seed, vsize=1):
    it = O.Interner(vsize, 64 * MB)
    ts = []
    for t in range(trees):
        tr = O.Tree(5)
        m, v = W.dense_to_batch(W.terrain_dense(W.terrain_heights(grid=2, chunk=32, seed=seed), t % 2, t // 2 % 2), 5)
        if vsize == 2:
            v = v.astype(np.uint16) * 300
        elif vsize == 4:
            v = v.astype(np.uint32) * np.uint32(2654435761)
        oracle_apply_dense(it, tr, m, v)
        ts.append(tr)
    return it, ts


def _section(it, trees, vsize):
    roots = [t.root for t in trees]
    leaves, branches = V.live_nodes(roots, it.children)
    pos = [(i, 0, -i) for i in range(len(trees))]
    return V.serialize(leaves, branches, roots, pos, it.children, it.get_value, vsize), leaves, branches, pos


def test_plain_checker_section_length():
    it = O.Interner(1, 64 * MB)
    t = O.Tree(5)
    m, v = W.dense_to_batch(W.plain_checker(5), 5)
    oracle_apply_dense(it, t, m, v)
    data, leaves, branches, _ = _section(it, [t], 1)
    assert (len(leaves), len(branches)) == (2, 5)
    # L u32 + 2 x (id, value) + B u32 + 5 x (id, mask, 8 child ids, lod) + count u32 + one chunk record
    assert len(data) == 4 + 2 * 2 + 4 + 5 * 11 + 4 + (12 + 12 + 1)
    assert struct.unpack_from(">I", data, 0)[0] == 2 and struct.unpack_from(">I", data, 8)[0] == 5
    lv, br, ch = V.deserialize(data, 1)
    assert sorted(lv.values()) == [1, 2] and all(m == 0xFF for m, _, _ in br.values())
    assert ch == [((0, 0, 0), 7)]  # the root is the last branch by node index


def test_roundtrip_rebuilds_every_tree():
    for vsize in (1, 2, 4):
        it, trees = _oracle_world(5, seed=3, vsize=vsize)
        data, leaves, branches, pos = _section(it, trees, vsize)
        parsed = V.deserialize(data, vsize)
        assert len(parsed[0]) == len(leaves) and len(parsed[1]) == len(branches)
        it2 = O.Interner(vsize, 64 * MB)
        roots, pos2 = V.rebuild(parsed, it2.get_or_create_leaf, it2.get_or_create_branch)
        assert [tuple(p) for p in pos2] == [tuple(p) for p in pos]
        for r, t in zip(roots, trees):
            t2 = O.Tree(5)
            t2.root = r
            assert t2.canonical_hash(it2) == t.canonical_hash(it)
        assert it2.counts() == (len(leaves), len(branches) + 1)


def test_planted_mistake_is_caught():
    """A section that lists a branch's children in reverse slot order rebuilds other trees."""
    it, trees = _oracle_world(2, seed=4)
    roots = [t.root for t in trees]
    leaves, branches = V.live_nodes(roots, it.children)
    rev = lambda b: list(reversed(it.children(b)))  # noqa: E731
    data = V.serialize(leaves, branches, roots, [(0, 0, 0)] * 2, rev, it.get_value, 1)
    it2 = O.Interner(1, 64 * MB)
    r2, _ = V.rebuild(V.deserialize(data, 1), it2.get_or_create_leaf, it2.get_or_create_branch)
    t2 = O.Tree(5)
    t2.root = r2[0]
    assert t2.canonical_hash(it2) != trees[0].canonical_hash(it)
