"""VTM snapshots on the GPU (vx_model_serialize / vx_model_deserialize) against the
oracle's restatement of VoxModel::serialize / deserialize (oracle/vtm_oracle.py).

Bar: a section the GPU writes, parsed and rebuilt by the oracle, gives every tree's
canonical hash; a section the oracle writes, imported by the GPU, gives the oracle's
hashes, reads and pattern counts, with reference counts = parent references + roots
(clearing every tree leaves only the EMPTY branch); export -> import -> export is
byte-identical; the imported interner keeps working (further applies match the
original world); the file wrapper checks magic, MD5 and refuses compressed data.
"""
import struct

import numpy as np
import pytest

import oracle_lib as O
from gpu_helpers import full_reads_gpu, full_reads_oracle, oracle_apply_dense
from oracle import vtm_oracle as V

pytestmark = pytest.mark.gpu

MB = 1024 * 1024


@pytest.fixture(scope="module")
def vx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    from voxelis_amd import api, vtm, workloads
    return torch, api, vtm, workloads


def _world(api, W, vsize, grid=3, seed=5):
    """A GPU forest and the same world in the oracle: terrain trees, one left empty, one
    filled with a single leaf."""
    n = grid * grid
    m, v, _, _ = W.terrain_batches(grid=grid, seed=seed)
    if vsize == 2:
        v = v.astype(np.uint16) * 257
    elif vsize == 4:  # i32 patterns over the whole 32-bit range
        v = v.astype(np.uint32) * np.uint32(2654435761)
    it = api.VoxInterner(vsize, 64 * MB)
    it_o = O.Interner(vsize, 64 * MB)
    f = api.Forest(it, [5] * n)
    hp = [i != 1 for i in range(n)]
    fills = [None] * n
    fills[2] = 9
    f.apply([api.to_device(x) for x in m], [api.to_device(x) for x in v], fills=fills,
            has_patches=[h and i != 2 for i, h in enumerate(hp)])
    trees = [O.Tree(5) for _ in range(n)]
    for i in range(n):
        if i == 1:
            continue
        if i == 2:
            trees[i].fill(it_o, 9)
            continue
        oracle_apply_dense(it_o, trees[i], m[i], v[i])
    pos = np.array([[i % grid, 0, i // grid] for i in range(n)], dtype=np.int32)
    return it, f, it_o, trees, pos


@pytest.mark.parametrize("vsize", [1, 2, 4])
def test_gpu_export_rebuilds_in_oracle(vx, vsize):
    torch, api, vtm, W = vx
    it, f, it_o, trees, pos = _world(api, W, vsize)
    data = f.serialize(pos)
    parsed = V.deserialize(data, vsize)
    leaves, branches = V.live_nodes([t.root for t in trees], it_o.children)
    assert (len(parsed[0]), len(parsed[1])) == (len(leaves), len(branches))
    assert it.counts() == (len(leaves), len(branches) + 1)
    it2 = O.Interner(vsize, 64 * MB)
    roots, pos2 = V.rebuild(parsed, it2.get_or_create_leaf, it2.get_or_create_branch)
    assert np.array_equal(np.array(pos2), pos)
    g = f.canonical_hashes()
    for i, r in enumerate(roots):
        t2 = O.Tree(5)
        t2.root = r
        assert t2.canonical_hash(it2) == trees[i].canonical_hash(it_o) == g[i], i


@pytest.mark.parametrize("vsize", [1, 2, 4])
def test_oracle_export_imports_on_gpu(vx, vsize):
    torch, api, vtm, W = vx
    _, _, it_o, trees, pos = _world(api, W, vsize, seed=6)
    roots = [t.root for t in trees]
    leaves, branches = V.live_nodes(roots, it_o.children)
    data = V.serialize(leaves, branches, roots, pos, it_o.children, it_o.get_value, vsize)
    it = api.VoxInterner(vsize, 64 * MB)
    f, pos2 = api.Forest.deserialize(it, data, 5)
    assert np.array_equal(pos2, pos)
    assert it.counts() == (len(leaves), len(branches) + 1)
    g = f.canonical_hashes()
    r = f.roots_host()
    for i, t in enumerate(trees):
        assert g[i] == t.canonical_hash(it_o), i
        assert np.array_equal(full_reads_gpu(f, i, 5), full_reads_oracle(it_o, t)), i
        if t.root:
            assert it.get_ref(int(r[i])) == sum(1 for x in roots if x == t.root) + \
                sum(1 for b in branches for c in it_o.children(b) if c == t.root)
    # every count is parent references + roots: clearing every tree frees everything
    f.fill(np.zeros(len(trees), dtype=np.uint64))
    assert it.counts() == (0, 1)


def test_roundtrip_is_byte_identical_and_keeps_working(vx):
    torch, api, vtm, W = vx
    it, f, it_o, trees, pos = _world(api, W, 1, seed=7)
    data = f.serialize(pos)
    it2 = api.VoxInterner(1, 64 * MB)
    f2, pos2 = api.Forest.deserialize(it2, data, 5)
    assert f2.serialize(pos2) == data
    # the imported interner is a working one: the same applies on both worlds agree
    # (some trees hit the leaf-brick panic, H1, on both: the statuses must agree)
    m, v, _, _ = W.terrain_batches(grid=3, seed=8)
    st = [ff.apply([api.to_device(x) for x in m], [api.to_device(x) for x in v]) for ff in (f, f2)]
    assert np.array_equal(st[0], st[1]) and (st[0] == 1).any()
    assert f.canonical_hashes() == f2.canonical_hashes()
    assert it.counts() == it2.counts()


def test_file_wrapper(vx, tmp_path):
    torch, api, vtm, W = vx
    it, f, it_o, trees, pos = _world(api, W, 1, seed=9)
    p = tmp_path / "world.vtm"
    n = vtm.export_model_to_vtm("terrain 3x3", p, f, pos, chunk_world_size=32.0, world_bounds=(2, 0, 2))
    raw = p.read_bytes()
    assert raw[:12] == b"VoxTreeModel" and struct.unpack(">H", raw[12:14])[0] == 0x0100
    meta, it2, f2, pos2 = vtm.import_model_from_vtm(p, 64 * MB)
    assert meta["name"] == "terrain 3x3" and meta["max_depth"] == 5 and meta["world_bounds"] == (2, 0, 2)
    assert meta["chunk_world_size"] == 32.0 and len(raw) == 42 + len("terrain 3x3") + 20 + n
    assert f2.canonical_hashes() == f.canonical_hashes()
    bad = bytearray(raw)
    bad[-1] ^= 0xFF
    p.write_bytes(bytes(bad))
    with pytest.raises(ValueError):
        vtm.import_model_from_vtm(p, 64 * MB)
    comp = bytearray(raw)
    comp[14:16] = struct.pack(">H", vtm.FLAG_COMPRESSED)
    p.write_bytes(bytes(comp))
    with pytest.raises(NotImplementedError):
        vtm.import_model_from_vtm(p, 64 * MB)


@pytest.mark.timeout(600)
def test_export_with_more_trees_and_nodes_than_a_scan_tile(vx):
    """2304 trees and more than 2048 nodes: every one of the export's five scans takes a
    second level of tile sums (2048 elements per tile), the most scratch arrays an export
    of this size allocates."""
    torch, api, vtm, W = vx
    it, f, it_o, trees, pos = _world(api, W, 2, grid=48, seed=9)
    assert len(trees) > 2048 and sum(it.counts()) > 2048
    data = f.serialize(pos)
    parsed = V.deserialize(data, 2)
    leaves, branches = V.live_nodes([t.root for t in trees], it_o.children)
    assert (len(parsed[0]), len(parsed[1])) == (len(leaves), len(branches))
    it2 = O.Interner(2, 64 * MB)
    roots, pos2 = V.rebuild(parsed, it2.get_or_create_leaf, it2.get_or_create_branch)
    assert np.array_equal(np.array(pos2), pos)
    g = f.canonical_hashes()
    for i in range(0, len(roots), 97):
        t2 = O.Tree(5)
        t2.root = roots[i]
        assert t2.canonical_hash(it2) == trees[i].canonical_hash(it_o) == g[i], i
