"""vx_model_serialize_noalloc, called through the C ABI, against vx_model_serialize:
the same bytes on two calls in a row (the second with a smaller world, so the shared
scratch holds the first call's contents), and a world too large for the fixed scratch
rejected with VX_E_SCRATCH, leaving the interner as it was."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MB = 1024 * 1024


@pytest.fixture(scope="module")
def vx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    from voxelis_amd import api, workloads
    from voxelis_amd._lib import lib
    return torch, api, workloads, lib


def _terrain_world(api, W, grid, seed, vsize):
    m, v, _, _ = W.terrain_batches(grid=grid, seed=seed)
    if vsize == 2:
        v = v.astype(np.uint16) * 257
    it = api.VoxInterner(vsize, 256 * MB)
    f = api.Forest(it, [5] * (grid * grid))
    st = f.apply([api.to_device(x) for x in m], [api.to_device(x) for x in v])
    assert (st == 1).all()
    pos = np.array([[i % grid, 0, i // grid] for i in range(grid * grid)], dtype=np.int32)
    return it, f, pos


def _call(fn, it, f, pos, cap):
    pos = np.ascontiguousarray(pos, dtype=np.int32).reshape(-1)
    buf = (C.c_uint8 * max(cap, 1))()
    size = C.c_size_t(0)
    rc = fn(it.h, f.n, f.roots.data_ptr(), pos.ctypes.data, buf if cap else None, cap, C.byref(size))
    return rc, size.value, bytes(memoryview(buf)[:size.value]) if cap else b""


def test_noalloc_export_equals_export_twice(vx):
    torch, api, W, lib = vx
    big = _terrain_world(api, W, grid=8, seed=5, vsize=1)
    small = _terrain_world(api, W, grid=3, seed=9, vsize=2)
    for it, f, pos in (big, small, big):
        ref = f.serialize(pos)
        rc, size, _ = _call(lib().vx_model_serialize_noalloc, it, f, pos, 0)  # size query
        assert rc == 0 and size == len(ref)
        rc, size, data = _call(lib().vx_model_serialize_noalloc, it, f, pos, size)
        assert rc == 0 and data == ref
        rc, _, _ = _call(lib().vx_model_serialize_noalloc, it, f, pos, size - 1)
        assert rc == -4  # VX_E_RANGE: the caller's buffer is too small


def test_noalloc_export_rejects_a_world_beyond_its_scratch(vx):
    torch, api, W, lib = vx
    rng = np.random.default_rng(4)
    n, d, k = 96, 6, 100000  # random u16 values: about 35k branch nodes per tree, 3.4 M in all
    it = api.VoxInterner(2, 4096 * MB)
    f = api.Forest(it, [d] * n)
    xyz = np.concatenate([rng.integers(0, 64, size=(k, 3)) for _ in range(n)]).astype(np.int32)
    vals = rng.integers(1, 65536, size=len(xyz)).astype(np.uint16)
    st = f.apply_edits(np.arange(n + 1, dtype=np.uint64) * k, api.to_device(xyz), api.to_device(vals))
    assert (st == 1).all()
    assert it.alive_nodes() > 3_000_000
    before = (it.counts(), it.alive_nodes(), f.canonical_hashes())
    pos = np.zeros((n, 3), dtype=np.int32)
    rc, size, _ = _call(lib().vx_model_serialize_noalloc, it, f, pos, 0)
    assert rc == -12 and size > 32 * MB  # VX_E_SCRATCH, with the size it would take
    rc, _, _ = _call(lib().vx_model_serialize_noalloc, it, f, pos, size)
    assert rc == -12
    assert (it.counts(), it.alive_nodes(), f.canonical_hashes()) == before
    # the allocating entry point exports the same world
    assert len(f.serialize(pos)) > 0
    # and the no-allocation one still works on a world that fits
    it2, f2, pos2 = _terrain_world(api, W, grid=2, seed=3, vsize=1)
    rc, size, _ = _call(lib().vx_model_serialize_noalloc, it2, f2, pos2, 0)
    rc, size, data = _call(lib().vx_model_serialize_noalloc, it2, f2, pos2, size)
    assert rc == 0 and data == f2.serialize(pos2)
