"""Four-byte voxels (VoxInterner<i32>, the type of every reference test and benchmark:
core/voxel.rs:93, voxtree_bench.rs:46) on the GPU against the oracle.  Values travel as
their 32-bit patterns (the path only compares values for equality and treats 0 as the
default), so negative i32 values are exercised through patterns with the top bit set.
The leaf table is hashed (interner/mod.rs:627-710): many distinct values, leaves created
and freed across applies, dry runs and rolled-back applies whose requested leaves must
leave the table again."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_helpers import assert_same, oracle_apply_dense
from test_gpu_parity import REFERENCE_BATCH_CASES, _check_tree, _run_tree_case

pytestmark = pytest.mark.gpu

MB = 1024 * 1024
K = np.uint32(2654435761)  # odd: v -> v*K is a bijection of the 32-bit patterns keeping 0


def i32(v):
    """Spread small test values over the whole 32-bit range (0 stays the default)."""
    return (np.asarray(v, dtype=np.uint64).astype(np.uint32) * K).astype(np.uint32)


@pytest.fixture(scope="module")
def vx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    from voxelis_amd import api, workloads
    return torch, api, workloads


@pytest.mark.parametrize("name", sorted(REFERENCE_BATCH_CASES))
def test_reference_batch_cases_i32(vx, name):
    _, api, _ = vx
    depth, steps = REFERENCE_BATCH_CASES[name]()
    steps = [(s[0], None if s[1] is None else int(i32([s[1]])[0]), s[2], i32(s[3])) if s[0] == "batch" else s
             for s in steps]
    it, t, it_o, t_o, err = _run_tree_case(api, depth, steps, value_size=4)
    assert err is None
    _check_tree(api, it, t, it_o, t_o)
    assert all(int(x) < 2 ** 32 for x in t.get_many(it, O.all_coords(2))[0])


def _shape(name, n):
    """The reference's criterion shapes (voxtree_bench.rs) as dense [y, z, x] i32 values."""
    y, z, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    if name == "set_uniform":
        return np.full((n, n, n), 1, dtype=np.int64)
    if name == "set_sum":
        return x + y + z + 1
    if name == "set_checkerboard":
        return ((x + y + z) % 2 == 0).astype(np.int64)
    return np.where((x % 4 == 0) & (y % 4 == 0) & (z % 4 == 0), 1, 0)  # set_sparse_fill


@pytest.mark.parametrize("name", ["set_uniform", "set_sum", "set_checkerboard", "set_sparse_fill"])
@pytest.mark.parametrize("depth", [5, 6])
def test_calibration_shapes_alternating_i32(vx, name, depth):
    """Two prebuilt batches applied alternately to one tree (voxtree_bench.rs:553-579),
    values negated in the second (i32 patterns with the top bit set)."""
    torch, api, W = vx
    dense = _shape(name, 1 << depth)
    a = dense.astype(np.int32).view(np.uint32)
    b = (-dense).astype(np.int32).view(np.uint32)
    it = api.VoxInterner(4, 256 * MB)
    it_o = O.Interner(4, 256 * MB)
    f = api.Forest(it, [depth])
    t_o = O.Tree(depth)
    bat = [W.dense_to_batch(x, depth) for x in (a, b)]
    dev = [(api.to_device(m), api.to_device(v)) for m, v in bat]
    for step in range(4):
        m, v = dev[step % 2]
        st = f.apply([m], [v])
        assert st[0] == 1, (step, st)
        oracle_apply_dense(it_o, t_o, *bat[step % 2])
        assert_same(it, f, it_o, [t_o], full_reads=(step >= 2), what=f"{name} {depth} step {step}")


def test_random_values_churn_dry_run_and_rollback_i32(vx):
    """Sparse random 32-bit values over several trees and applies (leaves created and
    freed every step), an exact dry-run count, and an apply rolled back for its node limit
    whose requested leaves leave the table (a retry then matches the oracle)."""
    torch, api, W = vx
    rng = np.random.default_rng(77)
    depths = [4, 5, 3, 5, 4, 2]
    it = api.VoxInterner(4, 64 * MB)
    it_o = O.Interner(4, 64 * MB)
    f = api.Forest(it, depths)
    trees_o = [O.Tree(d) for d in depths]

    def batches():
        xyzs, vals, offs = [], [], [0]
        for d in depths:
            k = int(rng.integers(1, 400))
            xyzs.append(rng.integers(0, 1 << d, size=(k, 3)).astype(np.int32))
            vals.append(rng.integers(1, 1 << 32, size=k, dtype=np.uint64).astype(np.uint32))
            offs.append(offs[-1] + k)
        return xyzs, vals, offs

    def oracle_apply(xyzs, vals):
        for t, xyz, v in zip(trees_o, xyzs, vals):
            b = O.Batch(t.depth)
            b.set_many(xyz, v.astype(np.uint64))
            assert t.apply_batch(it_o, b)

    for step in range(6):
        xyzs, vals, offs = batches()
        args = (offs, api.to_device(np.concatenate(xyzs)), api.to_device(np.concatenate(vals)))
        if step == 3:
            before = (it.counts(), f.canonical_hashes())
            st = f.apply_edits(*args, dry_run=True, counts=True)
            assert (st == 1).all()
            need = f.counts_host()
            assert (it.counts(), f.canonical_hashes()) == before
            st = f.apply_edits(*args, node_limit=max(1, need[0] + need[1] - 3))
            assert (st == -8).all(), st
            assert (it.counts(), f.canonical_hashes()) == before
        st = f.apply_edits(*args)
        assert (st == 1).all(), (step, st)
        oracle_apply(xyzs, vals)
        assert_same(it, f, it_o, trees_o, full_reads=(step % 2 == 1), what=f"step {step}")


def test_fill_set_and_to_vec_i32(vx):
    torch, api, W = vx
    it = api.VoxInterner(4, 64 * MB)
    it_o = O.Interner(4, 64 * MB)
    depths = [3, 4]
    f = api.Forest(it, depths)
    trees_o = [O.Tree(d) for d in depths]
    fv = [int(i32([5])[0]), int(i32([0x80000001])[0])]
    f.fill(fv)
    for t, v in zip(trees_o, fv):
        t.fill(it_o, v)
    assert_same(it, f, it_o, trees_o, what="fill")
    rng = np.random.default_rng(3)
    for rnd in range(8):
        xyz = np.stack([rng.integers(0, 1 << d, 3) for d in depths]).astype(np.int32)
        vals = i32(rng.integers(0, 4, len(depths)))  # includes 0: remove
        st = f.set_voxels(torch.from_numpy(xyz.reshape(-1)).cuda(), api.to_device(vals))
        for t, t_o in enumerate(trees_o):
            ch = t_o.set(it_o, *(int(c) for c in xyz[t]), int(vals[t]))
            assert st[t] == (1 if ch else 0), (rnd, t, st[t], ch)
        assert_same(it, f, it_o, trees_o, what=f"set round {rnd}")
    vec = f.to_vec(0)
    for t, v in zip(trees_o, vec):
        assert np.array_equal(v.reshape(-1).astype(np.uint64), t.to_vec(it_o).reshape(-1))
