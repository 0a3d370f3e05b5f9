"""Equal brick keys are merged per wavefront before the probe (one candidate per key
and tree), so a tree that fails in the brick pass drops only its own candidates.

Sixteen depth-2 trees (8 bricks each) share wavefronts.  Tree 0 partially edits a
brick that its first apply collapsed into a leaf (the reference's panic at
block_id.rs:273-276, reached from voxtree.rs:829-834) while every tree writes the
same brick patterns elsewhere: tree 0 is left as it was (VX_E_LEAF_BRICK) and every
other tree matches the oracle, for one and for four brick slots per lane.
"""
import numpy as np
import pytest

import oracle_lib as O
from gpu_helpers import assert_same, oracle_apply_dense

pytestmark = pytest.mark.gpu

MB = 1024 * 1024


@pytest.fixture(scope="module")
def vx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    from voxelis_amd import api, workloads
    return torch, api, workloads


def _dense_steps():
    """(first apply, second apply of tree 0, second apply of the others); (y, z, x) order."""
    first = np.zeros((4, 4, 4), dtype=np.uint8)
    first[0:2, 0:2, 0:2] = 5  # brick 0 uniform: a leaf at brick level
    first[3, 3, 3] = 1
    first[3, 3, 2] = 2  # both values' leaves exist before the second apply (no deferred brick)
    others = np.zeros((4, 4, 4), dtype=np.uint8)
    for y in range(4):
        for z in range(4):
            for x in range(4):
                if x >= 2 or y >= 2 or z >= 2:
                    others[y, z, x] = 1 + (x & 1)  # the same mixed pattern in bricks 1..7
    tree0 = others.copy()
    tree0[0, 0, 0] = 6  # one voxel of the collapsed brick: the leaf-brick panic
    return first, tree0, others


@pytest.mark.parametrize("brick_rounds", [1, 4])
def test_failed_tree_keeps_other_trees_keys(vx, brick_rounds):
    torch, api, W = vx
    n, depth = 16, 2
    it = api.VoxInterner(1, 64 * MB)
    it.set_tuning(brick_rounds=brick_rounds)
    it_o = O.Interner(1, 64 * MB)
    f = api.Forest(it, [depth] * n)
    trees_o = [O.Tree(depth) for _ in range(n)]
    first, tree0, others = _dense_steps()
    m1, v1 = W.dense_to_batch(first, depth)
    st = f.apply([api.to_device(m1)] * n, [api.to_device(v1)] * n)
    assert (st == 1).all(), st
    for t in trees_o:
        assert oracle_apply_dense(it_o, t, m1, v1)[0]
    assert_same(it, f, it_o, trees_o, what="first apply")
    roots_before = f.roots_host()
    m0, v0 = W.dense_to_batch(tree0, depth)
    mo, vo = W.dense_to_batch(others, depth)
    st = f.apply([api.to_device(m0)] + [api.to_device(mo)] * (n - 1),
                 [api.to_device(v0)] + [api.to_device(vo)] * (n - 1))
    assert st[0] == -5 and (st[1:] == 1).all(), st
    with pytest.raises(O.OraclePanic):
        oracle_apply_dense(it_o, trees_o[0], m0, v0)
    for t in trees_o[1:]:
        assert oracle_apply_dense(it_o, t, mo, vo)[0]
    assert f.roots_host()[0] == roots_before[0]
    assert_same(it, f, it_o, trees_o, what="second apply")
    assert it.stats()["merged"] > 0  # equal keys of one tree did share candidates
