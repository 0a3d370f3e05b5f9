"""Batch features of the growth predictor (voxelis_amd/admission.py) against brute force
on small batches: touched node positions per height from the Morton paths of the written
voxels, complete and uniform bricks, distinct bricks, and the hard bound."""
import math

import numpy as np

from voxelis_amd.admission import batch_features, leaf_allowance, with_state
from voxelis_amd.morton import apply_edits_to_dense, dense_batch_arrays, encode_path


def _brute(depth, xyz, vals):
    """(set, clear, positions per height, full, uniform, distinct, bound) from the edit list
    with the last write to a voxel winning, as Batch::just_set records it."""
    last = {}
    for (x, y, z), v in zip(xyz.tolist(), vals.tolist()):
        last[(x, y, z)] = v
    bricks = {}
    for (x, y, z), v in last.items():
        m = int(encode_path(np.array([[x, y, z]], dtype=np.int32))[0])
        b = bricks.setdefault(m >> 3, {})
        b[m & 7] = v
    nset = sum(1 for v in last.values() if v != 0)
    nclr = sum(1 for v in last.values() if v == 0)
    per_h = []
    for h in range(1, depth + 1):
        per_h.append(len({p >> (3 * (h - 1)) for p in bricks}))
    full = [b for b in bricks.values() if len(b) == 8 and all(v != 0 for v in b.values())]
    uni = [b for b in full if len(set(b.values())) == 1]
    keys = set()
    for b in bricks.values():
        sm = sum(1 << s for s, v in b.items() if v != 0)
        cm = sum(1 << s for s, v in b.items() if v == 0)
        keys.add((sm, cm) + tuple(b.get(s, 0) for s in range(8)))
    return nset, nclr, per_h, len(full) / len(bricks), len(uni) / len(bricks), len(keys), sum(per_h)


def test_batch_features_match_brute_force():
    rng = np.random.default_rng(4)
    for depth, n, vmax in ((3, 300, 3), (4, 50, 200), (5, 3000, 2), (2, 64, 1)):
        side = 1 << depth
        xyz = rng.integers(0, side, size=(n, 3)).astype(np.int32)
        vals = rng.integers(0, vmax + 1, size=n).astype(np.uint16)
        m, v = dense_batch_arrays(depth, np.uint16)
        apply_edits_to_dense(m, v, xyz, vals)
        f = batch_features(m, v, depth, 2)
        nset, nclr, per_h, full, uni, distinct, bound = _brute(depth, xyz, vals)
        assert f[0] == np.float32(math.log1p(nset)) and f[1] == np.float32(math.log1p(nclr))
        assert f[2] == depth
        for h in range(1, min(depth, 6) + 1):
            assert f[2 + h] == np.float32(math.log1p(per_h[h - 1])), (depth, h)
        assert abs(f[9] - full) < 1e-6 and abs(f[10] - uni) < 1e-6
        assert f[11] == np.float32(math.log1p(distinct))
        assert f[14] == 2
        assert f[15] == np.float32(math.log1p(bound))


def test_empty_batch_and_state_features():
    m, v = dense_batch_arrays(4, np.uint8)
    f = batch_features(m, v, 4, 1)
    assert f[0] == 0 and f[1] == 0 and f[15] == 0 and f[14] == 1 and f[2] == 4
    X = with_state(np.stack([f, f]), [True, False], [0, 41])
    assert X[0, 12] == 1 and X[1, 12] == 0
    assert X[0, 13] == 0 and X[1, 13] == np.float32(math.log1p(41))
    assert f[12] == 0  # the batch features are left unchanged
    assert leaf_allowance(1) == 256 and leaf_allowance(2) == 65536 and leaf_allowance(4, 17) == 17
