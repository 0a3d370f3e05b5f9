"""Exact node counts, reservations and rollback of the apply (through the C ABI).

- A dry run changes nothing and counts exactly the nodes the real apply creates
  (checked against the real apply's own allocation count, and against the
  oracle's get_or_create misses where the reference frees nothing transiently).
- An apply that exceeds its node_limit, or the interner's capacity, is rolled
  back: every tree keeps its root, the interner keeps its patterns and reference
  counts (a later clear of every tree leaves exactly what the oracle leaves), and
  the interner stays usable.
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
    return api, workloads


def _batches(api, W, grid):
    """Step A: terrain; step B: the same terrain with the ids swapped (so no batch edits
    part of a collapsed brick, which the reference rejects), and fills on some trees
    (fill + patches on trees 0-3, fill only on 4-6)."""
    n = grid * grid
    mA, vA, _, _ = W.terrain_batches(grid=grid, seed=21)
    mB = mA
    vB = vA.copy()
    vB[vA == 1] = 2
    vB[vA == 2] = 1
    fills = [None] * n
    patches = [True] * n
    for i in range(4):
        fills[i] = 5
    fills[4], fills[5], fills[6] = 7, 9, 11
    patches[4] = patches[5] = patches[6] = False
    dev = lambda arrs: [api.to_device(a) for a in arrs]
    return (dev(mA), dev(vA), mA, vA), (dev(mB), dev(vB), mB, vB, fills, patches)


def _oracle_step(it_o, trees_o, m, v, fills=None, patches=None):
    for i, t in enumerate(trees_o):
        f = None if fills is None else fills[i]
        if patches is not None and not patches[i]:
            b = t.create_batch()
            b.just_fill(f)
            t.apply_batch(it_o, b)
        else:
            oracle_apply_dense(it_o, t, m[i], v[i], fill=f)


def _misses(it_o):
    s = it_o.stats()
    return s["leaf_misses"] + s["branch_misses"], s["transient_free"]


def _state(it, f):
    return (it.counts(), it.alive_nodes(), f.roots_host(), f.canonical_hashes())


@pytest.mark.parametrize("brick_rounds", [0, 4])
def test_dry_run_counts_exactly_and_changes_nothing(vx, brick_rounds):
    api, W = vx
    grid = 4
    n = grid * grid
    (mA, vA, hmA, hvA), (mB, vB, hmB, hvB, fills, patches) = _batches(api, W, grid)
    it = api.VoxInterner(1, 64 * MB)
    it.set_tuning(brick_rounds=brick_rounds)  # 0: one brick slot per lane at this size
    it_o = O.Interner(1, 64 * MB)
    f = api.Forest(it, [5] * n)
    trees_o = [O.Tree(5) for _ in range(n)]
    # A on empty trees: dry, then real
    empty = _state(it, f)
    st_dry = f.apply(mA, vA, dry_run=True, counts=True)
    dA = f.counts_host()
    assert _state(it, f) == empty
    st = f.apply(mA, vA, counts=True)
    rA = f.counts_host()
    assert (st == st_dry).all() and (st == 1).all()
    assert dA[0] == rA[0] and dA[1] == rA[1], (dA[:2], rA[:2])
    # per tree: a node shared by several trees counts for whichever tree inserted it first
    assert dA[2].sum() == dA[0] and rA[2].sum() == rA[0]
    m0, tr0 = _misses(it_o)
    _oracle_step(it_o, trees_o, hmA, hvA)
    m1, tr1 = _misses(it_o)
    if tr1 == tr0:
        assert m1 - m0 == rA[0] + rA[1]
    assert_same(it, f, it_o, trees_o, full_reads=False, what="A")
    # B over A: the dry run leaves every root, pattern and count as it was
    before = _state(it, f)
    st_dry = f.apply(mB, vB, fills=fills, has_patches=patches, dry_run=True, counts=True)
    dB = f.counts_host()
    assert _state(it, f) == before
    st = f.apply(mB, vB, fills=fills, has_patches=patches, counts=True)
    rB = f.counts_host()
    assert np.array_equal(st, st_dry), (st, st_dry)
    assert dB[0] == rB[0] and dB[1] == rB[1], (dB[:2], rB[:2])
    assert dB[2].sum() == dB[0] and rB[2].sum() == rB[0]
    m0, tr0 = _misses(it_o)
    _oracle_step(it_o, trees_o, hmB, hvB, fills, patches)
    m1, tr1 = _misses(it_o)
    if tr1 == tr0:
        assert m1 - m0 == rB[0] + rB[1]
    assert_same(it, f, it_o, trees_o, full_reads=True, what="B")


def _clear_all_and_compare(it, f, it_o, trees_o):
    f.fill([0] * f.n)
    for t in trees_o:
        t.clear(it_o)
    assert it.counts() == it_o.counts()
    assert f.roots_host() == [0] * f.n


@pytest.mark.parametrize("brick_rounds", [0, 4])
def test_node_limit_rolls_back(vx, brick_rounds):
    api, W = vx
    grid = 4
    n = grid * grid
    (mA, vA, hmA, hvA), (mB, vB, hmB, hvB, fills, patches) = _batches(api, W, grid)
    it = api.VoxInterner(1, 64 * MB)
    it.set_tuning(brick_rounds=brick_rounds)
    it_o = O.Interner(1, 64 * MB)
    f = api.Forest(it, [5] * n)
    trees_o = [O.Tree(5) for _ in range(n)]
    assert (f.apply(mA, vA) == 1).all()
    _oracle_step(it_o, trees_o, hmA, hvA)
    f.apply(mB, vB, fills=fills, has_patches=patches, dry_run=True, counts=True)
    b, l, _ = f.counts_host()
    need = b + l
    assert need > 1
    before = _state(it, f)
    stats_before = it.stats()
    st = f.apply(mB, vB, fills=fills, has_patches=patches, node_limit=need - 1, counts=True)
    assert (st == -8).all(), st
    assert f.counts_host()[0] is None
    assert _state(it, f) == before
    assert it.stats()["recycled"] == stats_before["recycled"]
    # the same apply with the exact reservation goes through and matches the oracle
    st = f.apply(mB, vB, fills=fills, has_patches=patches, node_limit=need, counts=True)
    assert (st == 1).all(), st
    assert f.counts_host()[:2] == (b, l)
    _oracle_step(it_o, trees_o, hmB, hvB, fills, patches)
    assert_same(it, f, it_o, trees_o, full_reads=True, what="after rollback")
    # reference counts survived the rollback: clearing every tree frees the same nodes
    _clear_all_and_compare(it, f, it_o, trees_o)


@pytest.mark.parametrize("brick_rounds", [0, 4])
def test_capacity_oom_rolls_back_and_interner_stays_usable(vx, brick_rounds):
    api, W = vx
    grid = 4
    n = grid * grid
    (mA, vA, hmA, hvA), (mB, vB, hmB, hvB, fills, patches) = _batches(api, W, grid)
    # size the interner so that A fits and B does not
    probe = api.VoxInterner(1, 64 * MB)
    fp = api.Forest(probe, [5] * n)
    assert (fp.apply(mA, vA) == 1).all()
    alive_a = probe.alive_nodes()
    fp.apply(mB, vB, fills=fills, has_patches=patches, dry_run=True, counts=True)
    b, l, _ = fp.counts_host()
    probe.close()
    cap = alive_a + b + l - 1
    budget = cap * (78 + 1)
    it = api.VoxInterner(1, budget)
    it.set_tuning(brick_rounds=brick_rounds)
    assert it.capacity() == cap
    it_o = O.Interner(1, budget)
    f = api.Forest(it, [5] * n)
    trees_o = [O.Tree(5) for _ in range(n)]
    assert (f.apply(mA, vA) == 1).all()
    _oracle_step(it_o, trees_o, hmA, hvA)
    before = _state(it, f)
    st = f.apply(mB, vB, fills=fills, has_patches=patches)
    assert (st == -2).all(), st
    assert _state(it, f) == before
    it.sync()  # not poisoned
    # a smaller apply (B on the first half of the trees only) still fits
    half = n // 2
    fh = api.Forest(it, [5] * half)
    fh.roots.copy_(f.roots[:half])
    st = fh.apply(mB[:half], vB[:half], fills=fills[:half], has_patches=patches[:half])
    assert (st == 1).all(), st
    f.roots[:half] = fh.roots
    _oracle_step(it_o, trees_o[:half], hmB[:half], hvB[:half], fills[:half], patches[:half])
    assert_same(it, f, it_o, trees_o, full_reads=False, what="after OOM")
    _clear_all_and_compare(it, f, it_o, trees_o)


def test_edit_list_dry_run_matches_real(vx):
    api, W = vx
    import torch
    from voxelis_amd import serve
    world = serve.World(2, 256 * MB, (4, 4, 2), depth=5)
    rng = np.random.default_rng(5)
    for rnd in range(3):
        # a few spheres, some carving (value 0), as one edit-list apply
        per_tree = {}
        for _ in range(4):
            c = (rng.random(3) * np.array([128, 128, 64])).astype(np.int64)
            val = 0 if (rnd > 0 and rng.random() < 0.5) else int(rng.integers(1, 65536))
            for t, (loc, v) in serve.sphere_edits(world, c, float(rng.uniform(3, 20)), val).items():
                a = per_tree.setdefault(t, ([], []))
                a[0].append(loc)
                a[1].append(v)
        trees = sorted(per_tree)
        offs = [0]
        xyz, vals = [], []
        for t in trees:
            loc = np.concatenate(per_tree[t][0])
            v = np.concatenate(per_tree[t][1])
            xyz.append(loc)
            vals.append(v)
            offs.append(offs[-1] + len(v))
        xyz_t = api.to_device(np.concatenate(xyz).astype(np.int32))
        val_t = api.to_device(np.concatenate(vals).astype(np.uint16))
        idx = torch.tensor(trees, dtype=torch.int64, device=world.dev)
        sub = api.Forest(world.it, [5] * len(trees))
        sub.roots.copy_(world.forest.roots[idx])
        before = (world.it.counts(), sub.roots_host())
        st_d = sub.apply_edits(offs, xyz_t, val_t, dry_run=True, counts=True)
        d = sub.counts_host()
        assert (world.it.counts(), sub.roots_host()) == before
        st = sub.apply_edits(offs, xyz_t, val_t, counts=True)
        r = sub.counts_host()
        assert np.array_equal(st, st_d)
        assert d[:2] == r[:2] and d[2].sum() == d[0] and r[2].sum() == r[0], (d[:2], r[:2])
        world.forest.roots[idx] = sub.roots


def test_batch_admission_predicts_and_falls_back_exactly(vx):
    """Predictor-driven admission (voxelis_amd/admission.py): an untrained network with a
    margin far too small under-predicts, the library rolls the apply back (VX_E_BUDGET) and
    the exact dry-run recount admits it; a generous margin admits directly.  Either way
    the trees equal the oracle's after every apply."""
    api, W = vx
    from voxelis_amd.admission import BatchAdmission, batch_features, with_state
    from voxelis_amd.predictor import GrowthMLP, Predictor
    (mA_t, vA_t, mA, vA), (mB_t, vB_t, mB, vB, _, _) = _batches(api, W, 4)
    n = len(mA)
    static = np.stack([batch_features(mA[i], vA[i], 5, 1) for i in range(n)])
    pred = Predictor(GrowthMLP(seed=3), np.zeros(16, np.float32), np.ones(16, np.float32))
    it = api.VoxInterner(1, 64 * MB)
    f = api.Forest(it, [5] * n)
    it_o = O.Interner(1, 64 * MB)
    trees_o = [O.Tree(5) for _ in range(n)]
    views = [f.make_views(mA_t, vA_t), f.make_views(mB_t, vB_t)]
    last = np.zeros(n)
    for k, (margin, want_fallback) in enumerate(((1e-6, True), (1e-6, True), (1e6, False))):
        adm = BatchAdmission(pred, margin, 1)
        adm.value_size = 0  # no leaf allowance: the tiny margin must under-predict
        X = with_state(static, f.roots.cpu().numpy() == 0, last)
        per = adm.apply(f, views[k % 2], X)
        assert adm.fallbacks == (1 if want_fallback else 0), k
        assert (f.status.cpu().numpy() == 1).all()
        m, v = (mA, vA) if k % 2 == 0 else (mB, vB)
        _oracle_step(it_o, trees_o, m, v)
        assert_same(it, f, it_o, trees_o, full_reads=(k == 2), what=f"admitted apply {k}")
        assert per.sum() == adm.last[0]  # the per-tree counts add up to the apply's new branches
        last = per
    it.close()
