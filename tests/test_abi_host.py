"""CPU-side checks of the product: the C-ABI library loads and exports every symbol
include/voxelis_hip.h declares (no compute calls here), and the host logic that
prepares device inputs agrees with the oracle's Batch."""
import ctypes
import os
import re

import numpy as np

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "voxelis_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from voxelis_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 19
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the ctypes table covers the whole header
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)
    assert _lib.lib().vx_status_string(-6) == b"apply produced the current root"


def test_ctypes_mirrors_the_header_structs():
    """The Python mirrors of vx_tuning and of the VX_STAT_* list follow the header."""
    from voxelis_amd import _lib
    text = open(os.path.join(ROOT, "include", "voxelis_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = re.search(r"typedef struct \{([^}]*)\} vx_tuning;", text).group(1)
    assert re.findall(r"uint32_t (\w+);", body) == [f for f, _ in _lib.Tuning._fields_]
    assert ctypes.sizeof(_lib.Tuning) == 4 * len(_lib.Tuning._fields_)
    stats = re.search(r"enum \{\s*(VX_STAT_\w+ = 0,.*?)VX_STAT_COUNT", text, re.S).group(1)
    assert len(re.findall(r"VX_STAT_\w+", stats)) == len(_lib.STAT_NAMES)


def test_morton_encoder_matches_oracle():
    from voxelis_amd.morton import encode_path
    rng = np.random.default_rng(5)
    xyz = rng.integers(0, 1024, size=(2000, 3)).astype(np.int32)
    ours = encode_path(xyz)
    ref = np.array([O.encode_path(*map(int, p)) for p in xyz], dtype=np.uint32)
    assert np.array_equal(ours, ref)


def test_dense_batch_matches_oracle_batch():
    # last-write-wins with duplicates and clears (batch.rs:145-175)
    from voxelis_amd.morton import apply_edits_to_dense, dense_batch_arrays
    rng = np.random.default_rng(9)
    xyz = rng.integers(0, 16, size=(3000, 3)).astype(np.int32)
    vals = rng.integers(0, 5, size=3000)  # includes the default value (clears)
    masks, values = dense_batch_arrays(4, np.uint8)
    apply_edits_to_dense(masks, values, xyz, vals)
    b = O.Batch(4)
    b.set_many(xyz, vals.astype(np.uint64))
    assert np.array_equal(masks, b.masks())
    assert np.array_equal(values.astype(np.uint64), b.values())


def test_workload_batches_match_oracle_just_set():
    from voxelis_amd import workloads as W
    dense = W.plain_checker(3)
    masks, values = W.dense_to_batch(dense, 3)
    b = O.Batch(3)
    xyz = O.all_coords(3)
    b.set_many(xyz, dense[xyz[:, 1], xyz[:, 2], xyz[:, 0]].astype(np.uint64))
    assert np.array_equal(masks, b.masks())
    assert np.array_equal(values.astype(np.uint64), b.values())


def test_terrain_heights_deterministic_and_in_range():
    from voxelis_amd import workloads as W
    h1 = W.terrain_heights(grid=4, seed=11)
    h2 = W.terrain_heights(grid=4, seed=11)
    assert np.array_equal(h1, h2)
    assert h1.min() >= 8 and h1.max() <= 23
    assert len(np.unique(h1)) > 4


def test_committed_tuning_picks_are_valid_launch_shapes():
    """The autotuner's committed picks (voxelis_amd/tuning.json, loaded by bench.py) hold
    values vx_interner_set_tuning accepts, and name the sweep that chose them."""
    import os
    from voxelis_amd.autotune import TUNING_FIELDS, committed_pick
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for name in ("uneven_4096_u16", "terrain_4096x32_u8"):
        p = committed_pick(name)
        assert p is not None, name
        assert set(TUNING_FIELDS) <= set(p)
        assert 0.0 < p["load_factor"] < 1.0
        assert p["tile"] in (128, 256, 512)
        assert 0 <= p["dedup_rounds"] <= 16 and 0 <= p["brick_merge_rounds"] <= 16
        assert p["brick_rounds"] in (0, 1, 2, 4)
        assert 0 < p["commit_grid"] <= 1 << 20 and 0 < p["gc_grid"] <= 1 << 20
        assert os.path.exists(os.path.join(root, p["sweep"])), p["sweep"]
