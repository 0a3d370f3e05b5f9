"""Pins the oracle's core types against the reference's own known answers.

BlockId: block_id.rs doctest golden raw value (405-412) and unit tests (474-609).
PATH_MASKS: the reference table voxtree.rs:41-105 (tests/golden/path_masks.json).
Morton path: brute-force bit interleave (utils/common.rs:24-55 semantics).
calc_average: brute-force mode definition (core/voxel.rs:96-141).
Allocator: pool_allocator_lite.rs tests (LIFO reuse, OOM) through the interner.
"""
import itertools
import json
import os

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_blockid_golden_raw():
    # block_id.rs:405-412
    assert O.bid_new_leaf(123, 456) == 0x800001C80000007B


def test_blockid_fields():
    b = O.bid_new_branch(789, 1011, 0xAB, 0xCD)
    assert b & 0xFFFFFFFF == 789
    assert (b >> 32) & 0x7FFF == 1011
    assert (b >> 55) & 0xFF == 0xAB
    assert (b >> 47) & 0xFF == 0xCD
    assert b >> 63 == 0
    # has_child(0) true, has_child(1) false for mask 0xCD (block_id.rs:308-310)
    assert ((b >> 47) & 0xFF) & 1 and not ((b >> 47) & 0xFF) & 2
    # raw roundtrip value of new_branch(123,456,0xAA,0x55) (block_id.rs:536-541)
    r = O.bid_new_branch(123, 456, 0xAA, 0x55)
    assert r == (0xAA << 55) | (0x55 << 47) | (456 << 32) | 123
    # max values are not INVALID (block_id.rs:518-533)
    m = O.bid_new_branch(0xFFFFFFFF, 0x7FFE, 0xFF, 0xFF) | (1 << 63)
    assert m != O.INVALID
    # EMPTY is the branch with index 0, generation 0, no children
    assert O.bid_new_branch(0, 0, 0, 0) == O.EMPTY
    # generation overflow is rejected (block_id.rs:217, 565-567)
    assert O.bid_new_leaf(0, 0x7FFF) == O.INVALID


def test_path_masks_match_reference_table():
    table = json.load(open(os.path.join(GOLDEN, "path_masks.json")))["path_masks"]
    for d in range(7):
        for k in range(6):
            assert O.path_mask(d, k) == table[d][k], (d, k)


def _interleave(x, y, z):
    r = 0
    for b in range(10):
        r |= ((x >> b) & 1) << (3 * b)
        r |= ((y >> b) & 1) << (3 * b + 1)
        r |= ((z >> b) & 1) << (3 * b + 2)
    return r


def test_morton_path_bruteforce():
    rng = np.random.default_rng(7)
    pts = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1023, 1023, 1023), (5, 9, 17)]
    pts += [tuple(int(v) for v in rng.integers(0, 1024, 3)) for _ in range(500)]
    for x, y, z in pts:
        assert O.encode_path(x, y, z) == _interleave(x, y, z)
    # the low digit is the slot inside a 2x2x2 brick: x | y<<1 | z<<2
    for x, y, z in itertools.product(range(2), repeat=3):
        assert O.encode_path(x, y, z) == x | (y << 1) | (z << 2)


def _mode_reference(vals):
    # restated from the definition: most frequent value, first seen wins ties,
    # but a non-default value displaces a default leader of equal count
    order, counts = [], {}
    for v in vals:
        if v not in counts:
            order.append(v)
            counts[v] = 0
        counts[v] += 1
    best = order[0]
    for v in order[1:]:
        if counts[v] > counts[best] or (counts[v] == counts[best] and best == 0 and v != 0):
            best = v
    return best


def test_calc_average_bruteforce():
    # exhaustive over 8 children drawn from {0,1,2} (6561 cases)
    for vals in itertools.product(range(3), repeat=8):
        assert O.calc_average(list(vals)) == _mode_reference(vals), vals
    # specific tie rules
    assert O.calc_average([0, 0, 0, 0, 5, 5, 5, 5]) == 5
    assert O.calc_average([5, 5, 5, 5, 0, 0, 0, 0]) == 5
    assert O.calc_average([3, 3, 3, 3, 5, 5, 5, 5]) == 3
    assert O.calc_average([0] * 8) == 0


def test_interner_capacity_formula():
    # interner/mod.rs:49-53, 158-164: node = 4 + 2 + 64 + sizeof(T) + 8
    it = O.Interner(1, 256 * 1024 * 1024)
    assert it.capacity() == (256 * 1024 * 1024) // 79 == 3397917
    it2 = O.Interner(2, 64 * 1024 ** 3)
    assert it2.capacity() == (64 * 1024 ** 3) // 80 == 858993459
    with pytest.raises(O.OraclePanic):
        O.Interner(1, 10)


def test_interner_empty_state():
    it = O.Interner(1, 1024)
    assert it.counts() == (0, 1)  # EMPTY branch pattern pre-installed (mod.rs:97)


def test_allocator_lifo_reuse_and_oom():
    # pool_allocator_lite.rs tests: LIFO reuse order, OOM panic at capacity
    it = O.Interner(1, 79 * 4)  # capacity 4: EMPTY + 3 nodes
    a = it.get_or_create_leaf(1)
    b = it.get_or_create_leaf(2)
    c = it.get_or_create_leaf(3)
    assert [x & 0xFFFFFFFF for x in (a, b, c)] == [1, 2, 3]
    with pytest.raises(O.OraclePanic) as e:
        it.get_or_create_leaf(4)
    assert e.value.kind == "OOM"
    it2 = O.Interner(1, 79 * 4)
    a = it2.get_or_create_leaf(1)
    b = it2.get_or_create_leaf(2)
    it2.get_or_create_leaf(3)
    it2.dec_ref_recursive(b)  # frees index 2
    it2.dec_ref_recursive(a)  # frees index 1
    n1 = it2.get_or_create_leaf(7)
    n2 = it2.get_or_create_leaf(8)
    assert n1 & 0xFFFFFFFF == 1 and n2 & 0xFFFFFFFF == 2  # reverse free order
    # recycled slots carry generation + 1 (mod.rs:585-595)
    assert (n1 >> 32) & 0x7FFF == 1


def test_leaf_and_branch_hash_consing():
    it = O.Interner(1, 1 << 20)
    l1 = it.get_or_create_leaf(5)
    l2 = it.get_or_create_leaf(5)
    assert l1 == l2 and it.get_ref(l1) == 2
    ch = [l1] + [O.EMPTY] * 7
    it.get_or_create_leaf(5)  # the caller-owned child ref for the branch
    b1 = it.get_or_create_branch(ch, 1, 1)
    assert it.get_ref(b1) == 1 and it.get_ref(l1) == 3
    it.get_or_create_leaf(5)
    b2 = it.get_or_create_branch(ch, 1, 1)  # hit: child refs handed back
    assert b1 == b2 and it.get_ref(b1) == 2 and it.get_ref(l1) == 3
    assert it.counts() == (1, 2)
    assert it.get_value(b1) == 0  # LOD mode of [5, 0 x7]: default wins on count
