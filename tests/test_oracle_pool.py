"""The worker-process oracle (tests/oracle_pool.py) against one shared oracle interner:
the union of the workers' live patterns is the shared interner's pattern set, and the
per-tree hashes and reads are the same."""
import numpy as np

import oracle_lib as O
import oracle_pool as OP


def test_pool_union_equals_shared_interner():
    from voxelis_amd import workloads as W
    spec = W.uneven_spec(24, depths=(3, 4, 5), p_dense=0.5, seed=11)
    hashes, reads, status, nl, nb = OP.run_uneven(spec, workers=3, budget=1 << 28, n_reads=32)
    it = O.Interner(2, 1 << 28)
    for i, row in enumerate(spec):
        d = row[0]
        m, v = OP.uneven_tree_batch(row)
        t = O.Tree(d)
        for ph in range(2):
            b = O.Batch(d)
            b.load_dense(m, v if ph == 0 else OP.next_phase(v))
            assert t.apply_batch(it, b) == status[i][ph]
        assert t.canonical_hash(it) == hashes[i]
        vals, pres = t.get_many(it, OP.read_coords(d, 32, 1000 + i))
        assert np.array_equal(np.where(pres == 1, vals, 0), reads[i])
    assert it.counts() == (nl, nb)
    assert nb > 10 and nl > 10
