"""Pins the oracle's sorted pool insert (InsertIntoPool +
parallel_try_insert, src/kgraph.cpp:261-298, 948-967) by brute force: after
any insert sequence the pool holds exactly the L smallest distinct ids,
sorted, and the return codes follow the reference's contract (position,
L for the overflow slot, L + 1 for a duplicate)."""
import numpy as np

import oracle


def test_keeps_best_L_distinct_sorted(rng):
    for trial in range(200):
        L = int(rng.integers(1, 40))
        n = int(rng.integers(1, 200))
        ids = rng.integers(0, 80, n).astype(np.uint32)
        # a given id always has the same distance (symmetric metric)
        table = rng.permutation(80).astype(np.float32) + 0.5
        d = table[ids]
        oi, od, _ = oracle.pool_insert_sequence(ids, d, L)
        uniq = np.unique(ids)
        best = uniq[np.argsort(table[uniq], kind="stable")][:L]
        assert list(oi) == list(best), trial
        assert np.all(np.diff(od) >= 0)


def test_duplicate_returns_L_plus_one():
    ids = np.array([5, 7, 5], np.uint32)
    d = np.array([1.0, 2.0, 1.0], np.float32)
    oi, od, ret = oracle.pool_insert_sequence(ids, d, 4)
    assert list(oi) == [5, 7]
    assert ret[2] == 2 + 1  # L (=2 at that point) + 1


def test_ties_go_after_existing_equal_entries():
    # equal distances keep arrival order (insert after equals), except at the
    # last slot where the reference inserts before (App. A-2)
    ids = np.array([1, 2, 3, 4], np.uint32)
    d = np.array([1.0, 3.0, 1.0, 2.0], np.float32)
    oi, od, _ = oracle.pool_insert_sequence(ids, d, 10)
    assert list(od) == [1.0, 1.0, 2.0, 3.0]
    assert list(oi[:2]) == [1, 3]


def test_full_pool_rejects_worse_and_updates_radius():
    ids = np.arange(10, dtype=np.uint32)
    d = np.arange(10, dtype=np.float32)
    oi, od, ret = oracle.pool_insert_sequence(ids, d, 5)
    assert list(oi) == [0, 1, 2, 3, 4]
    # entries 5..9 are beyond the radius once the pool is full
    assert all(r >= 5 for r in ret[5:])
