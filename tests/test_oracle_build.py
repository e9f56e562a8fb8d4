"""Pins the oracle's NN-descent and full build by what the mathematics fixes:
KNN-graph recall against brute force, graph invariants (out-lists unique and
self-free, degrees bounded, every node reachable from the entry point), the
stop rule, and the sync_prune2 / sync_prune equivalence the paper claims
(reported, PDF p.12)."""
import numpy as np
import pytest

import oracle
from fastkcna_amd import FK_PG_ALPHAPG, FK_PG_HNSW_LAYER, FK_PG_NSG, FK_PG_TAUMNG, IndexParams


def brute_knn(x, K):
    ids, _ = oracle.knn(x, x, K, self_ids=np.arange(len(x)))
    return ids


def small_data(n=3000, d=16, seed=7):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 100, size=(20, d))
    lab = rng.integers(0, 20, n)
    return np.rint(c[lab] + rng.normal(0, 8, size=(n, d))).clip(0, 255).astype(np.float32)


def reachable(off, ids, ep):
    n = len(off) - 1
    seen = np.zeros(n, bool)
    seen[ep] = True
    stack = [ep]
    while stack:
        u = stack.pop()
        for v in ids[off[u]:off[u + 1]]:
            if not seen[v]:
                seen[v] = True
                stack.append(int(v))
    return seen


def test_knn_brute_force_ties_by_id():
    x = np.array([[0.0], [1.0], [1.0], [2.0]], np.float32)
    ids, d = oracle.knn(x, x[:1], 3, self_ids=np.array([0]))
    assert list(ids[0]) == [1, 2, 3] and list(d[0]) == [1.0, 1.0, 4.0]


def test_nndescent_recall():
    x = small_data()
    p = IndexParams(K=20, L=30, S=10, R=50, iterations=8, controls=50)
    pools, plen, st = oracle.nndescent(x, p, threads=4)
    gt = brute_knn(x, 20)
    _, pid = oracle.unpack(pools[:, :20])
    rec = np.mean([len(set(pid[i]) & set(gt[i])) / 20 for i in range(len(x))])
    assert rec > 0.9, rec
    assert st["nnd_iterations"] >= 1
    d, _ = oracle.unpack(pools)
    assert np.all(np.diff(d[:, :20], axis=1) >= 0)


@pytest.mark.parametrize("pg", [FK_PG_NSG, FK_PG_TAUMNG, FK_PG_ALPHAPG, FK_PG_HNSW_LAYER])
def test_build_invariants(pg):
    x = small_data(2000)
    p = IndexParams(K=40, L=40, S=10, R=50, iterations=5, nsg_R=16, search_L=40, search_K=40, tau=1.0, angle=60)
    off, ids, ep, width, st, _ = oracle.build(pg, x, p, threads=4)
    n = len(x)
    deg = np.diff(off)
    for i in range(n):
        nb = ids[off[i]:off[i + 1]]
        assert i not in nb
        assert len(set(nb.tolist())) == len(nb)
    assert width >= p.nsg_R and deg.max() <= width
    if pg != FK_PG_HNSW_LAYER:
        assert reachable(off, ids, ep).all()
    assert st["n_loops"] == p.loop_i + 1
    assert st["d_join"] > 0 and st["d_prune"] > 0 and st["d_bv"] > 0


@pytest.mark.parametrize("pg,tau,angle", [(FK_PG_NSG, 0.0, 60.0), (FK_PG_TAUMNG, 2.0, 60.0),
                                          (FK_PG_ALPHAPG, 0.0, 50.0)])
@pytest.mark.parametrize("loop_i", [2, 3])
def test_prune2_matches_full_rule(pg, tau, angle, loop_i):
    """The paper says the incremental rule (*_sync_prune2, applied from the
    second loop on to nodes the previous loop did not re-prune) is
    result-preserving; the GPU relies on it by always applying the full rule.
    In every loop after the first the oracle runs both rules on the same pool
    state (after a BridgeView pass, so isnew and pid are set) and counts the
    nodes whose out-lists differ: there must be none, over a non-trivial
    number of compared nodes."""
    x = small_data(2000)
    p = IndexParams(K=40, L=40, S=10, R=50, iterations=5, nsg_R=16, search_L=30, search_K=40, loop_i=loop_i,
                    tau=tau, angle=angle)
    *_, (diff, compared) = oracle.build(pg, x, p, threads=1, check_prune2=True)
    print("prune2: differing nodes", diff, "of", compared)
    assert compared >= loop_i * len(x) // 2  # the incremental rule ran in loops 2 .. loop_i + 1
    assert diff == 0
