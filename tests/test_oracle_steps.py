"""Pins the oracle's stateful steps (src/kgraph.cpp: BridgeView :2478-2663,
InterInsert / Link :2051-2143, UpdateNhoods :3537-3584, the stop rule
:1365-1428, ComputeEp :1528-1623) on small hand-built states whose results
follow from the reference's definitions by hand:

* BridgeView with empty out-lists computes nothing, so its rounds depend only
  on how many window entries are flagged: every entry is flagged by the
  initialisation, a round expands at most massq_S per node, and the round
  whose capped count is below half of N * massq_S expands the rest and is
  the last.  This fixes the window W = min(search_L + (loop - 1) step, MAXL),
  the budget and the strict "< 0.5" of the 50 % rule.
* BridgeView on two far clusters whose out-lists point across: no candidate
  can enter a pool, and the evaluated-pair count is the sum over expanded
  window entries of the out-list entries that pass the one-sided skip rule
  ("both old" pairs skipped; isnew of a window entry = its old flag inside
  [pre_L, cur_L) from the second BridgeView on).
* Link on a 1-D chain where each pool holds only nodes to the right: the NSG
  rule keeps the nearest right neighbour, InterInsert adds the left one, and
  with nsg_R = 1 the reversed prune keeps the nearer of the two.
* The connectivity pass (tree_grow / findroot / DFS, :2145-2260) on an
  8-node graph of four components: the attached edges, in order, follow by
  hand from "first unreached id, first reached pool entry with fewer than
  nsg_R out-edges" (one reached through an earlier attach, one full node
  skipped), each with the reference distance of its two rows.
"""
import numpy as np
import pytest

import oracle
from fastkcna_amd import FK_PG_ALPHAPG, FK_PG_HNSW_LAYER, FK_PG_NSG, FK_PG_TAUMNG, IndexParams
from fastkcna_amd._ffi import StepState


def exact_pools(x, cap, rng=None, lens=None):
    """Pools of the exact nearest neighbours (self excluded), ties by id."""
    ids, d = oracle.knn(x, x, cap, self_ids=np.arange(len(x)))
    s = StepState(len(x), cap, 8)
    s.pool[:] = oracle.pack(d, ids)
    s.pool_len[:] = cap if lens is None else lens
    return s


def bv_rounds_closed_form(F, S):
    """Rounds of a BridgeView pass in which nothing is inserted: F[i]
    flagged window entries per node, at most S expanded per node and round;
    the round whose capped count falls below N S / 2 expands the rest."""
    rem = np.array(F, np.int64)
    n = len(rem)
    rounds = 0
    while True:
        capped = int(np.minimum(rem, S).sum())
        rem -= np.minimum(rem, S)
        rounds += 1
        if np.float32(capped) / np.float32(n) / np.float32(S) < np.float32(0.5):
            return rounds


@pytest.mark.parametrize("loopcount", [1, 2, 3])
@pytest.mark.parametrize("lens_kind", ["full", "ragged"])
def test_bridgeview_rounds_window_and_budget(loopcount, lens_kind):
    rng = np.random.default_rng(31)
    n, cap = 300, 60
    x = rng.normal(size=(n, 8)).astype(np.float32)
    lens = np.full(n, cap) if lens_kind == "full" else rng.integers(1, cap + 1, n)
    s = exact_pools(x, cap, lens=lens)
    s.pool_flags[:] = rng.integers(0, 4, size=s.pool_flags.shape).astype(np.uint8)
    old_flags = s.pool_flags.copy()
    s.loopcount = loopcount
    p = IndexParams(K=cap, L=cap, S=10, search_L=17, search_K=cap, step=9, massq_S=7)
    before = s.pool.copy()
    st = oracle.step_bridgeview(x, p, s)
    W = min(17 + (loopcount - 1) * 9, cap)
    F = np.minimum(W, lens)
    assert st["bv_rounds"][0] == bv_rounds_closed_form(F, 7)
    assert st["d_bv"] == 0
    assert np.array_equal(s.pool_len, lens)
    assert np.array_equal(s.pool, before)
    pre = 0 if loopcount == 1 else min(17 + (loopcount - 2) * 9, cap)
    for i in range(n):
        j = np.arange(lens[i])
        flag = (j >= W).astype(np.uint8)  # the window was expanded, the rest stays flagged
        isnew = ((j >= pre) & (j < W) & (loopcount > 1) & ((old_flags[i, j] & 1) == 1)).astype(np.uint8)
        assert np.array_equal(s.pool_flags[i, :lens[i]], flag | (isnew << 1)), i


def test_bridgeview_rounds_rule_is_strict():
    """F = 25 flagged per node, massq_S = 10: rounds expand 10, 10, 5 (5 / 10
    = 0.5 is not below 0.5, so the third round is not the last) and a fourth
    round finds nothing; with F = 24 the third round (4 / 10) is the last."""
    rng = np.random.default_rng(32)
    n = 100
    x = rng.normal(size=(n, 8)).astype(np.float32)
    for F, want in ((25, 4), (24, 3), (5, 2), (4, 1)):
        s = exact_pools(x, 40, lens=np.full(n, 40))
        s.loopcount = 1
        p = IndexParams(K=40, L=40, S=10, search_L=F, search_K=40, massq_S=10)
        st = oracle.step_bridgeview(x, p, s)
        assert st["bv_rounds"][0] == want, (F, st["bv_rounds"][0])


@pytest.mark.parametrize("loopcount", [1, 2])
def test_bridgeview_skip_rule_counts(loopcount):
    rng = np.random.default_rng(33)
    half, cap = 120, 30
    a = rng.normal(size=(half, 8))
    b = rng.normal(size=(half, 8)) + 1000.0  # far: nothing from the other cluster enters a pool
    x = np.concatenate([a, b]).astype(np.float32)
    n = 2 * half
    s = StepState(n, cap, 12)
    for lo in (0, half):
        part = x[lo:lo + half]
        ids, d = oracle.knn(part, part, cap, self_ids=np.arange(half))
        s.pool[lo:lo + half] = oracle.pack(d, ids + lo)
    s.pool_len[:] = cap
    s.pool_flags[:] = rng.integers(0, 2, size=s.pool_flags.shape).astype(np.uint8)
    old_flags = s.pool_flags.copy()
    for i in range(n):
        other = np.arange(half) + (half if i < half else 0)
        k = int(rng.integers(0, 13))
        s.out_len[i] = k
        s.out_id[i, :k] = rng.choice(other, k, replace=False)
        s.out_isnew[i, :k] = rng.integers(0, 2, k)
    s.loopcount = loopcount
    p = IndexParams(K=cap, L=cap, S=10, search_L=12, search_K=cap, step=8, massq_S=4)
    before = s.pool.copy()
    st = oracle.step_bridgeview(x, p, s)
    assert np.array_equal(s.pool, before)
    W = min(12 + (loopcount - 1) * 8, cap)
    pre = 0 if loopcount == 1 else 12
    want = 0
    for i in range(n):
        for j in range(W):
            q = int(before[i, j] & np.uint64(0xFFFFFFFF))
            isnew = loopcount > 1 and pre <= j < W and (old_flags[i, j] & 1)
            e_new = s.out_isnew[q, :s.out_len[q]]
            want += int(s.out_len[q]) if isnew else int(np.count_nonzero(e_new))
    assert st["d_bv"] == want


def test_bridgeview_pools_improve_and_stay_exact():
    """From random pools, a pass can only improve each pool position, and
    every pool stays strictly (dist, id) ordered with exact distances."""
    rng = np.random.default_rng(34)
    n, cap = 400, 40
    x = rng.normal(size=(n, 16)).astype(np.float32)
    s = StepState(n, cap, 10)
    for i in range(n):
        cand = rng.choice(np.delete(np.arange(n), i), cap, replace=False)
        d = oracle.pair_distances(x, np.full(cap, i), cand)
        o = np.lexsort((cand, d))
        s.pool[i] = oracle.pack(d[o], cand[o])
        s.out_len[i] = 10
        s.out_id[i] = rng.choice(np.delete(np.arange(n), i), 10, replace=False)
        s.out_isnew[i] = 1
    s.pool_len[:] = cap
    s.loopcount = 1
    before_d, _ = oracle.unpack(s.pool.copy())
    p = IndexParams(K=cap, L=cap, S=10, search_L=20, search_K=cap, massq_S=10)
    st = oracle.step_bridgeview(x, p, s)
    d, ids = oracle.unpack(s.pool)
    assert st["d_bv"] > 0
    assert np.all(d <= before_d) and np.any(d < before_d)
    assert np.all(s.pool[:, 1:] > s.pool[:, :-1])
    rows = np.repeat(np.arange(n), cap)
    assert np.array_equal(oracle.pair_distances(x, rows, ids.ravel()).view(np.uint32), d.ravel().view(np.uint32))
    assert not np.any(ids == np.arange(n)[:, None])


def chain(n, rng):
    gaps = 1.0 + rng.permutation(n - 1) / n  # distinct gaps: no distance ties
    x = np.zeros((n, 8), np.float32)
    x[1:, 0] = np.cumsum(gaps)
    return x


@pytest.mark.parametrize("pg,tau,angle", [(FK_PG_NSG, 0.0, 60.0), (FK_PG_TAUMNG, 0.0, 60.0),
                                          (FK_PG_ALPHAPG, 0.0, 60.0), (FK_PG_HNSW_LAYER, 0.0, 60.0)])
@pytest.mark.parametrize("R", [1, 2, 3])
def test_link_on_a_chain(pg, tau, angle, R):
    rng = np.random.default_rng(35)
    n, m = 50, 6
    x = chain(n, rng)
    s = StepState(n, m, 4)
    for i in range(n):
        right = np.arange(i + 1, min(n, i + 1 + m))
        nb = right if len(right) else np.array([i - 1])
        d = oracle.pair_distances(x, np.full(len(nb), i), nb)
        s.pool[i, :len(nb)] = oracle.pack(d, nb)
        s.pool_len[i] = len(nb)
    p = IndexParams(K=m, L=m, S=2, nsg_R=R, search_L=m, search_K=m, tau=tau, angle=angle)
    st = oracle.step_link(pg, x, p, s)
    for i in range(n):
        got = list(s.out_id[i, :s.out_len[i]])
        if i == n - 1:
            want = [n - 2]
        elif i == 0:
            want = [1]
        else:
            left, right = i - 1, i + 1
            if R >= 2:
                # kept prefix, then the new reverse edge
                want = [right, left]
            else:
                dl, dr = x[i, 0] - x[left, 0], x[right, 0] - x[i, 0]
                want = [left if dl < dr else right]
        assert got == want, (i, got, want)
    # with room for a second edge, each node with r right neighbours checks
    # its r - 1 other candidates against the one kept edge; with nsg_R = 1
    # the prune stops at the first kept edge
    r = np.array([min(m, n - 1 - i) for i in range(n)])
    assert st["d_prune"] == (int(np.sum(np.maximum(r - 1, 0))) if R >= 2 else 0)
    # R = 1: the reversed prune keeps the nearer entry and stops (nothing checked)
    assert st["d_prune_rev"] == 0


def test_update_nhoods_by_hand():
    s = StepState(3, 1, 4, loopcount=1)
    s.out_len[:] = [3, 2, 0]
    s.out_id[0, :3] = [3, 5, 7]
    s.out_id[1, :2] = [1, 2]
    s.prev_len[:] = [2, 0, 1]
    s.prev_id[0, :2] = [5, 9]
    s.prev_id[2, :1] = [4]
    oracle.step_update_nhoods(s)
    assert list(s.out_isnew[0, :3]) == [1, 0, 1]
    assert list(s.out_isnew[1, :2]) == [1, 1]
    assert list(s.prev_len) == [3, 2, 0]
    assert list(s.prev_id[0, :3]) == [3, 5, 7] and list(s.prev_id[1, :2]) == [1, 2]
    s0 = StepState(1, 1, 4, loopcount=0)
    s0.out_len[:] = [2]
    s0.out_id[0, :2] = [8, 9]
    s0.prev_len[:] = [2]
    s0.prev_id[0, :2] = [8, 9]
    oracle.step_update_nhoods(s0)
    assert list(s0.out_isnew[0, :2]) == [1, 1]  # the first loop marks everything new


def test_stop_stats_by_hand():
    s = StepState(8, 6, 1)  # 8 nodes: more than K, L, search_L and search_K (nothing shrinks)
    s.pool_len[:] = [6, 6, 3, 6, 6, 6, 6, 6]
    s.pool_flags[0] = [1, 1, 0, 0, 1, 1]  # first K = 4: two flagged
    s.pool_flags[1] = [0, 0, 0, 0, 1, 1]  # none
    s.pool_flags[2] = [1, 1, 1, 1, 1, 1]  # only 3 valid entries
    s.pool_flags[3] = [3, 1, 2, 1, 0, 0]  # bit 0 only: three flagged
    for i in range(8):
        s.pool[i] = oracle.pack(np.arange(6, dtype=np.float32), np.arange(10 * i, 10 * i + 6) % 8)
    s.pool[3] = oracle.pack(np.arange(6, dtype=np.float32), [0, 1, 2, 4, 5, 6])
    p = IndexParams(K=4, L=6, S=2, search_L=6, search_K=6)
    gt = np.array([[0, 1, 2, 3], [2, 1, 7, 3]], np.uint32)
    delta, recall = oracle.step_stop_stats(p, s, np.array([0, 3]), gt)
    assert delta == np.float32((2 / 4 + 0 + 3 / 4 + 3 / 4) / 8)
    assert recall == np.float32((4 / 4 + 2 / 4) / 2)


def test_entry_point_is_a_local_minimum():
    """ComputeEp's beam search expands its best result, so no out-neighbour
    of the entry point is closer to the centroid.  HNSW-layer builds add no
    connectivity edges after the search, so the final graph is the searched
    one.  Integer rows with small sums make the float centroid exact."""
    rng = np.random.default_rng(36)
    n, d = 3000, 16
    c = rng.uniform(0, 60, size=(12, d))
    x = np.rint(c[rng.integers(0, 12, n)] + rng.normal(0, 6, size=(n, d))).clip(0, 255).astype(np.float32)
    assert x.sum(axis=0).max() < 2 ** 24
    center = (x.astype(np.float64).sum(axis=0).astype(np.float32) / np.float32(n)).astype(np.float32)
    p = IndexParams(K=30, L=30, S=10, R=30, iterations=5, nsg_R=12, search_L=30, search_K=30)
    off, ids, ep, *_ = oracle.build(FK_PG_HNSW_LAYER, x, p, threads=4)
    nb = ids[off[ep]:off[ep + 1]]
    assert len(nb) > 0
    dep = oracle.l2(center, x[ep])
    for v in nb:
        assert dep <= oracle.l2(center, x[v]), v


def test_search_is_exhaustive_when_the_beam_covers_the_graph():
    """With a beam at least as wide as a connected graph, best-first search
    evaluates every node once and returns the exact nearest neighbours (ties
    by id); on a complete graph, expanding the entry point alone does."""
    rng = np.random.default_rng(37)
    n = 300
    x = rng.normal(size=(n, 12)).astype(np.float32)
    qs = rng.normal(size=(40, 12)).astype(np.float32)
    gt, _ = oracle.knn(x, qs, 10)
    ring = [np.array([(i + 1) % n, (i - 1) % n]) for i in range(n)]
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum([len(v) for v in ring])
    res, nd = oracle.search(x, off, np.concatenate(ring), 17, qs, n, 10)
    assert np.array_equal(res, gt) and nd == 40 * n
    full = [np.delete(np.arange(n), i) for i in range(n)]
    off[1:] = np.cumsum([len(v) for v in full])
    res, nd = oracle.search(x, off, np.concatenate(full), 5, qs, 20, 10)
    assert np.array_equal(res, gt) and nd == 40 * n


def test_search_on_a_chain_is_greedy():
    """1-D points joined to their two neighbours: from either end a beam of 2
    walks monotonically to the query's nearest point."""
    n = 200
    x = np.zeros((n, 8), np.float32)
    x[:, 0] = np.arange(n) * 1.5
    ring = [np.array([j for j in (i - 1, i + 1) if 0 <= j < n]) for i in range(n)]
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum([len(v) for v in ring])
    ids = np.concatenate(ring)
    qs = np.zeros((5, 8), np.float32)
    qs[:, 0] = [0.2, 77.9, 150.1, 298.0, 120.7]
    for ep in (0, n - 1):
        res, _ = oracle.search(x, off, ids, ep, qs, 2, 1)
        assert list(res[:, 0]) == [int(np.argmin(np.abs(x[:, 0] - q))) for q in qs[:, 0]]


def test_connectivity_attach_order_by_hand():
    """ep = 0 reaches {0, 1}; components {2, 3}, {4, 5}, {6, 7} are not
    reached; nsg_R = 2.  Out-lists: 0 -> 1, 1 -> 0, 2 -> 3, 4 -> 5, 5 -> 4,
    6 -> 7.  Pools (only their ids and order matter): 2: [5, 1, 0],
    4: [6, 3, 0], 6: [1, 0, 2].  By hand: id 2 is the first unreached id, 5
    is unreached, 1 is reached with 1 < 2 out-edges: edge 1 -> 2 (3 is then
    reached through 2).  Id 4: 6 is unreached, 3 is reached (through the
    first attach) with no out-edge: edge 3 -> 4.  Id 6: 1 now has 2 = nsg_R
    out-edges and is skipped, 0 has one: edge 0 -> 6.  Three edges, no
    random draw."""
    rng = np.random.default_rng(61)
    x = rng.integers(0, 20, size=(8, 8)).astype(np.float32)
    s = StepState(8, 3, 4)
    out = {0: [1], 1: [0], 2: [3], 4: [5], 5: [4], 6: [7]}
    for i, lst in out.items():
        s.out_len[i] = len(lst)
        s.out_id[i, :len(lst)] = lst
        s.out_dist[i, :len(lst)] = [oracle.l2(x[i], x[j]) for j in lst]
    pools = {i: [(i + 1) % 8, (i + 2) % 8, (i + 3) % 8] for i in range(8)}
    pools.update({2: [5, 1, 0], 4: [6, 3, 0], 6: [1, 0, 2]})
    for i, ids in pools.items():
        s.pool[i, :3] = oracle.pack(np.float32([1.0, 2.0, 3.0]), ids)
        s.pool_len[i] = 3
    p = IndexParams(K=3, L=3, S=2, R=3, nsg_R=2, search_L=3, search_K=3)
    st = oracle.step_connectivity(x, p, s, 0)
    assert st["n_attached"] == 3 and st["d_conn"] == 3
    want = {0: [1, 6], 1: [0, 2], 2: [3], 3: [4], 4: [5], 5: [4], 6: [7], 7: []}
    for i, lst in want.items():
        assert s.out_len[i] == len(lst), i
        assert list(s.out_id[i, :len(lst)]) == lst, i
    for src, dst in ((1, 2), (3, 4), (0, 6)):
        k = list(s.out_id[src, :s.out_len[src]]).index(dst)
        assert s.out_dist[src, k] == np.float32(oracle.l2(x[src], x[dst]))
