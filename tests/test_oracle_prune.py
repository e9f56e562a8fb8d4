"""Pins the oracle's occlusion pruners (src/kgraph.cpp:1809-1851,
2833-2876, 3157-3203 and the *_reversed_prune forms).

The keep/drop decisions are re-derived from the rule's definition rather than
re-simulated: the output must (a) start with the first non-self candidate,
(b) contain no entry occluded by (or equal in id to) an earlier kept entry,
(c) drop every examined candidate that some kept entry ranked before it
occludes, (d) stop exactly when nsg_R entries are kept or the examined range
(min(len, search_K); all for the reversed form) is exhausted.  Those four
properties determine the greedy output uniquely.  Closed forms: tau = 0 is
the NSG rule; angle = 60 is the NSG rule away from equilateral triangles;
orthogonal candidates are never occluded; duplicate ids are dropped.
"""
import numpy as np
import pytest

import oracle
from fastkcna_amd import FK_PG_ALPHAPG, FK_PG_NSG, FK_PG_TAUMNG, IndexParams


def make_pools(x, n_tasks, L, rng, dup=False):
    n = x.shape[0]
    q = rng.choice(n, n_tasks, replace=False).astype(np.uint32)
    pools = np.zeros((n_tasks, L), np.uint64)
    plen = np.zeros(n_tasks, np.uint32)
    for t, qq in enumerate(q):
        cand = rng.choice(n, L, replace=False)
        if dup and L > 4:
            cand[3] = cand[1]
        d = oracle.pair_distances(x, np.full(L, qq), cand)
        order = np.lexsort((cand, d))
        pools[t] = oracle.pack(d[order], cand[order])
        plen[t] = L
    return pools, plen, q


def occludes(pg, djk, pd, dqt, p):
    djk, pd, dqt = np.float32(djk), np.float32(pd), np.float32(dqt)
    if pg == FK_PG_NSG:
        return djk < pd
    if pg == FK_PG_TAUMNG:
        return djk < pd - np.float32(3) * np.float32(p.tau)
    with np.errstate(divide="ignore", invalid="ignore"):
        c = ((dqt + djk) - pd) / np.float32(2) / np.sqrt(np.float32(dqt * djk))
    thr = np.float32(np.cos(np.float64(np.float32(p.angle) / np.float32(180)) * np.pi))
    return bool(djk < pd and c < thr)


def check_decisions(pg, x, keys, q, out, p, maxc):
    d, ids = oracle.unpack(keys)
    start = 1 if ids[0] == q else 0
    assert out[0] == ids[start]
    kept = [start]
    pos = {int(i): k for k, i in enumerate(ids)}
    limit = min(len(ids), maxc)
    examined_to = start
    for k in range(start + 1, limit):
        if len(kept) >= p.nsg_R:
            break
        examined_to = k
        occl = False
        for t in kept:
            if ids[t] == ids[k]:
                occl = True
                break
            djk = oracle.pair_distances(x, [ids[t]], [ids[k]])[0]
            if occludes(pg, djk, d[k], d[t], p):
                occl = True
                break
        if not occl:
            kept.append(k)
    assert [int(ids[k]) for k in kept] == [int(v) for v in out], "greedy re-derivation"
    for j, k in enumerate(kept[1:], 1):
        for t in kept[:j]:
            djk = oracle.pair_distances(x, [ids[t]], [ids[k]])[0]
            assert not occludes(pg, djk, d[k], d[t], p)
    del pos, examined_to


@pytest.mark.parametrize("pg", [FK_PG_NSG, FK_PG_TAUMNG, FK_PG_ALPHAPG])
@pytest.mark.parametrize("reversed_", [False, True])
def test_decisions_follow_the_rule(pg, reversed_, rng):
    x = rng.normal(size=(400, 24)).astype(np.float32)
    p = IndexParams(nsg_R=12, search_K=40, tau=0.3, angle=70)
    pools, plen, q = make_pools(x, 30, 60, rng, dup=True)
    oi, od, ol, cc = oracle.prune(pg, x, pools, plen, q, 400, p, reversed_=reversed_)
    maxc = 1 << 30 if reversed_ else p.search_K
    for t in range(30):
        check_decisions(pg, x, pools[t], q[t], oi[t, :ol[t]], p, maxc)
    assert cc > 0


def test_tau_zero_is_nsg_rule(rng):
    x = rng.normal(size=(300, 16)).astype(np.float32)
    pools, plen, q = make_pools(x, 40, 50, rng)
    p = IndexParams(nsg_R=10, search_K=50, tau=0.0)
    a = oracle.prune(FK_PG_NSG, x, pools, plen, q, 300, p)
    b = oracle.prune(FK_PG_TAUMNG, x, pools, plen, q, 300, p)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2]) and a[3] == b[3]


def test_angle_60_matches_nsg_on_generic_data(rng):
    x = rng.normal(size=(300, 16)).astype(np.float32)
    pools, plen, q = make_pools(x, 40, 50, rng)
    p = IndexParams(nsg_R=10, search_K=50, angle=60)
    a = oracle.prune(FK_PG_NSG, x, pools, plen, q, 300, p)
    b = oracle.prune(FK_PG_ALPHAPG, x, pools, plen, q, 300, p)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2])


def test_alpha_threshold_constant():
    # (float)cos((double)(60.f / 180) * pi): the float quotient is 0.3333333432...
    thr = np.float32(np.cos(np.float64(np.float32(60.0) / np.float32(180.0)) * np.pi))
    assert thr == np.float32(0.49999997)


def test_orthogonal_candidates_are_kept():
    d = 20
    x = np.zeros((d + 1, d), np.float32)
    for i in range(d):
        x[i + 1, i] = 1.0 + i * 0.1
    q = 0
    cand = np.arange(1, d + 1)
    dist = oracle.pair_distances(x, np.zeros(d, int), cand)
    order = np.lexsort((cand, dist))
    pools = oracle.pack(dist[order], cand[order])[None, :]
    p = IndexParams(nsg_R=8, search_K=100)
    oi, od, ol, cc = oracle.prune(FK_PG_NSG, x, pools, np.array([d], np.uint32), np.array([q]), d + 1, p)
    assert ol[0] == 8 and list(oi[0]) == list(cand[order][:8])


def test_self_entry_and_duplicates_dropped(rng):
    x = rng.normal(size=(50, 8)).astype(np.float32)
    q = 7
    cand = np.array([7, 3, 3, 9, 11], np.uint32)
    dist = oracle.pair_distances(x, np.full(5, q), cand)
    dist[0] = 0.0
    order = np.lexsort((cand, dist))
    pools = oracle.pack(dist[order], cand[order])[None, :]
    p = IndexParams(nsg_R=10, search_K=100)
    oi, od, ol, _ = oracle.prune(FK_PG_NSG, x, pools, np.array([5], np.uint32), np.array([q]), 50, p)
    got = list(oi[0, :ol[0]])
    assert q not in got
    assert len(got) == len(set(got))


def test_ids_beyond_oracle_size_are_skipped(rng):
    x = rng.normal(size=(60, 8)).astype(np.float32)
    q = 0
    cand = np.arange(1, 60, dtype=np.uint32)
    dist = oracle.pair_distances(x, np.zeros(59, int), cand)
    order = np.lexsort((cand, dist))
    pools = oracle.pack(dist[order], cand[order])[None, :]
    p = IndexParams(nsg_R=59, search_K=100)
    oi, od, ol, _ = oracle.prune(FK_PG_NSG, x, pools, np.array([59], np.uint32), np.array([q]), 30, p)
    got = oi[0, :ol[0]]
    assert np.all(got[1:] < 30)  # pool[start] itself is always kept
