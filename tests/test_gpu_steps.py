"""Step-level GPU parity through the C ABI: one Link step, UpdateNhoods, one
BridgeView pass and the stop-rule statistics run on the same supplied state
by the GPU (fk_step_*) and by the oracle (whose steps are pinned by
tests/test_oracle_steps.py), compared entry by entry.

Bars: Link, UpdateNhoods and the stop statistics bit-exact on every dataset
(out-list ids, distances and lengths, distance counts).  BridgeView: round
counts exact; every pool's lengths and distances bit-exact and its entries
and flags equal in the canonical (dist, id) order, except which ids of the
pool's last distance stay (insertion-order dependent in the reference); the
evaluated-pair count exact on float data and within 1 % on integer data,
whose many equal distances let the insertion order move entries across the
BridgeView window.
"""
import sys

import numpy as np
import pytest

import oracle
from fastkcna_amd import FK_PG_ALPHAPG, FK_PG_HNSW_LAYER, FK_PG_NSG, FK_PG_TAUMNG, IndexParams
from fastkcna_amd._ffi import StepState
from fastkcna_amd.datasets import float_mixture, integer_mixture, unit_mixture

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


DATA = {
    "int128": lambda: integer_mixture(3000, 128, 24, seed=41),
    "unit96": lambda: unit_mixture(2500, 96, 12, seed=42, sigma=0.25),
    "float960": lambda: float_mixture(700, 960, 6, seed=43),
    "odd20": lambda: np.random.default_rng(44).normal(size=(1500, 20)).astype(np.float32),
}
FLOAT_DATA = ("unit96", "float960", "odd20")
TAU = {"int128": 200.0, "unit96": 0.02, "float960": 0.3, "odd20": 0.3}


def params(name, **kw):
    base = dict(K=40, L=40, S=10, R=40, iterations=4, nsg_R=16, search_L=24, search_K=60, step=10, massq_S=6,
                tau=TAU[name], angle=60)
    base.update(kw)
    return IndexParams(**base)


_STATES = {}


def start_state(device, name):
    """The GPU's own NN-descent pools of the dataset (shared by both sides),
    pool capacity MAXL = search_K = 60 so BridgeView can grow them."""
    if name not in _STATES:
        x = DATA[name]()
        device.upload(x)
        p = params(name)
        pools, plen, _ = device.nndescent(p)
        s = StepState(len(x), 60, 16)
        s.pool[:, :pools.shape[1]] = pools
        s.pool_len[:] = plen
        _STATES[name] = (x, s)
    x, s = _STATES[name]
    device.upload(x)
    return x, s.copy()


def assert_links_equal(g, c):
    assert np.array_equal(g.out_len, c.out_len)
    for i in range(g.n):
        k = c.out_len[i]
        assert np.array_equal(g.out_id[i, :k], c.out_id[i, :k]), i
        assert np.array_equal(g.out_dist[i, :k].view(np.uint32), c.out_dist[i, :k].view(np.uint32)), i


@pytest.mark.parametrize("pg", [FK_PG_NSG, FK_PG_TAUMNG, FK_PG_ALPHAPG, FK_PG_HNSW_LAYER])
@pytest.mark.parametrize("name", list(DATA))
def test_link_step_bitexact(device, pg, name):
    say("link step", pg, name)
    x, s = start_state(device, name)
    p = params(name)
    g, c = s.copy(), s.copy()
    gst = device.step_link(pg, p, g)
    cst = oracle.step_link(pg, x, p, c)
    assert_links_equal(g, c)
    assert gst["d_prune"] == cst["d_prune"] and gst["d_prune_rev"] == cst["d_prune_rev"]
    say("link", name, "kept", int(c.out_len.sum()), "dists", cst["d_prune"], "reversed", cst["d_prune_rev"])


def test_link_step_hub(device):
    """A hub that every other node's pool starts with: its in-degree takes the
    large-union path of the reverse-edge merge, then the reversed prune."""
    rng = np.random.default_rng(45)
    n, d = 1200, 32
    u = rng.normal(size=(n - 1, d))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    x = np.zeros((n, d), np.float32)
    x[1:] = (u * (50.0 + rng.uniform(0, 1, size=(n - 1, 1)))).astype(np.float32)
    device.upload(x)
    ids, dist = oracle.knn(x, x, 30, self_ids=np.arange(n))
    s = StepState(n, 30, 16)
    s.pool[:] = oracle.pack(dist, ids)
    s.pool_len[:] = 30
    p = IndexParams(K=30, L=30, S=10, R=30, nsg_R=16, search_L=30, search_K=30)
    g, c = s.copy(), s.copy()
    gst = device.step_link(FK_PG_NSG, p, g)
    cst = oracle.step_link(FK_PG_NSG, x, p, c)
    assert_links_equal(g, c)
    assert gst["d_prune_rev"] == cst["d_prune_rev"] > 0


def components_state(x, n_comp, nsg_R, rng):
    """Out-lists that split the nodes into n_comp groups (random edges
    inside a group, degrees below nsg_R except a few full nodes in the entry
    point's group) and pools of 16 random nodes in (dist, id) order that
    always hold node 0 (the entry point, in group 0), so every attach finds
    a reached pool entry and no random draw runs."""
    n = len(x)
    group = rng.integers(0, n_comp, n)
    group[0] = 0
    s = StepState(n, 16, nsg_R + 8)
    for i in range(n):
        mates = np.flatnonzero(group == group[i])
        mates = mates[mates != i]
        full = group[i] == 0 and i % 97 == 3
        k = min(len(mates), nsg_R if full else int(rng.integers(1, 4)))
        ids = rng.choice(mates, k, replace=False) if k else np.zeros(0, np.int64)
        s.out_len[i] = k
        s.out_id[i, :k] = ids
        s.out_dist[i, :k] = oracle.pair_distances(x, np.full(k, i), ids)
        others = np.delete(np.arange(n), i)
        cand = rng.choice(others, 15, replace=False)
        if i != 0 and 0 not in cand:
            cand[0] = 0
        d = oracle.pair_distances(x, np.full(len(cand), i), cand)
        order = np.lexsort((cand, d))
        s.pool[i, :len(cand)] = oracle.pack(d[order], cand[order])
        s.pool_len[i] = len(cand)
    return s


@pytest.mark.parametrize("name,n_comp", [("unit96", 40), ("int128", 25), ("odd20", 60)])
def test_connectivity_step_bitexact(device, name, n_comp):
    """The connectivity pass (GPU BFS + min-label propagation + the
    attach loop) against the oracle's DFS / findroot on states of many
    components whose pools always hold a reached node: the same edges in the
    same out-list slots with the same distances, the same counts."""
    x = DATA[name]()[:1500]
    rng = np.random.default_rng(47 + n_comp)
    s = components_state(x, n_comp, 64, rng)
    say("connectivity", name, n_comp)
    device.upload(x)
    p = params(name, nsg_R=64, search_K=16, K=16, L=16, R=16, search_L=16)
    g, c = s.copy(), s.copy()
    gst = device.step_connectivity(p, g, 0)
    cst = oracle.step_connectivity(x, p, c, 0)
    assert gst["n_attached"] == cst["n_attached"] > 0 and gst["d_conn"] == cst["d_conn"]
    assert_links_equal(g, c)
    say("connectivity", name, "attached", cst["n_attached"])


def random_out_lists(s, rng, p_keep=0.5):
    for i in range(s.n):
        k = int(rng.integers(0, s.out_cap + 1))
        s.out_len[i] = k
        s.out_id[i, :k] = rng.choice(s.n, k, replace=False)
        j = int(rng.integers(0, s.out_cap + 1))
        keep = rng.random(k) < p_keep
        prev = np.concatenate([s.out_id[i, :k][keep], rng.choice(s.n, s.out_cap, replace=False)])
        prev = np.unique(prev)[:j]
        s.prev_len[i] = len(prev)
        s.prev_id[i, :len(prev)] = rng.permutation(prev)


@pytest.mark.parametrize("loopcount", [0, 1, 3])
def test_update_nhoods_step_bitexact(device, loopcount):
    rng = np.random.default_rng(46)
    x = DATA["odd20"]()
    device.upload(x)
    s = StepState(len(x), 1, 24, loopcount=loopcount)
    random_out_lists(s, rng)
    g, c = s.copy(), s.copy()
    device.step_update_nhoods(g)
    oracle.step_update_nhoods(c)
    assert np.array_equal(g.prev_len, c.prev_len)
    for i in range(s.n):
        k = s.out_len[i]
        assert np.array_equal(g.out_isnew[i, :k], c.out_isnew[i, :k]), i
        assert np.array_equal(g.prev_id[i, :k], c.prev_id[i, :k]), i
    assert (c.out_isnew[s.out_len > 0].sum() > 0)


def bv_state(device, name, loopcount, rng):
    """NN-descent pools, random pool flags, the oracle's Link out-lists of
    those pools with random isnew bits."""
    x, s = start_state(device, name)
    p = params(name)
    oracle.step_link(FK_PG_NSG, x, p, s)
    s.pool_flags[:] = rng.integers(0, 4, size=s.pool_flags.shape).astype(np.uint8)
    s.out_isnew[:] = rng.integers(0, 2, size=s.out_isnew.shape).astype(np.uint8)
    s.loopcount = loopcount
    return x, s, p


def assert_pools_equal(g, c, exact):
    """The GPU keeps pools in the build's canonical (dist, id) order; the
    reference places an entry after equal distances already present, so its
    order of equal distances is the insertion order.  The oracle's pools are
    put in canonical order (flags moving with their entries) before the
    comparison.  Entries at the pool's last distance may be different ids of
    that distance (which of them stay depends on the insertion order), so
    there only the distances are compared."""
    assert np.array_equal(g.pool_len, c.pool_len)
    for i in range(g.n):
        k = c.pool_len[i]
        o = np.argsort(c.pool[i, :k], kind="stable")
        ck, cf = c.pool[i, :k][o], c.pool_flags[i, :k][o]
        gk, gf = g.pool[i, :k], g.pool_flags[i, :k]
        gd, _ = oracle.unpack(gk)
        cd, _ = oracle.unpack(ck)
        assert np.array_equal(gd.view(np.uint32), cd.view(np.uint32)), i
        below = cd < cd[-1] if k else cd.astype(bool)
        assert np.array_equal(gk[below], ck[below]), i
        if exact:
            assert np.array_equal(gf[below], cf[below]), i


@pytest.mark.parametrize("loopcount", [1, 2])
@pytest.mark.parametrize("name", list(DATA))
def test_bridgeview_step(device, name, loopcount):
    say("bridgeview step", name, loopcount)
    rng = np.random.default_rng(47 + loopcount)
    x, s, p = bv_state(device, name, loopcount, rng)
    g, c = s.copy(), s.copy()
    gst = device.step_bridgeview(p, g)
    cst = oracle.step_bridgeview(x, p, c)
    assert gst["bv_rounds"][0] == cst["bv_rounds"][0]
    exact = name in FLOAT_DATA
    assert_pools_equal(g, c, exact)
    if exact:
        assert gst["d_bv"] == cst["d_bv"]
    else:
        assert abs(gst["d_bv"] - cst["d_bv"]) <= 0.01 * cst["d_bv"]
    assert 0 < gst["d_bv_computed"] <= gst["d_bv"]
    grew = int((c.pool_len > s.pool_len).sum())
    say("bv", name, "rounds", cst["bv_rounds"][0], "pairs", cst["d_bv"], "computed", gst["d_bv_computed"],
        "pools grown", grew)
    assert grew > 0 and (c.pool != s.pool).any()


def canonical(s):
    """The state with every pool in (dist, id) order (flags moved along): the
    input form the step entry points document."""
    for i in range(s.n):
        k = s.pool_len[i]
        o = np.argsort(s.pool[i, :k], kind="stable")
        s.pool[i, :k] = s.pool[i, :k][o]
        s.pool_flags[i, :k] = s.pool_flags[i, :k][o]
    return s


def test_loop_trajectory_matches(device):
    """Two loops of Link -> UpdateNhoods -> BridgeView from the NN-descent
    state; each step starts both sides from the oracle's previous result
    (pools in canonical order) and their outputs must agree at every step."""
    rng = np.random.default_rng(49)
    x, s = start_state(device, "unit96")
    p = params("unit96")
    for loop in range(2):
        say("trajectory loop", loop)
        g, c = s.copy(), s.copy()
        device.step_link(FK_PG_ALPHAPG, p, g)
        oracle.step_link(FK_PG_ALPHAPG, x, p, c)
        assert_links_equal(g, c)
        s = c
        s.loopcount = loop
        g, c = s.copy(), s.copy()
        device.step_update_nhoods(g)
        oracle.step_update_nhoods(c)
        assert np.array_equal(g.prev_len, c.prev_len)
        for i in range(s.n):
            k = s.out_len[i]
            assert np.array_equal(g.out_isnew[i, :k], c.out_isnew[i, :k]) and np.array_equal(g.prev_id[i, :k],
                                                                                             c.prev_id[i, :k]), i
        s = c
        s.loopcount = loop + 1
        if loop == 0:
            s.pool_flags[:] = rng.integers(0, 2, size=s.pool_flags.shape).astype(np.uint8)
        g, c = s.copy(), s.copy()
        gst = device.step_bridgeview(p, g)
        cst = oracle.step_bridgeview(x, p, c)
        assert gst["bv_rounds"][0] == cst["bv_rounds"][0] and gst["d_bv"] == cst["d_bv"]
        assert_pools_equal(g, c, True)
        s = canonical(c)


@pytest.mark.parametrize("name", ["int128", "unit96"])
def test_stop_stats_step_bitexact(device, name):
    rng = np.random.default_rng(50)
    x, s = start_state(device, name)
    s.pool_flags[:] = (rng.random(s.pool_flags.shape) < 0.1).astype(np.uint8)
    p = params(name)
    ctrl = rng.choice(s.n, 50, replace=False).astype(np.uint32)
    gt, _ = oracle.knn(x, x[ctrl], p.K, self_ids=ctrl)
    gd, gr = device.step_stop_stats(p, s, ctrl, gt)
    cd, cr = oracle.step_stop_stats(p, s, ctrl, gt)
    say("stop stats", name, gd, cd, gr, cr)
    assert np.float32(gd) == np.float32(cd) and np.float32(gr) == np.float32(cr)
    assert 0 < cd < 1 and 0 < cr <= 1
    # the trigger the build takes from them
    for thr in (cd, np.nextafter(np.float32(cd), np.float32(0))):
        q = params(name, delta=float(thr))
        assert (device.step_stop_stats(q, s, ctrl, gt)[0] <= q.delta) == (cd <= q.delta)
