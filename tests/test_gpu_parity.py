"""GPU parity through the C ABI against the CPU oracle, at sizes the oracle
finishes in seconds.

Bars: distances bit-exact against the reference order (and within 1e-6
relative of FP64 on float data); prune ids, kept distances and distance
counts bit-exact; exact kNN ids bit-exact; NN-descent / full builds by
recall against brute force (no worse than the oracle, with a stated margin
for the sampling RNG) and graph invariants.
"""
import sys

import numpy as np
import pytest

import oracle
from fastkcna_amd import _ffi
from fastkcna_amd import FK_PG_ALPHAPG, FK_PG_HNSW_LAYER, FK_PG_NSG, FK_PG_TAUMNG, IndexParams
from fastkcna_amd.datasets import float_mixture, integer_mixture, unit_mixture

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def datasets():
    return {
        "int128": integer_mixture(4000, 128, 32, seed=11),
        "float960": float_mixture(600, 960, 8, seed=12),
        "unit96": unit_mixture(3000, 96, 16, seed=13, sigma=0.25),
        "odd100": np.random.default_rng(14).normal(size=(1500, 100)).astype(np.float32),
        # dpad % 32 of 24 and 16: the canonical order's tail blocks come first
        "odd20": np.random.default_rng(15).normal(size=(1500, 20)).astype(np.float32),
        "odd44": integer_mixture(2500, 44, 16, seed=16),
        # width 256: the prune's widest register-chain path (and variant 2's)
        "unit256": unit_mixture(2000, 256, 16, seed=17, sigma=0.3),
    }


@pytest.mark.parametrize("name", ["int128", "float960", "unit96", "odd100", "odd20", "odd44"])
def test_pair_distances_bitexact(device, name):
    x = datasets()[name]
    say("pair", name, x.shape)
    device.upload(x)
    rng = np.random.default_rng(1)
    a = rng.integers(0, len(x), 20000)
    b = rng.integers(0, len(x), 20000)
    g = device.pair_distances(a, b)
    c, c64 = oracle.pair_distances(x, a, b, with_f64=True)
    assert np.array_equal(g.view(np.uint32), c.view(np.uint32))
    nz = c64 > 0
    assert np.max(np.abs(g[nz] - c64[nz]) / c64[nz]) < 1e-6
    if name == "int128":
        assert np.array_equal(g.astype(np.float64), c64)


@pytest.mark.parametrize("name", ["int128", "unit96", "float960"])
def test_join_distances_bitexact(device, name):
    x = datasets()[name]
    say("join", name)
    device.upload(x)
    rng = np.random.default_rng(2)
    n_nodes = 300
    sizes = rng.integers(0, 140, n_nodes)
    sizes[:3] = [0, 1, 200]
    nn = np.minimum(rng.integers(0, 60, n_nodes), sizes)
    off = np.zeros(n_nodes + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    ids = rng.integers(0, len(x), int(off[-1])).astype(np.uint32)
    g = device.join_distances(ids, off, nn)
    a, b = [], []
    for t in range(n_nodes):
        lst = ids[off[t]:off[t + 1]]
        T = len(lst)
        for i in range(nn[t]):
            for j in range(i + 1, T):
                a.append(lst[i])
                b.append(lst[j])
    c = oracle.pair_distances(x, np.array(a), np.array(b))
    assert np.array_equal(g.view(np.uint32), c.view(np.uint32))


def make_pools(x, n_tasks, L, rng, q=None):
    n = len(x)
    q = rng.choice(n, n_tasks, replace=False).astype(np.uint32) if q is None else q
    pools = np.zeros((n_tasks, L), np.uint64)
    plen = np.zeros(n_tasks, np.uint32)
    for t, qq in enumerate(q):
        cand = rng.choice(n, L, replace=False)
        d = oracle.pair_distances(x, np.full(L, qq), cand)
        order = np.lexsort((cand, d))
        pools[t] = oracle.pack(d[order], cand[order])
        plen[t] = rng.integers(1, L + 1)
    return pools, plen, q


# prune kernel variants (the build picks one per shape by timing; variant 2
# applies to FP32 rows of width 96, 128 and 256, variant 3 -- Gram-form
# occlusion distances on MFMA -- to Gram-prepared datasets of width <= 256,
# and each runs as 0 elsewhere)
VARIANT_FLAGS = {0: _ffi.FK_FLAG_PRUNE_V0, 1: _ffi.FK_FLAG_PRUNE_V1, 2: _ffi.FK_FLAG_PRUNE_V2,
                 3: _ffi.FK_FLAG_PRUNE_V3 | _ffi.FK_FLAG_GRAM}


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("pg", [FK_PG_NSG, FK_PG_TAUMNG, FK_PG_ALPHAPG])
@pytest.mark.parametrize("reversed_", [False, True])
@pytest.mark.parametrize("name", ["int128", "unit96", "float960", "odd100", "odd20", "odd44", "unit256"])
def test_prune_bitexact(device, pg, reversed_, name, variant):
    """Both prune kernel variants (the build picks one per shape by timing)
    against the oracle."""
    x = datasets()[name]
    say("prune", pg, reversed_, name, "variant", variant)
    device.upload(x)
    rng = np.random.default_rng(3)
    tau = {"int128": 300.0, "unit96": 0.05, "float960": 0.5, "odd100": 1.0, "odd20": 0.2, "odd44": 100.0,
           "unit256": 0.05}[name]
    p = IndexParams(nsg_R=24, search_K=90, tau=tau, angle=63)
    p.flags = VARIANT_FLAGS[variant]
    pools, plen, q = make_pools(x, 400, 120, rng)
    # a self entry at the front of some pools
    for t in range(0, 400, 7):
        pools[t, 1:] = pools[t, :-1]
        pools[t, 0] = oracle.pack(np.float32([0.0]), [q[t]])[0]
    osz = len(x) - 50
    g = device.prune(pg, pools, plen, q, osz, p, reversed_=reversed_)
    c = oracle.prune(pg, x, pools, plen, q, osz, p, reversed_=reversed_)
    assert np.array_equal(g[2], c[2])
    for t in range(400):
        k = c[2][t]
        assert np.array_equal(g[0][t, :k], c[0][t, :k]), t
        assert np.array_equal(g[1][t, :k].view(np.uint32), c[1][t, :k].view(np.uint32)), t
    assert g[3] == c[3]


def test_prune_alpha_degenerate_cases(device):
    """Zero distances (duplicate rows) hit the 0/0 and x/0 paths of the
    alpha rule's cosine; GPU and oracle must agree on every one."""
    rng = np.random.default_rng(4)
    base = rng.integers(0, 4, size=(60, 8)).astype(np.float32)
    x = np.concatenate([base, base[:30]])  # duplicates
    device.upload(x)
    p = IndexParams(nsg_R=20, search_K=100, angle=60)
    pools, plen, q = make_pools(x, 80, 60, rng)
    g = device.prune(FK_PG_ALPHAPG, pools, plen, q, len(x), p)
    c = oracle.prune(FK_PG_ALPHAPG, x, pools, plen, q, len(x), p)
    assert np.array_equal(g[2], c[2])
    for t in range(80):
        assert np.array_equal(g[0][t, :c[2][t]], c[0][t, :c[2][t]])


@pytest.mark.parametrize("name", ["int128", "unit96", "float960", "odd20"])
def test_knn_bruteforce_matches_oracle(device, name):
    x = datasets()[name]
    say("knn", name)
    device.upload(x)
    rng = np.random.default_rng(5)
    sel = rng.choice(len(x), 37, replace=False).astype(np.uint32)
    gi, gd = device.knn(x[sel], 100, self_ids=sel)
    ci, cd = oracle.knn(x, x[sel], 100, self_ids=sel)
    assert np.array_equal(gi, ci)
    assert np.array_equal(gd.view(np.uint32), cd.view(np.uint32))


def knn_recall(pool_ids, gt, K):
    return float(np.mean([len(set(pool_ids[i, :K].tolist()) & set(gt[i, :K].tolist())) / K
                          for i in range(len(gt))]))


def test_nndescent_recall_not_worse_than_oracle(device):
    x = datasets()["int128"]
    say("nndescent int128")
    device.upload(x)
    p = IndexParams(K=30, L=40, S=10, R=50, iterations=6, controls=50)
    gp, gl, gst = device.nndescent(p)
    cp, cl, cst = oracle.nndescent(x, p, threads=8)
    gt, _ = oracle.knn(x, x, 30, self_ids=np.arange(len(x)))
    _, gids = oracle.unpack(gp)
    _, cids = oracle.unpack(cp)
    rg = knn_recall(gids, gt, 30)
    rc = knn_recall(cids, gt, 30)
    say("recall gpu", rg, "oracle", rc, "joins", gst["d_join"], cst["d_join"])
    assert rg >= rc - 0.01
    gd, _ = oracle.unpack(gp)
    for i in range(len(x)):
        row = gids[i, :gl[i]]
        assert i not in row
        assert len(set(row.tolist())) == len(row)
        assert np.all(np.diff(gd[i, :gl[i]]) >= 0)
        # pool distances are the exact distances of their ids
    sel = np.arange(0, len(x), 97)
    d_check = oracle.pair_distances(x, np.repeat(sel, 5), gids[sel, :5].ravel())
    assert np.array_equal(d_check.view(np.uint32), gd[sel, :5].ravel().view(np.uint32))


def reachable(off, ids, ep):
    n = len(off) - 1
    seen = np.zeros(n, bool)
    seen[ep] = True
    st = [ep]
    while st:
        u = st.pop()
        for v in ids[off[u]:off[u + 1]]:
            if not seen[v]:
                seen[v] = True
                st.append(int(v))
    return seen


@pytest.mark.parametrize("pg", [FK_PG_NSG, FK_PG_TAUMNG, FK_PG_ALPHAPG, FK_PG_HNSW_LAYER])
def test_full_build_invariants_and_recall(device, pg):
    x = datasets()["int128"]
    say("build", pg)
    device.upload(x)
    p = IndexParams(K=60, L=60, S=10, R=60, iterations=5, nsg_R=24, search_L=40, search_K=60, tau=20.0, angle=60)
    off, ids, ep, width, st = device.build(pg, p)
    n = len(x)
    for i in range(n):
        nb = ids[off[i]:off[i + 1]]
        assert i not in nb and len(set(nb.tolist())) == len(nb)
    assert np.diff(off).max() <= width
    if pg != FK_PG_HNSW_LAYER:
        assert reachable(off, ids, ep).all()
    # query recall@10 with the same evaluator for GPU and oracle graphs
    rng = np.random.default_rng(6)
    qs = (x[rng.choice(n, 200, replace=False)] + rng.normal(0, 3, size=(200, 128))).astype(np.float32)
    gt, _ = oracle.knn(x, qs, 10)
    res, _ = device.search(off, ids, ep, qs, 40, 10)
    rg = np.mean([len(set(res[i]) & set(gt[i])) / 10 for i in range(200)])
    coff, cids, cep, cw, cst, _ = oracle.build(pg, x, p, threads=8)
    resc, _ = device.search(coff, cids, cep, qs, 40, 10)
    rc = np.mean([len(set(resc[i]) & set(gt[i])) / 10 for i in range(200)])
    say("query recall@10 gpu", rg, "oracle", rc)
    assert rg >= rc - 0.02


@pytest.mark.parametrize("pg", [FK_PG_NSG, FK_PG_ALPHAPG, FK_PG_HNSW_LAYER])
def test_full_build_recall_within_seed_band(device, pg):
    """GPU and oracle builds differ only through their random streams (every
    deterministic step is bit-exact in tests/test_gpu_steps.py), so their
    quality must agree within the spread the seeds themselves produce.  On
    unclustered normal data (no islands: query recall is not decided by
    which cluster the entry point sits in) both sides build with three
    seeds; query recall@10 (one evaluator) and KNN recall after NN-descent
    must agree two-sided within the larger seed range plus 0.01."""
    rng = np.random.default_rng(71)
    x = rng.normal(size=(4000, 32)).astype(np.float32)
    qs = rng.normal(size=(300, 32)).astype(np.float32)
    say("recall band", pg)
    device.upload(x)
    gt, _ = oracle.knn(x, qs, 10)

    def qrec(off, ids, ep):
        res, _ = device.search(off, ids, ep, qs, 40, 10)
        return float(np.mean([len(set(res[i]) & set(gt[i])) / 10 for i in range(len(qs))]))

    g_q, c_q, g_k, c_k = [], [], [], []
    for seed in (1, 2, 3):
        p = IndexParams(K=40, L=40, S=10, R=40, iterations=5, nsg_R=16, search_L=40, search_K=40,
                        angle=50 if pg == FK_PG_ALPHAPG else 60, seed=seed)
        off, ids, ep, _, st = device.build(pg, p)
        g_q.append(qrec(off, ids, ep))
        g_k.append(float(st["recall"][int(st["nnd_iterations"])]))
        coff, cids, cep, _, cst, _ = oracle.build(pg, x, p, threads=8)
        c_q.append(qrec(coff, cids, cep))
        c_k.append(float(cst["recall"][int(cst["nnd_iterations"])]))
    say("  query recall gpu", g_q, "oracle", c_q)
    say("  knn recall after nn-descent gpu", g_k, "oracle", c_k)
    for g, c in ((g_q, c_q), (g_k, c_k)):
        band = max(max(g) - min(g), max(c) - min(c)) + 0.01
        assert abs(np.mean(g) - np.mean(c)) <= band, (g, c)


@pytest.mark.parametrize("pg", [FK_PG_NSG, FK_PG_TAUMNG, FK_PG_ALPHAPG])
@pytest.mark.parametrize("name", ["int128", "unit96", "odd44"])
def test_prune_variants_give_the_same_build(device, pg, name):
    """Builds with either prune variant forced and with the variant picked
    by timing are the same graph with the same statistics."""
    x = datasets()[name]
    say("variants", pg, name)
    device.upload(x)
    p = IndexParams(K=40, L=40, S=10, R=40, iterations=4, nsg_R=16, search_L=30, search_K=40, loop_i=1,
                    tau={"int128": 20.0, "unit96": 0.01, "odd44": 5.0}[name], angle=60)
    out = []
    for flags in (_ffi.FK_FLAG_PRUNE_V0, _ffi.FK_FLAG_PRUNE_V1, _ffi.FK_FLAG_PRUNE_V2, 0):
        off, ids, ep, width, st = device.build(pg, p.copy(flags=flags))
        out.append((off, ids, ep, st))
    for off, ids, ep, st in out[1:]:
        assert np.array_equal(off, out[0][0]) and np.array_equal(ids, out[0][1]) and ep == out[0][2]
        for k in ("d_prune", "d_prune_rev", "d_bv", "d_inter", "aod"):
            assert st[k] == out[0][3][k], k
    # the Gram-form prune inside a Gram build (whose graph is the plain one)
    off, ids, ep, width, st = device.build(pg, p.copy(flags=_ffi.FK_FLAG_GRAM | _ffi.FK_FLAG_PRUNE_V3))
    assert np.array_equal(off, out[0][0]) and np.array_equal(ids, out[0][1]) and ep == out[0][2]
    for k in ("d_prune", "d_prune_rev", "d_bv", "aod"):
        assert st[k] == out[0][3][k], k
    assert st["prune_variant"] == 4
    assert [o[3]["prune_variant"] for o in out[:3]] == [1, 2, 3]
    assert out[3][3]["prune_variant"] in (1, 2, 3)
    say("variants", pg, name, "picked", out[3][3]["prune_variant"] - 1, "tuning s", out[3][3]["t_tune"],
        "uncomputed (variant 2)", out[2][3]["d_prune_uncomputed"], "of", out[2][3]["d_prune"] + out[2][3]["d_prune_rev"])


@pytest.mark.parametrize("name", ["odd44", "odd20", "float960"])
def test_full_build_odd_widths(device, name):
    """Rows whose padded width is not a multiple of 32 (octet distances
    with the canonical tail blocks, generic chain path) and wide rows (d =
    960: tiled join block, wide brute-force controls)."""
    x = datasets()[name]
    say("build odd", name, x.shape)
    device.upload(x)
    n, d = x.shape
    p = IndexParams(K=40, L=40, S=10, R=40, iterations=5, nsg_R=16, search_L=40, search_K=40)
    off, ids, ep, width, st = device.build(FK_PG_NSG, p)
    for i in range(n):
        nb = ids[off[i]:off[i + 1]]
        assert i not in nb and len(set(nb.tolist())) == len(nb)
    assert reachable(off, ids, ep).all()
    rng = np.random.default_rng(7)
    qs = x[rng.choice(n, 100, replace=False)] + rng.normal(0, 0.05, size=(100, d)).astype(np.float32)
    gt, _ = oracle.knn(x, qs, 10)
    res, _ = device.search(off, ids, ep, qs, 40, 10)
    rg = np.mean([len(set(res[i]) & set(gt[i])) / 10 for i in range(100)])
    coff, cids, cep, _, _, _ = oracle.build(FK_PG_NSG, x, p, threads=8)
    resc, _ = device.search(coff, cids, cep, qs, 40, 10)
    rc = np.mean([len(set(resc[i]) & set(gt[i])) / 10 for i in range(100)])
    say("query recall@10 gpu", rg, "oracle", rc)
    assert rg >= rc - 0.03


_C2_CACHE = {}


def _c2_dumped_pool(device, p, n_dump=10_000):
    if "pool" not in _C2_CACHE:
        x = integer_mixture(1_000_000, 128, 1024, seed=1)
        device.upload(x)
        pools, plen, _ = device.nndescent(p)
        _C2_CACHE["pool"] = (x, np.ascontiguousarray(pools[:n_dump]), plen[:n_dump].copy())
        del pools
    return _C2_CACHE["pool"]


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("pg", [FK_PG_NSG, FK_PG_ALPHAPG])
def test_prune_bitexact_c2_dumped_pool(device, pg, variant):
    """The 1M x 128 NSG workload (1,024-centre integer mixture, seed 1,
    K=L=500 S=12 R=100 I=6, nsg_R=50, search_K=500, angle 60): the pools of
    nodes 0..9999 right after NN-descent are pruned on the GPU and by the
    oracle; ids, kept distances and distance counts must agree bit for bit.
    The pool keys themselves are checked against exact distances."""
    from fastkcna_amd.datasets import params_c2
    p = params_c2(IndexParams)
    say("c2 dumped pool", pg)
    x, pools, plen = _c2_dumped_pool(device, p)
    device.upload(x)
    n_dump = len(plen)
    assert (plen == 500).all()
    d, ids = oracle.unpack(pools)
    assert (pools[:, 1:] > pools[:, :-1]).all()  # strictly (dist, id) ordered
    rows = np.repeat(np.arange(0, n_dump, 10), 500)
    exact = oracle.pair_distances(x, rows, ids[::10].ravel())
    assert np.array_equal(exact.view(np.uint32), d[::10].ravel().view(np.uint32))
    q = np.arange(n_dump, dtype=np.uint32)
    p = p.copy(flags=VARIANT_FLAGS[variant])
    g = device.prune(pg, pools, plen, q, len(x), p)
    c = oracle.prune(pg, x, pools, plen, q, len(x), p)
    assert np.array_equal(g[2], c[2])
    for t in range(n_dump):
        k = c[2][t]
        assert np.array_equal(g[0][t, :k], c[0][t, :k]), t
        assert np.array_equal(g[1][t, :k].view(np.uint32), c[1][t, :k].view(np.uint32)), t
    assert g[3] == c[3]
    say("c2 dumped pool: kept", int(c[2].sum()), "distances", c[3])


def test_full_build_hub_in_degree(device):
    """One hub nearest to every other point: its in-degree (1,499) takes the
    large-union path of the reverse-edge merge and the reversed prune; the
    graph keeps the invariants and the oracle's query recall."""
    rng = np.random.default_rng(9)
    n, d = 1500, 32
    u = rng.normal(size=(n - 1, d))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    x = np.zeros((n, d), np.float32)
    x[1:] = (u * (100.0 + rng.uniform(0.0, 1.0, size=(n - 1, 1)))).astype(np.float32)
    say("build hub", x.shape)
    device.upload(x)
    p = IndexParams(K=40, L=40, S=10, R=40, iterations=5, nsg_R=16, search_L=40, search_K=40)
    off, ids, ep, width, st = device.build(FK_PG_NSG, p)
    for i in range(n):
        nb = ids[off[i]:off[i + 1]]
        assert i not in nb and len(set(nb.tolist())) == len(nb)
    assert np.diff(off).max() <= width
    assert reachable(off, ids, ep).all()
    into_hub = int(np.count_nonzero(ids == 0))
    say("out-lists holding the hub", into_hub, "reversed-prune distances", st["d_prune_rev"])
    assert into_hub > 256 and st["d_prune_rev"] > 0
    qs = (x[rng.choice(n, 100, replace=False)] * 0.98).astype(np.float32)
    gt, _ = oracle.knn(x, qs, 10)
    res, _ = device.search(off, ids, ep, qs, 40, 10)
    rg = np.mean([len(set(res[i]) & set(gt[i])) / 10 for i in range(100)])
    coff, cids, cep, _, _, _ = oracle.build(FK_PG_NSG, x, p, threads=8)
    resc, _ = device.search(coff, cids, cep, qs, 40, 10)
    rc = np.mean([len(set(resc[i]) & set(gt[i])) / 10 for i in range(100)])
    say("query recall@10 gpu", rg, "oracle", rc)
    assert rg >= rc - 0.03
