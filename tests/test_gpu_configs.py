"""The benchmark configurations at their full sizes on the GPU, each checked
against the oracle where the oracle can run it and by what the mathematics
fixes where it cannot.

C3  1,000,000 x 960 float mixture, tau-MNG tau = 10, K = L = 500 S = 12
    R = 100 I = 6 nsg_R = 50 (src/kgraph.cpp:2833-3141, :3854-3909): GPU
    NN-descent over all rows; the pools of nodes 0..9999 pruned by the GPU and
    the oracle with tau = 10 (on this data no candidate is occluded: the prune
    keeps the first nsg_R candidates, a closed form), tau = 0.5 and tau = 0
    (occlusion happens), bit-exact; the BF16 Gram-form distance against FP64 on
    1,000,000 random row pairs; NN-descent recall against the oracle's on a
    50,000-row sample.
C1  100,000 x 128, fasthnsw K = L = 200 S = 10 R = 100 I = 6 nsg_R = 40
    (src/fasthnsw.cpp:136-342): the fasthnsw tool's hnswlib index and an HNSW
    index built level by level by the oracle from the same level assignment,
    searched by one evaluator (greedy descent, level-0 beam of 40); and the
    level-0 layer's KNN recall after NN-descent and after each BridgeView.
"""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import oracle
from conftest import ROOT
from fastkcna_amd import FK_PG_HNSW_LAYER, FK_PG_TAUMNG, IndexParams, _ffi, hnsw
from fastkcna_amd.datasets import c1, c3, params_c1, params_c3
from fastkcna_amd.formats import write_lshkit
from test_gpu_cli import read_hnsw
from test_host_cli import minstd_levels

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "fastkcna_amd", "bin")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


# ---------------------------------------------------------------------------
# C3
# ---------------------------------------------------------------------------
_C3 = {}


def c3_state(device):
    """C3's rows (uploaded) and the GPU NN-descent pools of nodes 0..9999."""
    if "x" not in _C3:
        say("c3: generating 1,000,000 x 960")
        x, _ = c3(n_queries=0)
        device.upload(x)
        p = params_c3(IndexParams)
        say("c3: NN-descent")
        pools, plen, st = device.nndescent(p)
        _C3.update(x=x, pools=np.ascontiguousarray(pools[:10_000]), plen=plen[:10_000].copy(), st=st)
        # KNN recall@500 of 100 nodes' pools against brute force
        rng = np.random.default_rng(3)
        ctrl = rng.choice(len(x), 100, replace=False).astype(np.uint32)
        gt, _ = device.knn(x[ctrl], 500, self_ids=ctrl)
        _, ids = oracle.unpack(pools[ctrl])
        _C3["knn_recall"] = float(np.mean([len(set(ids[k].tolist()) & set(gt[k].tolist())) / 500
                                           for k in range(100)]))
        del pools
    device.upload(_C3["x"])
    return _C3


def test_c3_dumped_pool_prune(device):
    s = c3_state(device)
    x, pools, plen = s["x"], s["pools"], s["plen"]
    say("c3: NN-descent stats", {k: s["st"][k] for k in ("nnd_iterations", "d_join", "t_total")},
        "knn recall@500", s["knn_recall"])
    assert s["knn_recall"] > 0.5
    n_dump = len(plen)
    assert (plen == 500).all()
    assert (pools[:, 1:] > pools[:, :-1]).all()
    d, ids = oracle.unpack(pools)
    rows = np.repeat(np.arange(0, n_dump, 50), 500)
    exact = oracle.pair_distances(x, rows, ids[::50].ravel())
    assert np.array_equal(exact.view(np.uint32), d[::50].ravel().view(np.uint32))
    q = np.arange(n_dump, dtype=np.uint32)
    p10 = params_c3(IndexParams)
    for tau in (10.0, 0.05, 0.0):
        p = p10.copy(tau=tau)
        say("c3: tau-MNG prune of the dumped pools, tau", tau)
        g = device.prune(FK_PG_TAUMNG, pools, plen, q, len(x), p)
        c = oracle.prune(FK_PG_TAUMNG, x, pools, plen, q, len(x), p)
        assert np.array_equal(g[2], c[2])
        for t in range(n_dump):
            k = c[2][t]
            assert np.array_equal(g[0][t, :k], c[0][t, :k]), (tau, t)
            assert np.array_equal(g[1][t, :k].view(np.uint32), c[1][t, :k].view(np.uint32)), (tau, t)
        assert g[3] == c[3]
        say("c3: tau", tau, "kept", int(c[2].sum()), "distances", c[3])
        if tau == 10.0:
            # 3 tau = 30 exceeds every candidate distance here, so nothing is
            # occluded: the first nsg_R candidates other than the node itself,
            # each checked against all kept before it
            assert d.max() < 30.0
            for t in range(n_dump):
                first = [v for v in ids[t] if v != t][:50]
                assert list(c[0][t, :c[2][t]]) == first, t
            assert c[3] == n_dump * (50 * 49 // 2)
        elif tau == 0.0:
            assert c[2].sum() < n_dump * 50  # the NSG rule occludes here


def test_c3_bf16_gram_error_against_fp64(device):
    s = c3_state(device)
    x = s["x"]
    rng = np.random.default_rng(4)
    n = 1_000_000
    a = rng.integers(0, len(x), n).astype(np.uint32)
    b = rng.integers(0, len(x), n).astype(np.uint32)
    gd, lo, hi, mode = device.gram_pair_distances(a, b, flags=_ffi.FK_FLAG_GRAM | _ffi.FK_FLAG_GRAM_BF16)
    assert mode == _ffi.FK_GRAM_BF16
    fd = device.pair_distances(a, b)
    cf, c64 = oracle.pair_distances(x, a, b, with_f64=True)
    assert np.array_equal(fd.view(np.uint32), cf.view(np.uint32))  # the FP32 path is the canonical distance
    assert np.all(lo <= cf) and np.all(cf <= hi)  # the recheck bound holds on every pair
    nz = c64 > 0
    rel_bf16 = np.abs(gd[nz] - c64[nz]) / c64[nz]
    rel_fp32 = np.abs(cf[nz].astype(np.float64) - c64[nz]) / c64[nz]
    say("c3: BF16 Gram vs FP64 relative error: median %.2e p99 %.2e max %.2e; FP32 max %.2e"
        % (np.median(rel_bf16), np.quantile(rel_bf16, 0.99), rel_bf16.max(), rel_fp32.max()))
    assert rel_fp32.max() < 1e-5  # stated tolerance of the FP32 path against FP64
    assert np.median(rel_bf16) < 2e-3 and rel_bf16.max() < 5e-2


def test_c3_sample_nndescent_not_worse_than_oracle(device):
    x, _ = c3(n=50_000, n_queries=0)
    device.upload(x)
    p = params_c3(IndexParams)
    say("c3 sample: GPU and oracle NN-descent on 50,000 x 960")
    gp, gl, gst = device.nndescent(p)
    cp, cl, cst = oracle.nndescent(x, p, threads=0)
    rng = np.random.default_rng(5)
    sel = rng.choice(len(x), 300, replace=False).astype(np.uint32)
    gt, _ = device.knn(x[sel], 500, self_ids=sel)
    _, gi = oracle.unpack(gp[sel])
    _, ci = oracle.unpack(cp[sel])
    rg = float(np.mean([len(set(gi[k].tolist()) & set(gt[k].tolist())) / 500 for k in range(300)]))
    rc = float(np.mean([len(set(ci[k].tolist()) & set(gt[k].tolist())) / 500 for k in range(300)]))
    say("c3 sample: recall@500 gpu", rg, "oracle", rc, "joins", gst["d_join"], cst["d_join"])
    assert rg >= rc - 0.01


# ---------------------------------------------------------------------------
# C1
# ---------------------------------------------------------------------------
def csr(lists, n):
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(v) for v in lists])
    return off, (np.concatenate(lists) if off[-1] else np.zeros(0, np.uint32)).astype(np.int64)


def recall10(found, gt):
    return float(np.mean([len(set(f) & set(g.tolist())) / 10 for f, g in zip(found, gt)]))


def test_c1_fasthnsw_query_recall_vs_oracle(device, tmp_path):
    x, qs = c1()
    n, d = x.shape
    p = params_c1(IndexParams)
    device.upload(x)
    gt, _ = device.knn(qs, 10)
    # the GPU tool
    data, out, res = tmp_path / "c1.lshkit", tmp_path / "c1.hnsw", tmp_path / "c1.csv"
    write_lshkit(data, x)
    args = ["-K", p.K, "-L", p.L, "-S", p.S, "-R", p.R, "-I", p.iterations, "--nsg_R", p.nsg_R, "--search_L",
            p.search_L, "--search_K", p.search_K, "--result", res]
    say("c1: fasthnsw", args)
    r = subprocess.run([os.path.join(BIN, "fasthnsw"), str(data), str(out), *map(str, args)], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    hdr, lvl0, links = read_hnsw(out, d)
    spe, R = hdr["size_per_element"], hdr["maxM"]
    assert R == p.nsg_R and hdr["maxM0"] == 2 * p.nsg_R
    vec = np.zeros((n, d), np.float32)
    lab = np.zeros(n, np.int64)
    nb0 = []
    for i in range(n):
        rec = lvl0[i * spe:(i + 1) * spe]
        cnt = struct.unpack_from("<H", rec, 0)[0]
        nb0.append(np.frombuffer(rec, np.uint32, cnt, 4))
        vec[i] = np.frombuffer(rec, np.float32, d, hdr["offsetData"])
        lab[i] = struct.unpack_from("<Q", rec, hdr["label_offset"])[0]
    per = 4 + 4 * R
    graphs = [csr(nb0, n)]
    for level in range(1, hdr["maxlevel"] + 1):
        lists = []
        for i in range(n):
            blk = links[i][(level - 1) * per:level * per]
            c = struct.unpack_from("<H", blk, 0)[0] if blk else 0
            lists.append(np.frombuffer(blk, np.uint32, c, 4) if c else np.zeros(0, np.uint32))
        graphs.append(csr(lists, n))
    found = [lab[hnsw.search(graphs, hdr["enterpoint"], vec, q, p.search_L, 10)].tolist() for q in qs]
    rg = recall10(found, gt)
    # the oracle, level by level from the same level assignment
    lv = hnsw.levels(n, p.nsg_R, p.seed)
    assert np.array_equal(lv, np.array(minstd_levels(n, p.nsg_R, p.seed)))
    order = hnsw.level_order(lv)
    assert np.array_equal(vec, x[order])  # the tool's internal order is the same
    xr = np.ascontiguousarray(x[order])

    def oracle_layer(pl, size):
        say("c1: oracle layer on", size, "rows")
        off, ids, ep, *_ = oracle.build(FK_PG_HNSW_LAYER, xr, pl, n_used=size, threads=0)
        return off, ids, ep

    cgraphs, top_ep, _ = hnsw.build_layers(oracle_layer, n, lv, p)
    top = len(cgraphs) - 1
    cfound = [order[hnsw.search(cgraphs, top_ep, xr, q, p.search_L, 10)].tolist() for q in qs]
    rc = recall10(cfound, gt)
    say("c1: query recall@10 (ef 40) gpu", rg, "oracle", rc, "levels", top + 1)
    assert rg >= rc - 0.01


def test_c1_level0_layer_recall_vs_oracle(device):
    x, _ = c1()
    n = len(x)
    p = params_c1(IndexParams).copy(nsg_R=80)
    device.upload(x)
    say("c1: level-0 layer builds (nsg_R 80)")
    off, ids, ep, width, gst = device.build(FK_PG_HNSW_LAYER, p)
    coff, cids, cep, cwidth, cst, _ = oracle.build(FK_PG_HNSW_LAYER, x, p, threads=0)
    deg = np.diff(off)
    assert deg.max() <= 80 and width == 80
    for i in range(0, n, 97):
        nb = ids[off[i]:off[i + 1]]
        assert i not in nb and len(set(nb.tolist())) == len(nb)
    g_rec = [float(v) for v in gst["recall"][gst["nnd_iterations"]:gst["n_recall"]]]
    c_rec = [float(v) for v in cst["recall"][cst["nnd_iterations"]:cst["n_recall"]]]
    say("c1: KNN recall@MAXL after NN-descent and each BridgeView: gpu", g_rec, "oracle", c_rec)
    assert len(g_rec) == len(c_rec) == p.loop_i + 1
    for a, b in zip(g_rec, c_rec):
        assert a >= b - 0.01


# ---------------------------------------------------------------------------
# C4
# ---------------------------------------------------------------------------
def reachable(off, ids, ep):
    """Nodes reachable from ep over the directed CSR graph (SciPy BFS)."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import breadth_first_order
    n = len(off) - 1
    g = csr_matrix((np.ones(len(ids), np.int8), ids.astype(np.int64), off.astype(np.int64)), shape=(n, n))
    seen = np.zeros(n, bool)
    seen[breadth_first_order(g, ep, directed=True, return_predecessors=False)] = True
    return seen


def test_c4_all_sets_batched(device):
    """All 256 datasets of the batching workload (unit-normalised 16-centre
    mixtures, n log-uniform in [2,000, 200,000], d in {96, 128, 256}),
    alpha-PG K = L = 100 S = 10 R = 50 I = 6 nsg_R = 32, one fk_build_batch per
    width.  Every set's graph: no self edges or duplicates, degrees within the
    width, every node reachable from its entry point; query recall@10 of its
    1,000 queries (search_L 40) against brute force, reported; six sets
    bit-identical to their standalone builds; the two smallest sets of each
    width no worse than the oracle's builds of them under the same evaluator
    (these nearly uniform high-dimensional sets are hard: the recall is a
    property of the data, and parity with the reference is the bar)."""
    from fastkcna_amd import FK_PG_ALPHAPG
    from fastkcna_amd.datasets import c4_sets, params_c4
    say("c4: generating 256 sets")
    sets = c4_sets()
    p = params_c4(IndexParams)
    graphs = [None] * len(sets)
    for d in sorted({x.shape[1] for x, _ in sets}):
        idx = [k for k, (x, _) in enumerate(sets) if x.shape[1] == d]
        rows = np.concatenate([sets[k][0] for k in idx])
        off = np.zeros(len(idx) + 1, np.uint64)
        off[1:] = np.cumsum([len(sets[k][0]) for k in idx])
        device.upload(rows)
        say("c4: batch of", len(idx), "sets, d", d, "rows", len(rows))
        got, st = device.build_batch(FK_PG_ALPHAPG, p, off)
        for j, k in enumerate(idx):
            graphs[k] = got[j]
        del rows
    recalls = []
    for k, (x, q) in enumerate(sets):
        off, ids, ep, width = graphs[k]
        n = len(x)
        assert len(off) == n + 1 and np.diff(off.astype(np.int64)).max() <= width
        src = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
        assert not np.any(ids == src)
        key = src.astype(np.int64) * n + ids
        assert len(np.unique(key)) == len(key)
        assert reachable(off, ids, ep).all(), k
        device.upload(x)
        gt, _ = device.knn(q, 10)
        res, _ = device.search(off, ids, ep, q, 40, 10)
        recalls.append(recall10(res.tolist(), gt))
    recalls = np.array(recalls)
    say("c4: query recall@10 mean %.4f min %.4f max %.4f" % (recalls.mean(), recalls.min(), recalls.max()))
    sizes = np.array([len(x) for x, _ in sets])
    for k in [int(np.argmin(sizes)), int(np.argmax(sizes)), 3, 77, 150, 255]:
        x = sets[k][0]
        device.upload(x)
        off, ids, ep, width, _ = device.build(FK_PG_ALPHAPG, p)
        g = graphs[k]
        assert np.array_equal(off, g[0]) and np.array_equal(ids, g[1]) and ep == g[2] and width == g[3], k
    # the oracle on the two smallest sets of each width, the same evaluator
    gr, cr = [], []
    widths = np.array([x.shape[1] for x, _ in sets])
    for d in (96, 128, 256):
        pick = [k for k in np.argsort(sizes) if widths[k] == d][:2]
        for k in pick:
            x, q = sets[k]
            device.upload(x)
            gt, _ = device.knn(q, 10)
            coff, cids, cep, *_ = oracle.build(FK_PG_ALPHAPG, x, p, threads=0)
            cr.append(recall10(device.search(coff, cids, cep, q, 40, 10)[0].tolist(), gt))
            gr.append(float(recalls[k]))
    say("c4: smallest sets per width, recall gpu", gr, "oracle", cr)
    assert np.mean(gr) >= np.mean(cr) - 0.01


# ---------------------------------------------------------------------------
# C5
# ---------------------------------------------------------------------------
def test_c5_level0_plain_and_screen(device):
    """10,000,000 x 96 unit-normalised rows (4,096 centres, sigma 0.25 per
    coordinate: nearly uniform on the sphere, where six NN-descent iterations
    are far from converged for the reference as for this build).  The HNSW
    level-0 layer (nsg_R 64) built plain and with the learned screen: graph
    invariants on a sample of nodes; the screen's KNN recall over the
    controls no worse than plain - 0.01, its skipped pairs and audited losses
    reported.  On the first 200,000 rows the GPU and the oracle build the
    same layer: KNN recall after NN-descent and query recall@10 of 2,000
    queries (search_L 40, one evaluator) no worse than the oracle's - 0.01."""
    from fastkcna_amd.datasets import c5, params_c5
    say("c5: generating 10,000,000 x 96")
    x, q = c5(n_queries=2_000)
    p = params_c5(IndexParams)
    # the 200,000-row sample against the oracle
    xs = np.ascontiguousarray(x[:200_000])
    device.upload(xs)
    gt, _ = device.knn(q, 10)
    say("c5 sample: GPU and oracle level-0 builds on 200,000 rows")
    off, ids, ep, _, gst = device.build(FK_PG_HNSW_LAYER, p)
    coff, cids, cep, _, cst, _ = oracle.build(FK_PG_HNSW_LAYER, xs, p, threads=0)
    rg = recall10(device.search(off, ids, ep, q, 40, 10)[0].tolist(), gt)
    rc = recall10(device.search(coff, cids, cep, q, 40, 10)[0].tolist(), gt)
    kg, kc = float(gst["recall"][gst["nnd_iterations"]]), float(cst["recall"][cst["nnd_iterations"]])
    say("c5 sample: KNN recall after NN-descent gpu %.4f oracle %.4f; query recall@10 gpu %.4f oracle %.4f"
        % (kg, kc, rg, rc))
    assert kg >= kc - 0.01 and rg >= rc - 0.01
    del xs
    # the full layer
    device.upload(x)
    say("c5: plain level-0 build")
    off, ids, ep, width, st = device.build(FK_PG_HNSW_LAYER, p)
    n = len(x)
    deg = np.diff(off.astype(np.int64))
    assert width == 64 and deg.max() <= 64 and deg.min() >= 1
    for i in range(0, n, 100_003):
        nb = ids[off[i]:off[i + 1]]
        assert i not in nb and len(set(nb.tolist())) == len(nb)
    knn_plain = float(st["recall"][st["nnd_iterations"]])
    gt, _ = device.knn(q, 10)
    rq = recall10(device.search(off, ids, ep, q, 40, 10)[0].tolist(), gt)
    say("c5: plain build %.2f s, KNN recall@200 after NN-descent %.4f, final %.4f, query recall@10 %.4f"
        % (st["t_total"], knn_plain, st["recall"][st["n_recall"] - 1], rq))
    del off, ids
    say("c5: level-0 build with the learned screen")
    ps = p.copy(flags=_ffi.FK_FLAG_SCREEN)
    soff, sids, sep, swidth, sst = device.build(FK_PG_HNSW_LAYER, ps)
    knn_screen = float(sst["recall"][sst["nnd_iterations"]])
    lost = 64.0 * sst["screen_missed"] / max(sst["screen_cands"], 1)
    say("c5: screen build %.2f s, KNN recall after NN-descent %.4f, screened %d skipped %d (est. lost candidates "
        "%.2e), margin %.3f" % (sst["t_total"], knn_screen, sst["d_screened"], sst["d_screen_skip"], lost,
                                sst["screen_margin"]))
    assert knn_screen >= knn_plain - 0.01
