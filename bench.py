"""Benchmark: FastKCNA NSG build of 1,000,000 x 128 float32 points on one MI355X.

One step = one full graph build (sampled ground truth, NN-descent K=L=500 S=12
R=100 I=6, three NSG prune loops with reverse edges, entry point,
connectivity and two BridgeView passes) over the resident dataset, ending
with the graph in host memory.  The data is the integer-valued mixture of
1,024 Gaussians (seed 1) with 10,000 held-out queries; query recall@10 at
search_L = 80 is evaluated after the timed region.

Prints one JSON line (metric, value = distance evaluations per second of the
whole build, ms_per_step = build milliseconds, ...), with a `roofline` object
for the dominant kernel (timed with HIP events on the library's stream) and a
`cpu_baseline` object (the CPU oracle, the reference's algorithm restated in
C++/OpenMP, timed on this host on a bounded sample).
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "build sec @ recall@10 parity (1M×128 NSG); dist-evals/s; chosen vs plain FP32"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def recall_at_10(res, gt):
    return float(np.mean([len(set(res[i].tolist()) & set(gt[i].tolist())) / 10.0 for i in range(len(gt))]))


def recall_at_10_ties(base, queries, res, gt_d):
    """Tie-aware recall@10: results whose squared distance to the query is at
    most the 10th ground-truth distance, over 10 (integer data has many equal
    distances, so an id-based recall can miss an equally near result).  The
    distances of the results are taken in float64 on the host: exact on
    integer rows, within 1e-6 relative on float rows."""
    q = queries.astype(np.float64)
    hits = 0
    for i0 in range(0, len(res), 2048):
        r = res[i0:i0 + 2048]
        d = ((base[r].astype(np.float64) - q[i0:i0 + 2048, None, :]) ** 2).sum(-1)
        lim = gt_d[i0:i0 + 2048, 9].astype(np.float64) * (1 + 1e-6)
        hits += int(np.minimum((d <= lim[:, None]).sum(1), 10).sum())
    return hits / (10.0 * len(res))


def graph_digest(off, ids, ep):
    """SHA-1 of the CSR graph (offsets, ids, entry point): equal digests mean
    equal graphs, whatever kernels built them."""
    h = hashlib.sha1()
    h.update(np.ascontiguousarray(off, dtype=np.uint64).tobytes())
    h.update(np.ascontiguousarray(ids, dtype=np.uint32).tobytes())
    h.update(np.uint64(ep).tobytes())
    return h.hexdigest()


def total_dists(st, computed=False):
    """Distance evaluations of a build.  By default BridgeView pairs are
    counted as the reference counts them (every pair it evaluates); with
    computed=True only the distances the GPU computed in FP32: BridgeView
    skips targets already pooled or computed in the same round and, with
    FK_FLAG_GRAM, targets a 16-bit bound rejects; the prunes skip distances
    that cannot decide anything (tau-MNG candidates within 3 tau of the owner)
    or that a bound decides."""
    keys = ("d_init", "d_ctrl", "d_join", "d_prune", "d_prune_rev", "d_inter", "d_ep", "d_conn")
    bv = int(st.get("d_bv_computed", st["d_bv"])) - int(st.get("d_bv_rejected", 0)) if computed else int(st["d_bv"])
    skipped = int(st.get("d_prune_uncomputed", 0)) if computed else 0
    return sum(int(st[k]) for k in keys) + bv - skipped


def after_nndescent(st):
    """KNN-graph recall@MAXL over the controls right after NN-descent (the
    entry after the per-iteration recall@K entries)."""
    return float(st["recall"][int(st["nnd_iterations"])])


def host_cpu_info():
    """Logical CPUs, model, sockets, cores per socket and threads per core of
    this host (from lscpu), and the CPUs this process may run on."""
    import subprocess
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Socket(s)":
                info["sockets"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
            elif k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
    except Exception as e:  # lscpu missing: the counts above still hold
        info["lscpu_error"] = str(e)
    return info


def cpu_threads():
    """Threads for the CPU baseline: OMP_NUM_THREADS when set (the share of
    the host's CPUs given to this job), else every CPU this process may use."""
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return env if env > 0 else len(os.sched_getaffinity(0))


# kernel of each timed family (the family's dominant launch)
FAMILY_KERNEL = {"bv": "k_bv_round", "prune": "k_prune", "join": "k_join_count", "join_fill": "k_join_fill",
                 "merge": "k_nnd_merge", "controls": "k_topk_select"}
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "c2_nsg_build_hbm_traffic.md")


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by
    a device copy of known size) from the committed rocprofv3 counter summary
    of this workload; None when the summary has no row for it."""
    if kernel is None or not os.path.exists(TRAFFIC_SUMMARY):
        return None
    for line in open(TRAFFIC_SUMMARY):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if "(reversed)" in cells[0]:  # the family is the main prunes
            continue
        if len(cells) >= 5 and cells[0].strip("`").split("<")[0].endswith(kernel):
            return float(cells[4]) * 1e9
    return None


FP32_PEAK_TFLOPS = 157.3
# distance counters of each family (for its FP32 rate: 2 d flops per distance)
FAMILY_DISTS = {"prune": "d_prune", "prune_rev": "d_prune_rev", "bv": "d_bv_computed", "join": "d_join"}


def roofline(st, families, d):
    """Roofline of the dominant kernel family: algorithmic HBM bytes per
    launch (every row and key the work needs, each counted once) over its
    mean launch time (HIP events on the library's stream).  Beside it the
    same launch judged by its measured HBM traffic (rocprofv3 counters) and
    by its FP32 distance arithmetic: the algorithmic rate counts rows that
    the locality order serves from L2, so it is a consumption rate, and the
    measured-traffic fraction says how much of HBM the kernel really uses."""
    ms = st["prof_ms"]
    i = int(np.argmax(ms[:len(families)]))
    if ms[i] <= 0:
        return None
    launches = max(int(st["prof_launches"][i]), 1)
    gbs = st["prof_bytes"][i] / (ms[i] / 1000.0) / 1e9
    kernel = FAMILY_KERNEL.get(families[i])
    ms_launch = ms[i] / launches
    traffic = measured_traffic(kernel)
    out = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
           "traffic": traffic, "kernel": kernel or families[i], "family": families[i],
           "launches": launches, "bytes_per_launch": st["prof_bytes"][i] / launches,
           "ms_per_launch": ms_launch, "share_of_build": ms[i] / 1000.0 / max(st["t_total"], 1e-9),
           "traffic_source": os.path.relpath(TRAFFIC_SUMMARY, ROOT),
           "achieved_is": "algorithmic bytes (every examined candidate row and key once) per launch time; "
                          "most of them come from the XCD's L2, see measured_traffic_frac"}
    if traffic is not None:
        tg = traffic / (ms_launch / 1000.0) / 1e9
        out.update({"measured_traffic_gbs": tg, "measured_traffic_frac": tg / HBM_PEAK_GBS})
    # whole build: every family's algorithmic bytes at HBM peak against the
    # build's own time
    tb = float(np.sum(st["prof_bytes"][:len(families)]))
    out.update({"build_bytes_alg": tb, "build_roof_sec": tb / (HBM_PEAK_GBS * 1e9),
                "build_frac": tb / (HBM_PEAK_GBS * 1e9) / max(float(st["t_total"]), 1e-9)})
    key = FAMILY_DISTS.get(families[i])
    if key is not None and key in st:
        tf = 2.0 * d * float(st[key]) / (ms[i] / 1000.0) / 1e12
        out.update({"fp32_tflops": tf, "fp32_frac": tf / FP32_PEAK_TFLOPS})
    return out


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB (HIP runtime copy kernel): the
    reference point that scales FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    n = mib << 20
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: {mib} MiB device-to-device")


def cpu_baseline(n_sample, threads, dev, queries):
    """The CPU oracle on a bounded sample of the same workload; the GPU builds
    the same sample and both graphs are searched by the same evaluator."""
    import oracle
    from fastkcna_amd import FK_PG_NSG
    from fastkcna_amd.datasets import integer_mixture, params_c2
    from fastkcna_amd.kgraph import IndexParams
    x = integer_mixture(n_sample, 128, 1024, seed=1)
    p = params_c2(IndexParams)
    t0 = time.perf_counter()
    off, ids, ep, width, st, _ = oracle.build(FK_PG_NSG, x, p, threads=threads)
    dt = time.perf_counter() - t0
    nd = total_dists(st)
    # recall parity on the sample: one evaluator, one ground truth
    dev.upload(x)
    t1 = time.perf_counter()
    goff, gids, gep, _, gst = dev.build(FK_PG_NSG, p)
    gdt = time.perf_counter() - t1
    gt, _ = dev.knn(queries, 10)
    rc = recall_at_10(dev.search(off, ids, ep, queries, 80, 10)[0], gt)
    rg = recall_at_10(dev.search(goff, gids, gep, queries, 80, 10)[0], gt)
    # the same comparison below saturation: recall@10 at narrower beams
    curve = {}
    for sl in (10, 20, 40, 80):
        curve[str(sl)] = {"cpu": recall_at_10(dev.search(off, ids, ep, queries, sl, 10)[0], gt),
                          "gpu": recall_at_10(dev.search(goff, gids, gep, queries, sl, 10)[0], gt)}
    kc, kg = float(st["recall"][st["n_recall"] - 1]), float(gst["recall"][gst["n_recall"] - 1])
    kc0, kg0 = after_nndescent(st), after_nndescent(gst)
    # noise band of the KNN recall: each build estimates it on its own random
    # controls and random streams, so both sides are rebuilt with two more
    # seeds and the spread of each is reported beside the comparison
    band = {"seeds": [int(p.seed)], "cpu_after_nndescent": [kc0], "gpu_after_nndescent": [kg0],
            "cpu_final": [kc], "gpu_final": [kg]}
    rates = [nd / dt]
    for seed in (p.seed + 1, p.seed + 2):
        q = p.copy(seed=seed)
        t2 = time.perf_counter()
        *_, cst, _ = oracle.build(FK_PG_NSG, x, q, threads=threads)
        rates.append(total_dists(cst) / (time.perf_counter() - t2))
        *_, gst2 = dev.build(FK_PG_NSG, q)
        band["seeds"].append(int(seed))
        band["cpu_after_nndescent"].append(after_nndescent(cst))
        band["gpu_after_nndescent"].append(after_nndescent(gst2))
        band["cpu_final"].append(float(cst["recall"][cst["n_recall"] - 1]))
        band["gpu_final"].append(float(gst2["recall"][gst2["n_recall"] - 1]))
    return {"value": float(np.median(rates)), "unit": "dist-evals/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (C++/OpenMP restatement of the reference) NSG build of the first {n_sample:,} "
                      f"points of the same mixture, same parameters; median of 3 runs (seeds {int(p.seed)}, "
                      f"{int(p.seed) + 1}, {int(p.seed) + 2}); first run {dt:.1f} s, {nd:,} distance evaluations",
            "runs_dist_evals_per_sec": rates,
            "build_sec": dt, "host": host_cpu_info(),
            "sample_parity": {"cpu_build_sec": round(dt, 3), "gpu_build_sec": round(gdt, 4),
                              "speedup": round(dt / gdt, 1), "queries": int(len(queries)),
                              "cpu_recall_at_10": rc, "gpu_recall_at_10": rg,
                              "recall_at_10_by_search_L": curve,
                              "cpu_knn_recall_after_nndescent": kc0, "gpu_knn_recall_after_nndescent": kg0,
                              "cpu_knn_recall_at_maxl": kc, "gpu_knn_recall_at_maxl": kg,
                              "knn_recall_seed_band": band}}


def cpu_full_leg(dev, base, queries, gpu_graph, gpu_stats, threads):
    """The CPU oracle's build of the full workload (minutes of CPU time, not
    part of the default run): its build seconds beside the GPU's, and both
    graphs searched by the same evaluator against the same ground truth."""
    import oracle
    from fastkcna_amd import FK_PG_NSG
    from fastkcna_amd.datasets import params_c2
    from fastkcna_amd.kgraph import IndexParams
    p = params_c2(IndexParams)
    log(f"cpu full: oracle build of {len(base):,} rows on {threads} threads")
    t0 = time.perf_counter()
    off, ids, ep, width, st, _ = oracle.build(FK_PG_NSG, base, p, threads=threads, verbose=True)
    dt = time.perf_counter() - t0
    log(f"cpu full: {dt:.1f}s")
    dev.upload(base)
    gt, _ = dev.knn(queries, 10)
    goff, gids, gep = gpu_graph
    # each graph from its own entry point, and each from the other's: the
    # two NN-descent runs end in different (equally good) pools, so the
    # entry-point searches stop at different nodes, and on this data the
    # start decides much of recall@10 at small search_L
    curve = {}
    for sl in (80, 200, 500):
        curve[str(sl)] = {"cpu": recall_at_10(dev.search(off, ids, ep, queries, sl, 10)[0], gt),
                          "gpu": recall_at_10(dev.search(goff, gids, gep, queries, sl, 10)[0], gt),
                          "cpu_from_gpu_ep": recall_at_10(dev.search(off, ids, gep, queries, sl, 10)[0], gt),
                          "gpu_from_cpu_ep": recall_at_10(dev.search(goff, gids, ep, queries, sl, 10)[0], gt)}
    return {"cpu_build_sec": dt, "threads": threads, "host": host_cpu_info(), "cpu_ep": int(ep), "gpu_ep": int(gep),
            "cpu_dist_evals": total_dists(st), "gpu_build_sec": float(gpu_stats["t_total"]),
            "recall_at_10_by_search_L": curve,
            "cpu_knn_recall_after_nndescent": after_nndescent(st),
            "gpu_knn_recall_after_nndescent": after_nndescent(gpu_stats),
            "cpu_knn_recall_at_maxl": float(st["recall"][st["n_recall"] - 1]),
            "gpu_knn_recall_at_maxl": float(gpu_stats["recall"][gpu_stats["n_recall"] - 1]),
            "cpu_aod": int(st["aod"]), "gpu_aod": int(gpu_stats["aod"])}


def variant_leg(dev, p, args, extra_flags, plain_graph, gt, queries, plain_sec, label):
    """The same build with extra flags: same steps and warm-up, its graph
    compared with the plain path's and searched by the same evaluator."""
    from fastkcna_amd import FK_PG_NSG, _ffi
    import copy
    q = copy.copy(p)
    q.flags = p.flags | extra_flags
    for _ in range(args.warmup):
        dev.build(FK_PG_NSG, q)
    times, st, g = [], None, None
    for s in range(args.steps):
        t = time.perf_counter()
        off, ids, ep, width, st = dev.build(FK_PG_NSG, q)
        times.append(time.perf_counter() - t)
        g = (off, ids, ep)
        log(f"{label} step {s}: build {times[-1]:.3f}s")
    sec = float(np.mean(times))
    same = bool(np.array_equal(g[0], plain_graph[0]) and np.array_equal(g[1], plain_graph[1])
                and g[2] == plain_graph[2])
    rec = recall_at_10(dev.search(g[0], g[1], g[2], queries, 80, 10)[0], gt)
    out = {"build_sec": sec, "dist_evals_per_sec": total_dists(st) / sec, "speedup_vs_plain": plain_sec / sec,
           "recall_at_10": rec, "graph_identical_to_plain": same,
           "join_ms": st["prof_ms"][_ffi.PROF_FAMILIES.index("join")],
           "knn_recall_at_maxl": float(st["recall"][st["n_recall"] - 1]),
           "families_ms": {f: round(st["prof_ms"][i], 1) for i, f in enumerate(_ffi.PROF_FAMILIES)
                           if st["prof_launches"][i]}}
    return out, st


def chosen_leg(dev, p, args, plain_graph, gt, queries, plain_sec):
    """The chosen path: the local join in Gram form on the MFMA units."""
    from fastkcna_amd import _ffi
    out, st = variant_leg(dev, p, args, _ffi.FK_FLAG_GRAM, plain_graph, gt, queries, plain_sec, "chosen")
    out = {"path": "MFMA Gram local join (FK_FLAG_GRAM), FP32 recheck; everything else as the plain path",
           "gram_mode": _ffi.GRAM_MODE_NAMES[int(st["gram_mode"])], **out,
           "gram_pairs": int(st["d_gram"]), "fp32_rechecks": int(st["d_recheck"])}
    log(f"chosen: {out}")
    return out


def screen_leg(dev, p, args, plain_graph, gt, queries, plain_sec):
    """The plain join with the learned distance screen (FK_FLAG_SCREEN)."""
    from fastkcna_amd import _ffi
    out, st = variant_leg(dev, p, args, _ffi.FK_FLAG_SCREEN, plain_graph, gt, queries, plain_sec, "screen")
    out = {"path": "plain FP32 join with the learned distance screen (FK_FLAG_SCREEN)", **out,
           "screened_pairs": int(st["d_screened"]), "skipped_pairs": int(st["d_screen_skip"]),
           "est_lost_candidate_frac": 64.0 * st["screen_missed"] / max(st["screen_cands"], 1),
           "screen_sec": float(st["t_screen"]), "train_pairs": int(st["screen_train_pairs"]),
           "train_loss": float(st["screen_loss"]), "last_margin": float(st["screen_margin"])}
    log(f"screen: {out}")
    return out


def batching_leg(dev, n_sets, one_by_one=True):
    """The first n_sets datasets of the batching workload (n log-uniform in
    [2,000, 200,000], d in {96, 128, 256}, unit-normalised 16-centre
    mixtures; alpha-PG, K=L=100 S=10 R=50 I=6 nsg_R=32), grouped by width:
    one fk_build_batch per width against one fk_build per dataset (uploads
    outside both timings), and the batch again with FK_FLAG_GRAM (Gram-form
    prune); the graphs are compared.  one_by_one=False skips the per-dataset
    builds (profiling runs)."""
    from fastkcna_amd import FK_PG_ALPHAPG, _ffi
    from fastkcna_amd.datasets import c4_sets, params_c4
    from fastkcna_amd.kgraph import IndexParams
    t = time.perf_counter()
    sets = [x for x, _ in c4_sets(n_sets=n_sets, n_queries=0)]
    log(f"batching: {n_sets} sets, {sum(len(x) for x in sets):,} rows generated in {time.perf_counter() - t:.1f}s")
    p = params_c4(IndexParams)
    pg = p.copy(flags=_ffi.FK_FLAG_GRAM)
    t_batch = t_single = t_gram = 0.0
    same = same_gram = True
    gram_variants = {}
    for d in sorted({x.shape[1] for x in sets}):
        group = [x for x in sets if x.shape[1] == d]
        rows = np.concatenate(group)
        off = np.zeros(len(group) + 1, np.uint64)
        off[1:] = np.cumsum([len(x) for x in group])
        dev.upload(rows)
        dev.build_batch(FK_PG_ALPHAPG, p, off)  # warm-up
        t = time.perf_counter()
        got, _ = dev.build_batch(FK_PG_ALPHAPG, p, off)
        t_batch += time.perf_counter() - t
        dev.build_batch(FK_PG_ALPHAPG, pg, off)  # warm-up (prune variants measured here)
        t = time.perf_counter()
        ggot, gst = dev.build_batch(FK_PG_ALPHAPG, pg, off)
        t_gram += time.perf_counter() - t
        gram_variants[int(d)] = [int(gst["prune_variant"]) - 1, int(gst["prune_rev_variant"]) - 1]
        same_gram = same_gram and all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
                                      for a, b in zip(got, ggot))
        del ggot
        for k, x in enumerate(group if one_by_one else []):
            dev.upload(x)
            t = time.perf_counter()
            o, i, ep, w, _ = dev.build(FK_PG_ALPHAPG, p)
            t_single += time.perf_counter() - t
            same = same and np.array_equal(o, got[k][0]) and np.array_equal(i, got[k][1]) and ep == got[k][2]
    out = {"sets": n_sets, "rows": int(sum(len(x) for x in sets)), "batched_sec": t_batch,
           "one_by_one_sec": t_single if one_by_one else None,
           "speedup": t_single / t_batch if one_by_one else None, "graphs_identical": bool(same),
           "gram": {"batched_sec": t_gram, "speedup_vs_batched": t_batch / t_gram,
                    "speedup_vs_one_by_one": t_single / t_gram if one_by_one else None,
                    "graphs_identical": bool(same_gram),
                    "prune_variants_by_width": gram_variants}}
    log(f"batching: {out}")
    return out


def knn_recall_entries(st):
    """KNN-graph recall@MAXL over the controls after NN-descent and after
    each BridgeView."""
    return [round(float(v), 5) for v in st["recall"][int(st["nnd_iterations"]):int(st["n_recall"])]]


def c1_leg(dev):
    """C1: fasthnsw over 100,000 x 128 (64-centre integer mixture, seed 0),
    K = L = 200 S = 10 R = 100 I = 6 nsg_R = 40: every level built through
    the C ABI on the level-ordered rows (one warm-up), timed; query
    recall@10 of its 1,000 queries through the hnswlib search (ef 40)."""
    from fastkcna_amd import hnsw
    from fastkcna_amd.datasets import c1, params_c1
    from fastkcna_amd.kgraph import IndexParams
    x, qs = c1()
    p = params_c1(IndexParams)
    lv = hnsw.levels(len(x), p.nsg_R, p.seed)
    xr = np.ascontiguousarray(x[hnsw.level_order(lv)])
    dev.upload(xr)
    layer = hnsw.device_layer(dev)
    hnsw.build_layers(layer, len(x), lv, p)
    graphs, ep, sec = hnsw.build_layers(layer, len(x), lv, p)
    gt, _ = dev.knn(qs, 10)
    found = [hnsw.search(graphs, ep, xr, q, p.search_L, 10) for q in qs]
    rec = float(np.mean([len(set(f) & set(g.tolist())) / 10 for f, g in zip(found, gt)]))
    out = {"workload": "C1: fasthnsw, 100,000x128 integer mixture of 64 Gaussians (seed 0), K=L=200 S=10 R=100 "
                       "I=6 nsg_R=40 (level 0: 80) search_L=40 search_K=200",
           "build_sec": sec, "levels": len(graphs), "query_recall_at_10": rec, "queries": len(qs)}
    log(f"c1: {out}")
    return out


def c3_leg(dev):
    """C3: tau-MNG over 1,000,000 x 960 (256-centre float mixture, seed 2),
    K = L = 500 S = 12 R = 100 I = 6 nsg_R = 50 tau = 10: the plain build
    (kernels already loaded by the C2 builds) and the build with the
    Gram-form join on MFMA (BF16 operands, FP32 recheck), their graphs
    compared."""
    from fastkcna_amd import FK_PG_TAUMNG, _ffi
    from fastkcna_amd.datasets import c3, params_c3
    from fastkcna_amd.kgraph import IndexParams
    t = time.perf_counter()
    x, qs = c3(n_queries=1_000)
    log(f"c3: data {x.shape} in {time.perf_counter() - t:.1f}s")
    dev.upload(x)
    p = params_c3(IndexParams)
    p.flags = _ffi.FK_FLAG_PROFILE
    t = time.perf_counter()
    off, ids, ep, _, st = dev.build(FK_PG_TAUMNG, p)
    plain = time.perf_counter() - t
    log(f"c3: plain build {plain:.1f}s")
    gt, _ = dev.knn(qs, 10)
    rec = recall_at_10(dev.search(off, ids, ep, qs, 80, 10)[0], gt)
    q = p.copy(flags=p.flags | _ffi.FK_FLAG_GRAM)
    t = time.perf_counter()
    goff, gids, gep, _, gst = dev.build(FK_PG_TAUMNG, q)
    chosen = time.perf_counter() - t
    log(f"c3: Gram build {chosen:.1f}s")
    same = bool(np.array_equal(off, goff) and np.array_equal(ids, gids) and ep == gep)
    fam = lambda s: {f: round(s["prof_ms"][i], 1) for i, f in enumerate(_ffi.PROF_FAMILIES) if s["prof_launches"][i]}
    out = {"workload": "C3: tau-MNG, 1,000,000x960 float32 mixture of 256 Gaussians (seed 2, sigma 0.05), K=L=500 "
                       "S=12 R=100 I=6 nsg_R=50 tau=10 search_L=80 search_K=500",
           "build_sec": plain, "dist_evals_per_sec": total_dists(st) / plain,
           "dist_evals_computed_per_sec": total_dists(st, computed=True) / plain,
           "d_prune": int(st["d_prune"]), "d_prune_rev": int(st["d_prune_rev"]),
           "d_prune_uncomputed": int(st["d_prune_uncomputed"]),
           "prune_variant": int(st["prune_variant"]) - 1, "prune_rev_variant": int(st["prune_rev_variant"]) - 1,
           "tuning_sec": float(st["t_tune"]),
           "knn_recall_at_maxl": knn_recall_entries(st), "query_recall_at_10": rec,
           "families_ms": fam(st),
           "phase_sec": {k: round(float(st[k]), 4) for k in ("t_ctrl", "t_nnd", "t_prune", "t_inter", "t_ep", "t_conn",
                                                            "t_bv", "t_total")},
           "gram": {"build_sec": chosen, "speedup_vs_plain": plain / chosen,
                    "gram_mode": _ffi.GRAM_MODE_NAMES[int(gst["gram_mode"])], "graph_identical_to_plain": same,
                    "gram_pairs": int(gst["d_gram"]), "fp32_rechecks": int(gst["d_recheck"]),
                    "families_ms": fam(gst)}}
    log(f"c3: {out}")
    return out


def c5_leg(dev):
    """C5: the HNSW level-0 layer (nsg_R 64) of 10,000,000 x 96 unit-normalised
    rows (4,096 centres, sigma 0.25, seed 4), K = L = 200 S = 10 R = 100 I = 6
    search_K = 200, built plain, with FK_FLAG_GRAM (Gram-form prune) and with
    the learned screen."""
    from fastkcna_amd import FK_PG_HNSW_LAYER, _ffi
    from fastkcna_amd.datasets import c5, params_c5
    from fastkcna_amd.kgraph import IndexParams
    t = time.perf_counter()
    x, qs = c5(n_queries=1_000)
    log(f"c5: data {x.shape} in {time.perf_counter() - t:.1f}s")
    dev.upload(x)
    p = params_c5(IndexParams)
    t = time.perf_counter()
    off, ids, ep, _, st = dev.build(FK_PG_HNSW_LAYER, p.copy(flags=_ffi.FK_FLAG_PROFILE))
    plain = time.perf_counter() - t
    gt, _ = dev.knn(qs, 10)
    rec = recall_at_10(dev.search(off, ids, ep, qs, 40, 10)[0], gt)
    t = time.perf_counter()
    goff, gids, gep, _, gst = dev.build(FK_PG_HNSW_LAYER, p.copy(flags=_ffi.FK_FLAG_GRAM | _ffi.FK_FLAG_PROFILE))
    gram = time.perf_counter() - t
    gsame = bool(np.array_equal(off, goff) and np.array_equal(ids, gids) and ep == gep)
    del off, ids, goff, gids
    t = time.perf_counter()
    _, _, _, _, sst = dev.build(FK_PG_HNSW_LAYER, p.copy(flags=_ffi.FK_FLAG_SCREEN))
    scr = time.perf_counter() - t
    out = {"workload": "C5: HNSW level 0, 10,000,000x96 unit-normalised mixture of 4,096 Gaussians (seed 4, "
                       "sigma 0.25), K=L=200 S=10 R=100 I=6 nsg_R=64 search_L=40 search_K=200",
           "build_sec": plain, "dist_evals_per_sec": total_dists(st) / plain,
           "dist_evals_computed_per_sec": total_dists(st, computed=True) / plain,
           "knn_recall_at_maxl": knn_recall_entries(st), "query_recall_at_10_flat": rec,
           "families_ms": {f: round(st["prof_ms"][i], 1) for i, f in enumerate(_ffi.PROF_FAMILIES)
                           if st["prof_launches"][i]},
           "phase_sec": {k: round(float(st[k]), 4) for k in ("t_ctrl", "t_nnd", "t_prune", "t_inter", "t_ep", "t_conn",
                                                            "t_bv", "t_total")},
           "prune_variant": int(st["prune_variant"]) - 1, "prune_rev_variant": int(st["prune_rev_variant"]) - 1,
           "tuning_sec": float(st["t_tune"]),
           "gram": {"build_sec": gram, "speedup_vs_plain": plain / gram, "graph_identical_to_plain": gsame,
                    "gram_mode": _ffi.GRAM_MODE_NAMES[int(gst["gram_mode"])],
                    "prune_variant": int(gst["prune_variant"]) - 1,
                    "prune_rev_variant": int(gst["prune_rev_variant"]) - 1, "tuning_sec": float(gst["t_tune"]),
                    "d_prune_uncomputed": int(gst["d_prune_uncomputed"]),
                    "families_ms": {f: round(gst["prof_ms"][i], 1) for i, f in enumerate(_ffi.PROF_FAMILIES)
                                    if gst["prof_launches"][i]}},
           "learned_screen": {"build_sec": scr, "speedup_vs_plain": plain / scr,
                              "knn_recall_at_maxl": knn_recall_entries(sst),
                              "screened_pairs": int(sst["d_screened"]), "skipped_pairs": int(sst["d_screen_skip"]),
                              "est_lost_candidate_frac": 64.0 * sst["screen_missed"] / max(sst["screen_cands"], 1),
                              "train_loss": float(sst["screen_loss"]), "last_margin": float(sst["screen_margin"])}}
    log(f"c5: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--queries", type=int, default=10_000)
    ap.add_argument("--cpu-sample", type=int, default=100_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-full", action="store_true",
                    help="also build the full workload with the CPU oracle (minutes) and compare recall")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, help="MiB of a device-to-device copy to run first")
    ap.add_argument("--no-chosen", action="store_true", help="skip the chosen (MFMA Gram join) leg")
    ap.add_argument("--no-screen", action="store_true", help="skip the learned-screen leg")
    ap.add_argument("--batch-sets", type=int, default=256,
                    help="datasets of the batching workload timed batched vs one by one (0: skip)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other workloads (C1 fasthnsw, C3 tau-MNG 1M x 960, C5 10M x 96)")
    ap.add_argument("--leg", choices=["C1", "C3", "C4", "C5"],
                    help="run only this workload's leg and print its JSON (profiling runs)")
    ap.add_argument("--batch-only", action="store_true",
                    help="with --leg C4: only the batched builds, not the one-by-one comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the build does not shard "
              "across GPUs in this version)", file=sys.stderr)
        sys.exit(2)

    from fastkcna_amd import FK_PG_NSG, Device, _ffi
    from fastkcna_amd.datasets import integer_mixture, params_c2
    from fastkcna_amd.kgraph import IndexParams

    if args.leg:
        dev = Device(0)
        legs = {"C1": c1_leg, "C3": c3_leg, "C5": c5_leg,
                "C4": lambda d: batching_leg(d, args.batch_sets or 256, one_by_one=not args.batch_only)}
        print(json.dumps(legs[args.leg](dev)), flush=True)
        dev.close()
        return

    t = time.perf_counter()
    x = integer_mixture(args.n + args.queries, 128, 1024, seed=1)
    base, queries = x[:args.n], x[args.n:]
    log(f"data {base.shape} + {queries.shape} in {time.perf_counter() - t:.1f}s")
    p = params_c2(IndexParams)
    p.flags = _ffi.FK_FLAG_PROFILE | (_ffi.FK_FLAG_VERBOSE if args.verbose else 0)

    dev = Device(0)
    if args.copy_probe:
        copy_probe(args.copy_probe)
    t = time.perf_counter()
    dev.upload(base)
    upload_sec = time.perf_counter() - t
    log(f"upload {upload_sec:.2f}s")

    for w in range(args.warmup):
        t = time.perf_counter()
        dev.build(FK_PG_NSG, p)
        log(f"warmup {w} {time.perf_counter() - t:.2f}s")

    times, stats, graph = [], [], None
    for s in range(args.steps):
        t = time.perf_counter()
        off, ids, ep, width, st = dev.build(FK_PG_NSG, p)
        dt = time.perf_counter() - t
        times.append(dt)
        stats.append(st)
        graph = (off, ids, ep)
        log(f"step {s}: build {dt:.3f}s, dist-evals {total_dists(st):,}, aod {st['aod']}")

    off, ids, ep = graph
    digest = graph_digest(off, ids, ep)
    log(f"graph digest {digest}")
    t = time.perf_counter()
    gt, gt_d = dev.knn(queries, 10)
    res, _ = dev.search(off, ids, ep, queries, 80, 10)
    rec = recall_at_10(res, gt)
    rec_ties = recall_at_10_ties(base, queries, res, gt_d)
    log(f"recall@10 (search_L 80) {rec:.4f} (tie-aware {rec_ties:.4f}); evaluation {time.perf_counter() - t:.1f}s")

    chosen = screen = None
    if not args.no_chosen:
        chosen = chosen_leg(dev, p, args, graph, gt, queries, float(np.mean(times)))
    if not args.no_screen:
        screen = screen_leg(dev, p, args, graph, gt, queries, float(np.mean(times)))

    st = stats[-1]
    ms = 1000.0 * float(np.mean(times))
    nd = total_dists(st)
    value = nd / (ms / 1000.0)
    roof = roofline(st, _ffi.PROF_FAMILIES, 128)
    log("kernel families (ms, GB/s): " + ", ".join(
        f"{f}={st['prof_ms'][i]:.1f}/{st['prof_bytes'][i] / max(st['prof_ms'][i], 1e-9) / 1e6:.0f}"
        for i, f in enumerate(_ffi.PROF_FAMILIES) if st['prof_launches'][i]))
    batching = batching_leg(dev, args.batch_sets) if args.batch_sets else None
    configs = None
    if not args.no_configs:
        configs = {"C1": c1_leg(dev), "C3": c3_leg(dev), "C5": c5_leg(dev)}
    cpu = cpu_full = None
    threads = cpu_threads()
    if args.cpu_full:
        cpu_full = cpu_full_leg(dev, base, queries, graph, st, threads)  # uploads the workload's rows again
        log(f"cpu full: {cpu_full}")
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_sample, threads, dev, queries[:2000])
        log(f"cpu baseline {cpu['value']:.3e} dist-evals/s on {threads} threads ({cpu['build_sec']:.1f}s); "
            f"sample parity {cpu['sample_parity']}")
        cpu.pop("build_sec")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "dist-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"C2: NSG build, {args.n:,}x128 float32 integer-valued mixture of 1,024 "
                               "Gaussians (seed 1), K=L=500 S=12 R=100 I=6 nsg_R=50 search_L=80 search_K=500",
                   "build_sec": ms / 1000.0, "upload_sec": upload_sec,
                   "build_sec_with_upload": ms / 1000.0 + upload_sec,
                   "recall_at_10": rec, "recall_at_10_tie_aware": rec_ties, "search_L": 80, "queries": args.queries,
                   "dist_evals": nd, "dist_evals_computed": total_dists(st, computed=True),
                   "dist_evals_computed_per_sec": total_dists(st, computed=True) / (ms / 1000.0),
                   "d_bv": int(st["d_bv"]), "d_bv_computed": int(st["d_bv_computed"]),
                   "aod": int(st["aod"]), "nnd_iterations": int(st["nnd_iterations"]),
                   "prune_variant": int(st["prune_variant"]) - 1, "prune_rev_variant": int(st["prune_rev_variant"]) - 1,
                   "graph_digest": digest,
                   "phase_sec": {k: round(float(st[k]), 4) for k in ("t_ctrl", "t_nnd", "t_prune", "t_inter",
                                                                    "t_ep", "t_conn", "t_bv", "t_total")}},
        "roofline": roof,
        "chosen": chosen,
        "learned_screen": screen,
        "batching": batching,
        "configs": configs,
        "cpu_baseline": cpu,
        "cpu_full": cpu_full,
    }
    print(json.dumps(out), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
