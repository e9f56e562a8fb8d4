"""FastHNSW's layered build over the C ABI and an hnswlib-style search of its
result, in NumPy.

levels(): the level of every point as src/kgraph-data.h:179-222 draws it --
std::default_random_engine (minstd_rand0) seeded with `seed`, one
uniform_real_distribution<double>(0, 1) draw per point in point order (libstdc++
generate_canonical with two draws), level = (int)(-log(u) / log(nsg_R)).
build_layers(): the per-level prefix builds of src/fasthnsw.cpp:136-342 (rows
ordered by level, highest first; each level built over the prefix of points
at or above it, level 0 with nsg_R doubled).  search(): greedy descent
through the upper levels, then a best-first beam of width ef at level 0.
"""
import math
import time

import numpy as np

from ._ffi import FK_PG_HNSW_LAYER


def levels(n, M, seed):
    a, m = 16807, 2147483647
    state = seed % m or 1
    R = m - 1
    mult = 1 / math.log(1.0 * M)
    out = np.empty(n, np.int64)
    for i in range(n):
        s, tmp = 0.0, 1.0
        for _ in range(2):
            state = (a * state) % m
            s += (state - 1) * tmp
            tmp *= R
        u = s / tmp
        if u >= 1.0:
            u = math.nextafter(1.0, 0.0)
        out[i] = int(-math.log(u) * mult)
    return out


def level_order(lv):
    """Internal order of the rows: highest level first, then file order."""
    return np.lexsort((np.arange(len(lv)), -lv))


def build_layers(build_layer, n, lv, params):
    """Builds every level with build_layer(params, n_used) -> (offsets, ids,
    ep) over the level-ordered rows; returns ([(offsets, ids)] indexed by
    level, entry point of the top level, seconds)."""
    top = int(lv.max())
    counts = np.bincount(lv, minlength=top + 1)
    size, graphs, top_ep = 0, [None] * (top + 1), 0
    t0 = time.perf_counter()
    for level in range(top, -1, -1):
        size += int(counts[level])
        if size == 1:
            graphs[level] = (np.zeros(n + 1, np.int64), np.zeros(0, np.int64))
            continue
        p = params.copy(nsg_R=params.nsg_R * (2 if level == 0 else 1))
        off, ids, ep = build_layer(p, size)
        full = np.full(n + 1, int(off[-1]), np.int64)
        full[:size + 1] = off
        graphs[level] = (full, np.asarray(ids, np.int64))
        if level == top:
            top_ep = int(ep)
    return graphs, top_ep, time.perf_counter() - t0


def device_layer(dev):
    """build_layer over a Device holding the level-ordered rows."""
    def build(p, size):
        off, ids, ep, _, _ = dev.build(FK_PG_HNSW_LAYER, p, n_used=size)
        return off, ids, ep
    return build


def search(graphs, top_ep, vec, q, ef, K):
    cur = top_ep
    dc = float(np.sum((vec[cur] - q) ** 2))
    for off, ids in graphs[1:][::-1]:
        changed = True
        while changed:
            changed = False
            nb = ids[off[cur]:off[cur + 1]]
            if len(nb):
                dn = np.sum((vec[nb] - q) ** 2, axis=1)
                j = int(np.argmin(dn))
                if dn[j] < dc:
                    cur, dc, changed = int(nb[j]), float(dn[j]), True
    off, ids = graphs[0]
    visited = {cur}
    pool = [(dc, cur, False)]
    while True:
        pool.sort()
        pool = pool[:ef]
        i = next((k for k, e in enumerate(pool) if not e[2]), None)
        if i is None:
            break
        d0, u, _ = pool[i]
        pool[i] = (d0, u, True)
        nb = [int(v) for v in ids[off[u]:off[u + 1]] if int(v) not in visited]
        if nb:
            visited.update(nb)
            dn = np.sum((vec[nb] - q) ** 2, axis=1)
            pool.extend((float(dv), v, False) for dv, v in zip(dn, nb))
    return [e[1] for e in sorted(pool)[:K]]
