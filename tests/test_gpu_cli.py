"""The command-line tools end to end on the GPU: files in the reference's
formats, graphs that search well.

fastnsg / fasttaumng / fastalphapg: the NSG index (u32 width, u32 ep, per
node u32 k + ids, src/kgraph.cpp:1704-1724) and the CSV rows of SaveResult.
fasthnsw: the hnswlib index (src/fasthnsw.cpp:317-342) -- header fields,
level-0 records (links, vector, label), upper-level link lists -- searched
with hnswlib's greedy descent + level-0 beam, written here in NumPy.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle
from conftest import ROOT
from fastkcna_amd.datasets import integer_mixture
from fastkcna_amd.formats import load_nsg, write_lshkit

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "fastkcna_amd", "bin")


def run(tool, *args):
    print("[case]", tool, *args, flush=True)
    out = subprocess.run([os.path.join(BIN, tool), *map(str, args)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    return out


def beam_search(off, ids, ep, x, q, L, K):
    """NSG-style best-first search over a CSR graph (NumPy, exact distances)."""
    visited = {ep}
    d = lambda v: float(np.sum((x[v] - q) ** 2))
    pool = [(d(ep), ep, False)]
    while True:
        pool.sort()
        pool = pool[:L]
        i = next((k for k, e in enumerate(pool) if not e[2]), None)
        if i is None:
            break
        dist, u, _ = pool[i]
        pool[i] = (dist, u, True)
        for v in ids[off[u]:off[u + 1]]:
            v = int(v)
            if v not in visited:
                visited.add(v)
                pool.append((d(v), v, False))
    return [e[1] for e in sorted(pool)[:K]]


@pytest.mark.parametrize("tool,extra", [("fastnsg", []), ("fasttaumng", ["--tau", "5"]),
                                        ("fastalphapg", ["--angle", "60"])])
def test_flat_tools_write_nsg_and_csv(tmp_path, tool, extra):
    x = integer_mixture(3000, 32, 16, seed=21, sigma=40.0)
    data = tmp_path / "b.lshkit"
    write_lshkit(data, x)
    out, res = tmp_path / "g.nsg", tmp_path / "r.csv"
    run(tool, data, out, "-K", 40, "-L", 40, "-S", 10, "-R", 50, "-I", 5, "--nsg_R", 16, "--search_L", 40,
        "--search_K", 40, "--result", res, *extra)
    width, ep, off, ids = load_nsg(out)
    assert len(off) == 3001 and ep < 3000 and np.diff(off).max() <= width
    lines = res.read_text().splitlines()
    assert lines[0].startswith("PG_type,K,L,S,R,I")
    assert lines[1].split(",")[0] == {"fastnsg": "1", "fasttaumng": "3", "fastalphapg": "4"}[tool]
    assert [ln.split(",")[0] for ln in lines[2:]] == ["kgraph time", "recall", "prune time", "inter time",
                                                     "tree_grow time", "buildnsg time", "search time", "con time",
                                                     "prune cc", "search cc", "AOD", "final_prune_time"]
    rng = np.random.default_rng(1)
    qs = x[rng.choice(3000, 50, replace=False)] + rng.normal(0, 2, (50, 32)).astype(np.float32)
    gt, _ = oracle.knn(x, qs, 10)
    rec = np.mean([len(set(beam_search(off, ids, ep, x, qs[i], 40, 10)) & set(gt[i])) / 10 for i in range(50)])
    assert rec > 0.8, rec


def read_hnsw(path, d):
    raw = open(path, "rb").read()
    f = struct.unpack_from("<QQQQQQiIQQQdQ", raw, 0)
    hdr = dict(zip(["offsetLevel0", "max_elements", "cur_count", "size_per_element", "label_offset", "offsetData",
                    "maxlevel", "enterpoint", "maxM", "maxM0", "M", "mult", "ef_construction"], f))
    pos = struct.calcsize("<QQQQQQiIQQQdQ")
    n, spe = hdr["cur_count"], hdr["size_per_element"]
    lvl0 = raw[pos:pos + n * spe]
    pos += n * spe
    links = []
    for i in range(n):
        (sz,) = struct.unpack_from("<I", raw, pos)
        pos += 4
        links.append(raw[pos:pos + sz])
        pos += sz
    assert pos == len(raw)
    return hdr, lvl0, links


def test_fasthnsw_writes_hnswlib_index(tmp_path):
    n, d, R = 4000, 32, 12
    x = integer_mixture(n, d, 16, seed=22, sigma=40.0)
    data = tmp_path / "b.lshkit"
    write_lshkit(data, x)
    out, res = tmp_path / "g.hnsw", tmp_path / "r.csv"
    run("fasthnsw", data, out, "-K", 40, "-L", 40, "-S", 10, "-R", 50, "-I", 5, "--nsg_R", R, "--search_L", 40,
        "--search_K", 40, "--result", res)
    hdr, lvl0, links = read_hnsw(out, d)
    assert hdr["cur_count"] == n and hdr["maxM"] == R and hdr["maxM0"] == 2 * R and hdr["M"] == R
    assert hdr["size_per_element"] == 4 + 8 * R + 4 * d + 8
    assert hdr["offsetData"] == 4 + 8 * R and hdr["label_offset"] == 4 + 8 * R + 4 * d
    assert abs(hdr["mult"] - 1 / np.log(R)) < 1e-12 and hdr["ef_construction"] == 40
    spe = hdr["size_per_element"]
    vec = np.zeros((n, d), np.float32)
    lab = np.zeros(n, np.int64)
    nb0 = []
    for i in range(n):
        rec = lvl0[i * spe:(i + 1) * spe]
        cnt = struct.unpack_from("<H", rec, 0)[0]
        assert cnt <= 2 * R
        nb0.append(np.frombuffer(rec, np.uint32, cnt, 4))
        vec[i] = np.frombuffer(rec, np.float32, d, hdr["offsetData"])
        lab[i] = struct.unpack_from("<Q", rec, hdr["label_offset"])[0]
    assert sorted(lab.tolist()) == list(range(n))
    assert np.array_equal(vec, x[lab])  # internal order = rows reordered by level
    assert all(np.all(v < n) for v in nb0)
    # hnswlib search: greedy descent through the upper levels, beam at level 0
    per = 4 + 4 * R

    def upper(i, level):
        blk = links[i][(level - 1) * per:level * per]
        if not blk:
            return np.zeros(0, np.uint32)
        c = struct.unpack_from("<H", blk, 0)[0]
        return np.frombuffer(blk, np.uint32, c, 4)

    rng = np.random.default_rng(2)
    qs = x[rng.choice(n, 50, replace=False)] + rng.normal(0, 2, (50, d)).astype(np.float32)
    gt, _ = oracle.knn(x, qs, 10)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(v) for v in nb0])
    flat = np.concatenate(nb0)
    hits = 0
    for qi, q in enumerate(qs):
        cur = hdr["enterpoint"]
        for level in range(hdr["maxlevel"], 0, -1):
            changed = True
            while changed:
                changed = False
                for v in upper(cur, level):
                    if np.sum((vec[v] - q) ** 2) < np.sum((vec[cur] - q) ** 2):
                        cur, changed = int(v), True
        found = beam_search(off, flat, cur, vec, q, 40, 10)
        hits += len(set(lab[found].tolist()) & set(gt[qi].tolist()))
    assert hits / 500 > 0.8, hits / 500
    assert "fasthnsw indexing time" in res.read_text()


def test_fastnsg_gram_flag_writes_the_same_index(tmp_path):
    """--gram (local join on the MFMA units, exact form on integer rows) must
    write the byte-identical NSG index of the plain build."""
    x = integer_mixture(3000, 128, 16, seed=23, sigma=40.0)
    data = tmp_path / "b.lshkit"
    write_lshkit(data, x)
    outs = []
    for extra in ([], ["--gram"]):
        out = tmp_path / f"g{len(outs)}.nsg"
        run("fastnsg", data, out, "-K", 40, "-L", 40, "-S", 10, "-R", 50, "-I", 5, "--nsg_R", 16, "--search_L",
            40, "--search_K", 40, *extra)
        outs.append(out.read_bytes())
    assert outs[0] == outs[1]
