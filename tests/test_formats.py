"""Byte layouts of the reference's files (src/kgraph.cpp:1704-1807,
src/kgraph-data.h:262-287)."""
import struct

import numpy as np

from fastkcna_amd import IndexParams
from fastkcna_amd.formats import load_nsg, read_lshkit, save_nsg, save_result_csv, write_lshkit


def test_lshkit_round_trip(tmp_path, rng):
    x = rng.normal(size=(17, 5)).astype(np.float32)
    p = tmp_path / "a.lshkit"
    write_lshkit(p, x)
    raw = p.read_bytes()
    assert struct.unpack("<III", raw[:12]) == (4, 17, 5)
    assert np.array_equal(read_lshkit(p), x)


def test_nsg_layout(tmp_path):
    off = np.array([0, 2, 2, 5], np.uint64)
    ids = np.array([1, 2, 0, 1, 2], np.uint32)
    p = tmp_path / "g.nsg"
    save_nsg(p, 7, 1, off, ids)
    raw = np.fromfile(p, dtype=np.uint32)
    assert list(raw) == [7, 1, 2, 1, 2, 0, 3, 0, 1, 2]
    w, ep, o2, i2 = load_nsg(p)
    assert (w, ep) == (7, 1) and np.array_equal(o2, off) and np.array_equal(i2, ids)


def test_csv_rows(tmp_path):
    st = {"n_loops": 3, "n_kgraph": 2, "n_recall": 2, "kgraph_time": [1.5, 2.0], "recall": [0.9, 0.95],
          "prune_time": [0.1] * 3, "inter_time": [0.2] * 3, "tree_grow_time": [0.3] * 3,
          "buildnsg_time": [0.4] * 3, "search_time": [0.5] * 2, "con_time": [1, 2, 3],
          "prune_cc": [1e-4] * 3, "search_cc": [2e-4] * 2, "aod": 31}
    p = tmp_path / "r.csv"
    save_result_csv(p, 1, IndexParams(K=500, L=500), st)
    lines = p.read_text().splitlines()
    assert lines[0] == "PG_type,K,L,S,R,I,searchL,searchK,nsgR,massqS,tau,angle,step"
    assert lines[1].startswith("1,500,500,10,100,30,40,500,50,10,0,60,10,")
    assert lines[2] == "kgraph time,1.5,2,"
    assert lines[-2] == "AOD,31" and lines[-1] == "final_prune_time,-1"
