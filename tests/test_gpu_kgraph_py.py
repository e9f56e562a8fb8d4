"""The Python mirror of the reference's build interface (fastkcna_amd.kgraph:
KGraph.create().buildFast*, MatrixOracle with set_size, IndexParams defaults)
on the GPU: its NSG index is byte-identical to the fastnsg tool's on the same
rows and flags (src/kgraph.cpp:1704-1724), its CSV carries SaveResult's rows,
and an HNSW layer over an oracle prefix equals fk_build over that prefix."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from fastkcna_amd import FK_PG_HNSW_LAYER, Device, IndexParams, KGraph, MatrixOracle
from fastkcna_amd.datasets import integer_mixture
from fastkcna_amd.formats import write_lshkit

pytestmark = pytest.mark.gpu


def test_python_mirror_matches_the_tool(tmp_path):
    x = integer_mixture(3000, 64, 12, seed=71, sigma=30.0)
    data = tmp_path / "b.lshkit"
    write_lshkit(data, x)
    flags = ["-K", "40", "-L", "40", "-S", "10", "-R", "40", "-I", "5", "--nsg_R", "16", "--search_L", "40",
             "--search_K", "40"]
    tool_out = tmp_path / "tool.nsg"
    r = subprocess.run([os.path.join(ROOT, "fastkcna_amd", "bin", "fastnsg"), str(data), str(tool_out), *flags],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    g = KGraph.create()
    try:
        p = IndexParams(K=40, L=40, S=10, R=40, iterations=5, nsg_R=16, search_L=40, search_K=40)
        py_out, csv = tmp_path / "py.nsg", tmp_path / "py.csv"
        res = g.buildFastNSG(MatrixOracle(x), p, nsg_name=str(py_out), res_name=str(csv))
        assert py_out.read_bytes() == tool_out.read_bytes()
        assert res.ep < 3000 and res.width >= 16 and len(res.neighbors(0)) > 0
        lines = csv.read_text().splitlines()
        assert lines[0].startswith("PG_type,K,L,S,R,I") and lines[1].split(",")[0] == "1"
        # an HNSW layer over the first 1,000 rows (the oracle's size() prefix)
        o = MatrixOracle(x)
        o.set_size(1000)
        assert o.size() == 1000 and o.data_size() == 3000
        lay = g.buildHNSWLayer(o, p)
    finally:
        g.close()
    with Device(0) as dev:
        dev.upload(x)
        off, ids, ep, width, _ = dev.build(FK_PG_HNSW_LAYER, p, n_used=1000)
    assert np.array_equal(lay.offsets, off) and np.array_equal(lay.ids, ids) and lay.ep == ep
