"""C++ host layer checks that need no GPU: the command-line tools parse the
reference's flags, and the HNSW level assignment of kgraph-data.h reproduces
the reference's libstdc++ random stream (src/kgraph-data.h:179-222)."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "fastkcna_amd", "bin")


def tool(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.skip(f"{name} not built")
    return p


@pytest.mark.parametrize("name", ["fastnsg", "fasttaumng", "fastalphapg", "fasthnsw"])
def test_help_and_version(name):
    out = subprocess.run([tool(name), "--help"], capture_output=True, text=True, timeout=30)
    assert out.returncode == 0 and "Usage: index" in out.stdout
    for flag in ["--nsg_R", "--search_L", "--search_K", "--massq_S", "--loop_i", "--tau", "--angle", "-K", "-L",
                 "-S", "-R", "--iterations"]:
        assert flag in out.stdout
    out = subprocess.run([tool(name), "--version"], capture_output=True, text=True, timeout=30)
    assert out.returncode == 0 and "fastkcna-amd" in out.stdout


def test_unknown_flag_is_an_error():
    out = subprocess.run([tool("fastnsg"), "--bogus", "1", "x.lshkit"], capture_output=True, text=True, timeout=30)
    assert out.returncode == 1 and "unrecognised" in out.stderr


def test_missing_data_file_is_io_error(tmp_path):
    out = subprocess.run([tool("fastnsg"), str(tmp_path / "nope.lshkit"), str(tmp_path / "o.nsg")],
                         capture_output=True, text=True, timeout=30)
    assert out.returncode == 2 and "nope.lshkit" in out.stderr


def minstd_levels(n, M, seed):
    """std::default_random_engine (minstd_rand0) + uniform_real_distribution<
    double>(0,1) as libstdc++ evaluates them (generate_canonical with two
    draws), then (int)(-log(u) * 1/log(M))."""
    a, m = 16807, 2147483647
    state = seed % m
    if state == 0:
        state = 1
    R = m - 1  # max - min + 1
    mult = 1 / math.log(1.0 * M)
    out = []
    for _ in range(n):
        s = 0.0
        tmp = 1.0
        for _k in range(2):
            state = (a * state) % m
            s += (state - 1) * tmp
            tmp *= R
        u = s / tmp
        if u >= 1.0:
            u = math.nextafter(1.0, 0.0)
        out.append(int(-math.log(u) * mult))
    return out


def test_hnsw_levels_match_libstdcxx_stream(tmp_path):
    src = tmp_path / "lv.cpp"
    src.write_text('#include "kgraph-data.h"\n#include <cstdio>\nnamespace kgraph { unsigned verbosity = 0; }\n'
                   'int main(){ kgraph::Matrix<float> m; std::vector<int> d2f, f2d, pn;\n'
                   ' m.getlevel(5000, d2f, f2d, pn, 40, 2024);\n'
                   ' printf("%zu\\n", pn.size()); for (int v : pn) printf("%d\\n", v);\n'
                   ' for (int i = 0; i < 50; ++i) printf("%d\\n", d2f[i]); return 0; }\n')
    exe = tmp_path / "lv"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    nl = vals[0]
    pn = vals[1:1 + nl]
    d2f = vals[1 + nl:]
    lv = minstd_levels(5000, 40, 2024)
    assert pn == [lv.count(k) for k in range(max(lv) + 1)]
    expect = sorted(range(5000), key=lambda i: (-lv[i], i))[:50]
    assert d2f == expect
