"""C-ABI library checks that need no GPU: it loads, exports every function
include/fk_api.h declares, and its structures have the layout the bindings
assume."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HDR = os.path.join(ROOT, "include", "fk_api.h")
LIB = os.path.join(ROOT, "fastkcna_amd", "libfastkcna.so")


def declared_functions():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(fk_[a-z_0-9]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["fk_open", "fk_close", "fk_build", "fk_prune", "fk_join_distances", "fk_nndescent",
                 "fk_pair_distances", "fk_knn_bruteforce", "fk_search", "fk_dataset_upload"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libfastkcna.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"not exported: {missing}"


def test_library_loads_and_bindings_match():
    from fastkcna_amd import _ffi
    lib = _ffi.load()
    for name in declared_functions():
        assert hasattr(lib, name)
    assert set(_ffi.exported_symbols()) == set(declared_functions())
    assert lib.fk_version().decode().startswith("fastkcna-amd")


def test_struct_layout_matches_c(tmp_path):
    from fastkcna_amd._ffi import FkGraph, FkParams, FkStats
    prog = tmp_path / "sz.c"
    prog.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "fk_api.h"\n'
                    'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(fk_params), sizeof(fk_stats),'
                    ' sizeof(fk_graph), offsetof(fk_stats, kgraph_time), offsetof(fk_graph, stats));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert vals == [ctypes.sizeof(FkParams), ctypes.sizeof(FkStats), ctypes.sizeof(FkGraph),
                    FkStats.kgraph_time.offset, FkGraph.stats.offset]


def test_open_without_gpu_reports_nodevice():
    """On a host with no gfx950 device fk_open must fail cleanly."""
    from fastkcna_amd import _ffi
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is present")
    except Exception:
        pass
    lib = _ffi.load()
    h = ctypes.c_void_p()
    assert lib.fk_open(0, ctypes.byref(h)) == _ffi.FK_E_NODEVICE
    assert not h.value


def test_null_handle_is_rejected():
    from fastkcna_amd import _ffi
    lib = _ffi.load()
    assert lib.fk_dataset_upload(None, None, 0, 0, 0) == _ffi.FK_E_INVALID
    assert lib.fk_build(None, 1, 0, None, None) == _ffi.FK_E_INVALID
    assert lib.fk_last_error(None) == b"null handle"
