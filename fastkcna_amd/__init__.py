"""MI355X-native FastKCNA graph construction (NN-descent + NSG / HNSW-layer /
tau-MNG / alpha-PG pruners) behind the reference's KGraph build interface.

The compute path is libfastkcna.so (gfx950 HIP kernels behind the C ABI in
include/fk_api.h); this package binds it with ctypes and mirrors the
reference's Python-visible surface for tests and benchmarks.
"""
from ._ffi import (FK_PG_ALPHAPG, FK_PG_HNSW_LAYER, FK_PG_NSG, FK_PG_TAUMNG, FkError,  # noqa: F401
                   LIB_PATH, load)
from .engine import Device  # noqa: F401
from .kgraph import IndexParams, KGraph, MatrixOracle  # noqa: F401

__all__ = [
    "Device", "FkError", "IndexParams", "KGraph", "MatrixOracle", "LIB_PATH", "load",
    "FK_PG_NSG", "FK_PG_HNSW_LAYER", "FK_PG_TAUMNG", "FK_PG_ALPHAPG",
]
