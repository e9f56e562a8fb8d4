"""Python mirror of the reference's KGraph build interface.

Names, defaults and argument meaning follow src/kgraph.h (IndexParams
:116-170 with the defaults of :12-40; KGraph::buildFastNSG / buildFastTauMNG /
buildFastAlphaPG :230-232; IndexOracle :63-81; MatrixOracle
src/kgraph-data.h:349-433).  The build itself runs on the GPU through the C
ABI; the index and CSV writers keep the reference's byte formats
(src/kgraph.cpp:1704-1807).
"""
import numpy as np

from . import _ffi
from .engine import Device
from .formats import save_nsg, save_result_csv

# src/kgraph.h:12-40
default_iterations = 30
default_L = 100
default_K = 25
default_S = 10
default_R = 100
default_controls = 100
default_seed = 2024
default_delta = 0.002
default_recall = 0.95
default_loop_i = 2
default_nsgr = 50
default_searchl = 40
default_searchk = 500
default_step = 10
default_massqS = 10
default_tau = 0.0
default_angle = 60.0


class invalid_argument(ValueError):
    """kgraph::invalid_argument (src/kgraph.h:363-366)."""


class runtime_error(RuntimeError):
    """kgraph::runtime_error (src/kgraph.h:368-371)."""


class io_error(runtime_error):
    """kgraph::io_error (src/kgraph.h:373-376)."""


class IndexParams:
    """KGraph::IndexParams with the CLI defaults (src/fastnsg.cpp:53-89)."""

    def __init__(self, **kw):
        self.K = default_K
        self.L = default_L
        self.S = default_S
        self.R = default_R
        self.iterations = default_iterations
        self.controls = default_controls
        self.seed = default_seed
        self.delta = default_delta
        self.recall = default_recall
        self.nsg_R = default_nsgr
        self.search_L = default_searchl
        self.search_K = default_searchk
        self.step = default_step
        self.massq_S = default_massqS
        self.loop_i = default_loop_i
        self.tau = default_tau
        self.angle = default_angle
        self.flags = 0
        for k, v in kw.items():
            if not hasattr(self, k):
                raise invalid_argument(f"unknown IndexParams field {k}")
            setattr(self, k, v)
        if self.S == 0:  # src/fastnsg.cpp:125-127
            self.S = self.K

    def copy(self, **kw):
        p = IndexParams()
        p.__dict__.update(self.__dict__)
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def __repr__(self):
        return "IndexParams(" + ", ".join(f"{k}={v}" for k, v in self.__dict__.items()) + ")"


class MatrixOracle:
    """Oracle over a float32 matrix: size() is the prefix set by set_size()
    (HNSW layers), data_size() the full matrix."""

    def __init__(self, data):
        self.data = np.ascontiguousarray(data, dtype=np.float32)
        if self.data.ndim != 2:
            raise invalid_argument("bad array shape")
        self.oracle_size = self.data.shape[0]

    def size(self):
        return self.oracle_size

    def data_size(self):
        return self.data.shape[0]

    def dim(self):
        return self.data.shape[1]

    def set_size(self, n):
        self.oracle_size = min(n, self.data.shape[0])

    def __call__(self, i, j):
        a = self.data[i].astype(np.float64)
        b = self.data[j].astype(np.float64)
        return float(np.sum((a - b) ** 2))


class BuildResult:
    def __init__(self, offsets, ids, ep, width, stats):
        self.offsets, self.ids, self.ep, self.width, self.stats = offsets, ids, ep, width, stats

    def neighbors(self, i):
        return self.ids[self.offsets[i]:self.offsets[i + 1]]


class KGraph:
    """KGraph::create() analogue; one instance owns one device handle."""

    def __init__(self, device=0):
        self._dev = Device(device)
        self._uploaded = None
        self.last = None

    @staticmethod
    def create(device=0):
        return KGraph(device)

    @staticmethod
    def version():
        return _ffi.load().fk_version().decode()

    def _ensure(self, oracle):
        key = (id(oracle.data), oracle.data.shape)
        if self._uploaded != key:
            self._dev.upload(oracle.data)
            self._uploaded = key

    def _build(self, pg_type, oracle, params, nsg_name, res_name):
        self._ensure(oracle)
        off, ids, ep, width, st = self._dev.build(pg_type, params, n_used=oracle.size())
        self.last = BuildResult(off, ids, ep, width, st)
        if nsg_name:
            save_nsg(nsg_name, width, ep, off, ids)
        if res_name:
            save_result_csv(res_name, pg_type, params, st)
        return self.last

    def buildFastNSG(self, oracle, params, info=None, nsg_name=None, res_name=None):
        return self._build(_ffi.FK_PG_NSG, oracle, params, nsg_name, res_name)

    def buildFastTauMNG(self, oracle, params, info=None, nsg_name=None, res_name=None):
        return self._build(_ffi.FK_PG_TAUMNG, oracle, params, nsg_name, res_name)

    def buildFastAlphaPG(self, oracle, params, info=None, nsg_name=None, res_name=None):
        return self._build(_ffi.FK_PG_ALPHAPG, oracle, params, nsg_name, res_name)

    def buildHNSWLayer(self, oracle, params, res_name=None):
        """One FastHNSW layer build over oracle.size() rows (no connectivity
        pass, src/kgraph.cpp:2326-2337)."""
        return self._build(_ffi.FK_PG_HNSW_LAYER, oracle, params, None, res_name)

    def close(self):
        self._dev.close()
