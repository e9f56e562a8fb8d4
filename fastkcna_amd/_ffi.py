"""ctypes binding of libfastkcna.so (the C ABI declared in include/fk_api.h).

The structures mirror fk_api.h field for field; tests compare their sizes
with what a C compiler reports for the header.  Loading fails loudly when the
library has not been built: there is no CPU fallback in the product path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfastkcna.so")

FK_OK, FK_E_INVALID, FK_E_NOMEM, FK_E_DEVICE, FK_E_IO, FK_E_INTERNAL, FK_E_NODEVICE = range(7)
FK_PG_NSG, FK_PG_HNSW_LAYER, FK_PG_TAUMNG, FK_PG_ALPHAPG = 1, 2, 3, 4
FK_FLAG_VERBOSE = 0x1
FK_FLAG_NO_CONTROLS = 0x2
FK_FLAG_PROFILE = 0x4
FK_FLAG_GRAM = 0x8
FK_FLAG_GRAM_BF16 = 0x10
FK_FLAG_GRAM_F16 = 0x20
FK_FLAG_SCREEN = 0x40
FK_FLAG_SCREEN_NOTRAIN = 0x80
FK_FLAG_PRUNE_V0 = 0x100
FK_FLAG_PRUNE_V1 = 0x200
FK_FLAG_PRUNE_V2 = 0x400
FK_FLAG_PRUNE_V3 = 0x800
FK_GRAM_NONE, FK_GRAM_EXACT, FK_GRAM_BF16, FK_GRAM_F16 = range(4)
GRAM_MODE_NAMES = {FK_GRAM_NONE: "none", FK_GRAM_EXACT: "exact", FK_GRAM_BF16: "bf16", FK_GRAM_F16: "f16"}
FK_MAX_LOOPS = 16
FK_PROF_N = 16
PROF_FAMILIES = ["controls", "init", "join", "join_fill", "merge", "update", "prune", "prune_rev", "inter", "bv",
                 "ep", "conn", "order"]

STATUS_NAMES = {
    FK_OK: "FK_OK",
    FK_E_INVALID: "FK_E_INVALID",
    FK_E_NOMEM: "FK_E_NOMEM",
    FK_E_DEVICE: "FK_E_DEVICE",
    FK_E_IO: "FK_E_IO",
    FK_E_INTERNAL: "FK_E_INTERNAL",
    FK_E_NODEVICE: "FK_E_NODEVICE",
}

u32, u64, f32, f64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_double


class FkParams(ctypes.Structure):
    _fields_ = [
        ("K", u32), ("L", u32), ("S", u32), ("R", u32), ("iterations", u32), ("controls", u32), ("seed", u32),
        ("delta", f32), ("recall", f32),
        ("nsg_R", u32), ("search_L", u32), ("search_K", u32), ("step", u32), ("massq_S", u32), ("loop_i", u32),
        ("tau", f32), ("angle", f32),
        ("flags", u32),
    ]


class FkStats(ctypes.Structure):
    _fields_ = [
        ("d_init", u64), ("d_ctrl", u64), ("d_join", u64), ("d_prune", u64), ("d_prune_rev", u64),
        ("d_inter", u64), ("d_ep", u64), ("d_bv", u64), ("d_conn", u64),
        ("t_total", f64), ("t_ctrl", f64), ("t_nnd", f64), ("t_prune", f64), ("t_inter", f64), ("t_ep", f64),
        ("t_conn", f64), ("t_bv", f64), ("t_upload", f64),
        ("nnd_iterations", u32), ("nnd_stop", u32), ("nnd_recall", f32), ("nnd_delta", f32),
        ("n_loops", u32), ("n_kgraph", u32), ("n_recall", u32),
        ("kgraph_time", f64 * (FK_MAX_LOOPS + 64)),
        ("recall", f32 * (FK_MAX_LOOPS + 64)),
        ("prune_time", f64 * FK_MAX_LOOPS), ("inter_time", f64 * FK_MAX_LOOPS),
        ("tree_grow_time", f64 * FK_MAX_LOOPS),
        ("buildnsg_time", f64 * FK_MAX_LOOPS), ("search_time", f64 * FK_MAX_LOOPS),
        ("con_time", f64 * FK_MAX_LOOPS),
        ("prune_cc", f32 * FK_MAX_LOOPS), ("search_cc", f32 * FK_MAX_LOOPS),
        ("aod", u32), ("max_degree", u32), ("min_degree", u32), ("n_attached", u32),
        ("bv_rounds", u32 * FK_MAX_LOOPS),
        ("prof_ms", f64 * FK_PROF_N), ("prof_bytes", f64 * FK_PROF_N), ("prof_launches", u32 * FK_PROF_N),
        ("gram_mode", u32), ("reserved0", u32), ("d_gram", u64), ("d_recheck", u64),
        ("d_screened", u64), ("d_screen_skip", u64), ("screen_audit", u64), ("screen_missed", u64),
        ("screen_margin", f32), ("screen_loss", f32), ("screen_train_pairs", u32), ("reserved1", u32),
        ("t_screen", f64), ("screen_calib_pairs", u64), ("screen_cands", u64),
        ("d_bv_computed", u64), ("d_prune_uncomputed", u64),
        ("prune_variant", u32), ("prune_rev_variant", u32), ("t_tune", f64), ("d_bv_rejected", u64),
    ]

    def to_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class FkState(ctypes.Structure):
    _fields_ = [
        ("n", u32), ("pool_cap", u32), ("out_cap", u32), ("loopcount", u32),
        ("pool", ctypes.c_void_p), ("pool_len", ctypes.c_void_p), ("pool_flags", ctypes.c_void_p),
        ("out_id", ctypes.c_void_p), ("out_dist", ctypes.c_void_p), ("out_isnew", ctypes.c_void_p),
        ("out_len", ctypes.c_void_p), ("prev_id", ctypes.c_void_p), ("prev_len", ctypes.c_void_p),
    ]


class StepState:
    """A build state in host memory for the step entry points (fk_state):
    pools with their flags, out-lists with their isnew bits, and the
    previous out-lists, as numpy arrays the C structure points into."""

    def __init__(self, n, pool_cap, out_cap, loopcount=0):
        self.n, self.pool_cap, self.out_cap, self.loopcount = int(n), int(pool_cap), int(out_cap), int(loopcount)
        self.pool = np.zeros((n, pool_cap), np.uint64)
        self.pool_len = np.zeros(n, np.uint32)
        self.pool_flags = np.zeros((n, pool_cap), np.uint8)
        self.out_id = np.zeros((n, out_cap), np.uint32)
        self.out_dist = np.zeros((n, out_cap), np.float32)
        self.out_isnew = np.zeros((n, out_cap), np.uint8)
        self.out_len = np.zeros(n, np.uint32)
        self.prev_id = np.zeros((n, out_cap), np.uint32)
        self.prev_len = np.zeros(n, np.uint32)

    ARRAYS = ("pool", "pool_len", "pool_flags", "out_id", "out_dist", "out_isnew", "out_len", "prev_id", "prev_len")

    def copy(self):
        c = StepState.__new__(StepState)
        c.n, c.pool_cap, c.out_cap, c.loopcount = self.n, self.pool_cap, self.out_cap, self.loopcount
        for k in self.ARRAYS:
            setattr(c, k, getattr(self, k).copy())
        return c

    def struct(self):
        s = FkState(self.n, self.pool_cap, self.out_cap, self.loopcount)
        for k in self.ARRAYS:
            a = getattr(self, k)
            assert a.flags["C_CONTIGUOUS"] and a.shape[0] == self.n
            setattr(s, k, a.ctypes.data_as(ctypes.c_void_p))
        return s


class FkGraph(ctypes.Structure):
    _fields_ = [
        ("n", u32), ("ep", u32), ("width", u32),
        ("offsets", ctypes.POINTER(u64)),
        ("ids", ctypes.POINTER(u32)),
        ("stats", FkStats),
    ]


_P = ctypes.c_void_p
_PROTOS = {
    "fk_open": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    "fk_close": (None, [_P]),
    "fk_last_error": (ctypes.c_char_p, [_P]),
    "fk_version": (ctypes.c_char_p, []),
    "fk_dataset_upload": (ctypes.c_int, [_P, _P, ctypes.c_size_t, u32, u32]),
    "fk_build": (ctypes.c_int, [_P, ctypes.c_int, u32, ctypes.POINTER(FkParams), ctypes.POINTER(FkGraph)]),
    "fk_graph_free": (None, [ctypes.POINTER(FkGraph)]),
    "fk_build_batch": (ctypes.c_int, [_P, ctypes.c_int, u32, _P, ctypes.POINTER(FkParams), ctypes.POINTER(FkGraph)]),
    "fk_pair_distances": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t, _P]),
    "fk_gram_pair_distances": (ctypes.c_int, [_P, u32, _P, _P, ctypes.c_size_t, _P, _P, _P, ctypes.POINTER(u32)]),
    "fk_screen_param_count": (ctypes.c_size_t, [u32]),
    "fk_screen_grad": (ctypes.c_int, [_P, _P, _P, _P, _P, u32, f32, _P, ctypes.POINTER(f64)]),
    "fk_screen_predict": (ctypes.c_int, [_P, _P, _P, _P, u32, f32, _P]),
    "fk_join_distances": (ctypes.c_int, [_P, _P, _P, _P, u32, _P, ctypes.POINTER(u64)]),
    "fk_prune": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, u32, _P, u32, u32, ctypes.POINTER(FkParams),
                                ctypes.c_int, _P, _P, u32, _P, ctypes.POINTER(u64)]),
    "fk_prune_noalloc": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, u32, _P, u32, u32, ctypes.POINTER(FkParams),
                                        ctypes.c_int, _P, _P, u32, _P, ctypes.POINTER(u64)]),
    "fk_nndescent": (ctypes.c_int, [_P, u32, ctypes.POINTER(FkParams), _P, _P, u32, ctypes.POINTER(FkStats)]),
    "fk_knn_bruteforce": (ctypes.c_int, [_P, _P, u32, u32, u32, _P, _P, _P]),
    "fk_search": (ctypes.c_int, [_P, _P, _P, u32, u32, _P, u32, u32, u32, _P, ctypes.POINTER(u64)]),
    "fk_step_link": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(FkParams), ctypes.POINTER(FkState),
                                    ctypes.POINTER(FkStats)]),
    "fk_step_update_nhoods": (ctypes.c_int, [_P, ctypes.POINTER(FkState)]),
    "fk_step_bridgeview": (ctypes.c_int, [_P, ctypes.POINTER(FkParams), ctypes.POINTER(FkState),
                                          ctypes.POINTER(FkStats)]),
    "fk_step_entry_point": (ctypes.c_int, [_P, ctypes.POINTER(FkParams), ctypes.POINTER(FkState), u32,
                                           ctypes.POINTER(u32)]),
    "fk_step_connectivity": (ctypes.c_int, [_P, ctypes.POINTER(FkParams), ctypes.POINTER(FkState), u32,
                                            ctypes.POINTER(FkStats)]),
    "fk_step_stop_stats": (ctypes.c_int, [_P, ctypes.POINTER(FkParams), ctypes.POINTER(FkState), u32, _P, _P, u32,
                                          ctypes.POINTER(f32), ctypes.POINTER(f32)]),
}

_lib = None


class FkError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def load():
    """Load libfastkcna.so; raises if it is missing (build it first)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C fastkcna_amd/csrc`"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return sorted(_PROTOS)


def ptr(a):
    """Raw pointer of a C-contiguous numpy array (None for None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def params_struct(p):
    """FkParams from any object with the IndexParams attribute names."""
    s = FkParams()
    for name, _ in FkParams._fields_:
        setattr(s, name, getattr(p, name))
    return s
