"""CPU oracle for FastKCNA's graph-construction path -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this package, and only as the checker (or the timed CPU baseline); the
product path (fastkcna_amd, libfastkcna.so, the C++ host layer) never calls
it.  The restatement itself is fk_oracle.cpp (C++/OpenMP, each part citing
the reference file:line it follows); this module binds it with ctypes.

Pinning (what makes the oracle trustworthy without running the reference,
which cannot be built here: its every translation unit includes Boost, which
this image lacks):
  * distance: bit-equal to an AVX intrinsics kernel written in the
    reference's construction and compiled the reference's way (avx_pin.cpp);
    exact on integer data; within 1e-6 relative of FP64 on float data
  * pool insert: brute-force invariants (sorted, unique, best-L kept)
  * pruners: closed forms (tau = 0 and the NSG rule coincide; nothing
    occluded -> identity; duplicates dropped) and an independent
    re-derivation of every keep/drop decision from its definition
  * NN-descent / full build: recall against brute force, graph invariants
  * steps (tests/test_oracle_steps.py): BridgeView, Link / InterInsert,
    UpdateNhoods, the stop rule, ComputeEp and the connectivity pass's
    attach order on hand-built states
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

FK_MAX_LOOPS = 16


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libfkoracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        lib.fko_l2.restype = ctypes.c_float
        lib.fko_l2.argtypes = [P, P, ctypes.c_uint]
        lib.fko_l2_f64.restype = ctypes.c_double
        lib.fko_l2_f64.argtypes = [P, P, ctypes.c_uint]
        lib.fko_pair_distances.restype = None
        lib.fko_pair_distances.argtypes = [P, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint, P, P, ctypes.c_size_t,
                                           P, P]
        lib.fko_pool_insert_sequence.restype = ctypes.c_uint
        lib.fko_pool_insert_sequence.argtypes = [P, P, ctypes.c_uint, ctypes.c_uint, P, P, P]
        lib.fko_prune.restype = None
        lib.fko_prune.argtypes = [ctypes.c_int, P, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint, P, P, ctypes.c_uint,
                                  P, ctypes.c_uint, ctypes.c_uint, P, ctypes.c_int, P, P, ctypes.c_uint, P, P]
        lib.fko_knn.restype = None
        lib.fko_knn.argtypes = [P, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint, P, ctypes.c_uint, ctypes.c_uint, P,
                                P, P]
        lib.fko_build.restype = ctypes.c_int
        lib.fko_build.argtypes = [ctypes.c_int, P, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, P,
                                  P, P, P, P, P, P, P]
        lib.fko_nndescent.restype = ctypes.c_int
        lib.fko_nndescent.argtypes = [P, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint, P, ctypes.c_int, P, P,
                                      ctypes.c_uint, P]
        lib.fko_search.restype = None
        lib.fko_search.argtypes = [P, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint, P, P, ctypes.c_uint,
                                   ctypes.c_uint, P, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint64, P,
                                   P]
        lib.fko_step_entry_point.restype = ctypes.c_uint
        lib.fko_step_entry_point.argtypes = [P, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint, P, ctypes.c_void_p,
                                             ctypes.c_uint]
        PS = ctypes.c_void_p
        lib.fko_step_link.restype = ctypes.c_int
        lib.fko_step_link.argtypes = [ctypes.c_int, P, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint, P, PS, P]
        lib.fko_step_update_nhoods.restype = ctypes.c_int
        lib.fko_step_update_nhoods.argtypes = [PS]
        lib.fko_step_bridgeview.restype = ctypes.c_int
        lib.fko_step_bridgeview.argtypes = [P, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint, P, PS, P]
        lib.fko_step_connectivity.restype = ctypes.c_int
        lib.fko_step_connectivity.argtypes = [P, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint, P, PS, ctypes.c_uint,
                                              P]
        lib.fko_step_stop_stats.restype = None
        lib.fko_step_stop_stats.argtypes = [P, PS, ctypes.c_uint, P, P, ctypes.c_uint, P, P]
        lib.fko_free.restype = None
        lib.fko_free.argtypes = [P]
        _lib = lib
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _params(p):
    from fastkcna_amd._ffi import params_struct  # struct layout only; no library call
    return params_struct(p)


def _stats():
    from fastkcna_amd._ffi import FkStats
    return FkStats()


class BuildOpts(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_int), ("full_rule_only", ctypes.c_int), ("check_prune2", ctypes.c_int),
                ("verbose", ctypes.c_int)]


def pad8(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    d = x.shape[-1]
    dp = (d + 7) // 8 * 8
    if dp == d:
        return x
    out = np.zeros(x.shape[:-1] + (dp,), np.float32)
    out[..., :d] = x
    return out


def l2(a, b):
    """Reference-order squared L2 of two vectors."""
    a, b = pad8(a), pad8(b)
    return _load().fko_l2(_p(a), _p(b), a.shape[-1])


def l2_f64(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return _load().fko_l2_f64(_p(a), _p(b), a.shape[-1])


def pair_distances(rows, a, b, with_f64=False):
    rows = np.ascontiguousarray(rows, np.float32)
    a = np.ascontiguousarray(a, np.uint32)
    b = np.ascontiguousarray(b, np.uint32)
    out = np.empty(len(a), np.float32)
    out64 = np.empty(len(a), np.float64) if with_f64 else None
    _load().fko_pair_distances(_p(rows), rows.strides[0], rows.shape[0], rows.shape[1], _p(a), _p(b), len(a),
                               _p(out), _p(out64))
    return (out, out64) if with_f64 else out


def pool_insert_sequence(ids, dists, L):
    ids = np.ascontiguousarray(ids, np.uint32)
    dists = np.ascontiguousarray(dists, np.float32)
    oi = np.zeros(L + 1, np.uint32)
    od = np.zeros(L + 1, np.float32)
    ret = np.zeros(len(ids), np.int32)
    n = _load().fko_pool_insert_sequence(_p(ids), _p(dists), len(ids), L, _p(oi), _p(od), _p(ret))
    return oi[:n], od[:n], ret


def prune(pg_type, rows, pools, pool_len, q, oracle_size, params, reversed_=False):
    rows = np.ascontiguousarray(rows, np.float32)
    pools = np.ascontiguousarray(pools, np.uint64)
    pool_len = np.ascontiguousarray(pool_len, np.uint32)
    q = None if q is None else np.ascontiguousarray(q, np.uint32)
    n_tasks, stride = pools.shape
    R = params.nsg_R
    out_ids = np.zeros((n_tasks, R), np.uint32)
    out_d = np.zeros((n_tasks, R), np.float32)
    out_len = np.zeros(n_tasks, np.uint32)
    cc = ctypes.c_uint64()
    ps = _params(params)
    _load().fko_prune(pg_type, _p(rows), rows.strides[0], rows.shape[0], rows.shape[1], _p(pools), _p(pool_len),
                      stride, _p(q), n_tasks, oracle_size, ctypes.byref(ps), 1 if reversed_ else 0, _p(out_ids),
                      _p(out_d), R, _p(out_len), ctypes.byref(cc))
    return out_ids, out_d, out_len, cc.value


def knn(rows, queries, K, self_ids=None):
    rows = np.ascontiguousarray(rows, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    self_ids = None if self_ids is None else np.ascontiguousarray(self_ids, np.uint32)
    ids = np.empty((queries.shape[0], K), np.uint32)
    d = np.empty((queries.shape[0], K), np.float32)
    _load().fko_knn(_p(rows), rows.strides[0], rows.shape[0], rows.shape[1], _p(queries), queries.shape[0], K,
                    _p(self_ids), _p(ids), _p(d))
    return ids, d


def build(pg_type, rows, params, n_used=None, threads=0, full_rule_only=False, check_prune2=False, verbose=False):
    """Full CPU build; returns (offsets, ids, ep, width, stats dict, prune2)
    with prune2 = (nodes where sync_prune2 differed from the full rule,
    nodes where the two were compared) when check_prune2 is set."""
    rows = np.ascontiguousarray(rows, np.float32)
    n_used = rows.shape[0] if n_used is None else n_used
    lib = _load()
    off_p = ctypes.POINTER(ctypes.c_uint64)()
    ids_p = ctypes.POINTER(ctypes.c_uint32)()
    ep = ctypes.c_uint32()
    width = ctypes.c_uint32()
    st = _stats()
    diff = (ctypes.c_uint64 * 2)()
    opts = BuildOpts(threads, int(full_rule_only), int(check_prune2), int(verbose))
    ps = _params(params)
    lib.fko_build(pg_type, _p(rows), rows.strides[0], rows.shape[0], n_used, rows.shape[1], ctypes.byref(ps),
                  ctypes.byref(opts), ctypes.byref(off_p), ctypes.byref(ids_p), ctypes.byref(ep),
                  ctypes.byref(width), ctypes.byref(st), diff)
    off = np.ctypeslib.as_array(off_p, shape=(n_used + 1,)).copy()
    ne = int(off[-1])
    ids = np.ctypeslib.as_array(ids_p, shape=(max(ne, 1),))[:ne].copy()
    lib.fko_free(ctypes.cast(off_p, ctypes.c_void_p))
    lib.fko_free(ctypes.cast(ids_p, ctypes.c_void_p))
    return off, ids, ep.value, width.value, st.to_dict(), (int(diff[0]), int(diff[1]))


def nndescent(rows, params, threads=0):
    rows = np.ascontiguousarray(rows, np.float32)
    n = rows.shape[0]
    L = min(params.L, n - 1)
    pools = np.zeros((n, L), np.uint64)
    plen = np.zeros(n, np.uint32)
    st = _stats()
    ps = _params(params)
    _load().fko_nndescent(_p(rows), rows.strides[0], n, rows.shape[1], ctypes.byref(ps), threads, _p(pools), _p(plen),
                          L, ctypes.byref(st))
    return pools, plen, st.to_dict()


def pack(dists, ids):
    """Packed pool keys (float bits << 32 | id) from float32 dists and ids."""
    d = np.ascontiguousarray(dists, np.float32).view(np.uint32).astype(np.uint64)
    return (d << np.uint64(32)) | np.asarray(ids, np.uint64)


def unpack(keys):
    keys = np.asarray(keys, np.uint64)
    ids = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    d = (keys >> np.uint64(32)).astype(np.uint32).view(np.float32)
    return d, ids


# -- step entry points: one step of the loop on a supplied state -------------
# `state` is a fastkcna_amd._ffi.StepState (plain numpy arrays; the structure
# layout is shared with the library, no library call is made).
def step_link(pg_type, rows, params, state):
    """Link step (full prune rule, InterInsert, reversed prune) on the state;
    appended reverse edges of lists within nsg_R in (dist, id) order."""
    rows = np.ascontiguousarray(rows, np.float32)
    st = _stats()
    ps = _params(params)
    s = state.struct()
    rc = _load().fko_step_link(pg_type, _p(rows), rows.strides[0], rows.shape[0], rows.shape[1], ctypes.byref(ps),
                               ctypes.byref(s), ctypes.byref(st))
    assert rc == 0, "out-list above out_cap"
    return st.to_dict()


def step_update_nhoods(state):
    rc = _load().fko_step_update_nhoods(ctypes.byref(state.struct()))
    assert rc == 0


def step_bridgeview(rows, params, state):
    """One BridgeView pass on the state at state.loopcount (pools, lengths and
    flags updated in place)."""
    rows = np.ascontiguousarray(rows, np.float32)
    st = _stats()
    ps = _params(params)
    s = state.struct()
    rc = _load().fko_step_bridgeview(_p(rows), rows.strides[0], rows.shape[0], rows.shape[1], ctypes.byref(ps),
                                     ctypes.byref(s), ctypes.byref(st))
    assert rc == 0, "pool_cap must equal MAXL"
    return st.to_dict()


def step_entry_point(rows, params, state, init_node):
    """ComputeEp on the state (pools, out-lists) from init_node's pool."""
    rows = np.ascontiguousarray(rows, np.float32)
    ps = _params(params)
    return int(_load().fko_step_entry_point(_p(rows), rows.strides[0], rows.shape[0], rows.shape[1],
                                            ctypes.byref(ps), ctypes.byref(state.struct()), init_node))


def step_connectivity(rows, params, state, ep):
    """The connectivity pass (DFS from ep, findroot, attach) on the state's
    out-lists and pools; attached edges appended to the out-lists."""
    rows = np.ascontiguousarray(rows, np.float32)
    st = _stats()
    ps = _params(params)
    rc = _load().fko_step_connectivity(_p(rows), rows.strides[0], rows.shape[0], rows.shape[1], ctypes.byref(ps),
                                       ctypes.byref(state.struct()), ep, ctypes.byref(st))
    assert rc == 0, "out-list above out_cap"
    return st.to_dict()


def step_stop_stats(params, state, ctrl_ids, gt):
    ctrl_ids = np.ascontiguousarray(ctrl_ids, np.uint32)
    gt = np.ascontiguousarray(gt, np.uint32)
    delta, recall = ctypes.c_float(), ctypes.c_float()
    ps = _params(params)
    _load().fko_step_stop_stats(ctypes.byref(ps), ctypes.byref(state.struct()), len(ctrl_ids), _p(ctrl_ids), _p(gt),
                                gt.shape[1] if gt.ndim == 2 else 0, ctypes.byref(delta), ctypes.byref(recall))
    return delta.value, recall.value


SEARCH_SEED = 0x9E3779B97F4A7C15  # the evaluator's fill stream seed (fk_search)


def search(rows, offsets, ids, ep, queries, search_L, K):
    """CPU restatement of the query evaluator (fk_search): best-first beam
    from ep over a CSR graph; returns (ids [n_q, K], distances computed)."""
    rows = np.ascontiguousarray(rows, np.float32)
    offsets = np.ascontiguousarray(offsets, np.uint64)
    ids = np.ascontiguousarray(ids, np.uint32)
    queries = np.ascontiguousarray(queries, np.float32)
    n = len(offsets) - 1
    out = np.empty((len(queries), K), np.uint32)
    nd = ctypes.c_uint64()
    _load().fko_search(_p(rows), rows.strides[0], rows.shape[0], rows.shape[1], _p(offsets), _p(ids), n, ep,
                       _p(queries), len(queries), search_L, K, SEARCH_SEED, _p(out), ctypes.byref(nd))
    return out, nd.value
