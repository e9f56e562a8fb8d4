"""NumPy-facing wrapper of one fk_handle (one GPU, one HIP stream, one
resident dataset).  Every method calls the C ABI; inputs and outputs are host
NumPy arrays."""
import ctypes

import numpy as np

from . import _ffi
from ._ffi import FkError, FkGraph, FkParams, FkStats, ptr


class _GraphMemory:
    """Owns a graph's host arrays (library allocations) until the numpy
    views of them are gone."""

    def __init__(self, lib, g):
        self.lib, self.g = lib, g

    def __del__(self):
        self.lib.fk_graph_free(ctypes.byref(self.g))


def _graph_arrays(lib, g):
    """(offsets, ids) of a built graph as numpy arrays over the library's
    own host buffers (no copy); the buffers are freed with the arrays."""
    mem = _GraphMemory(lib, g)
    n = g.n
    ob = (ctypes.c_uint64 * (n + 1)).from_address(ctypes.cast(g.offsets, ctypes.c_void_p).value)
    ob.owner = mem
    off = np.frombuffer(ob, dtype=np.uint64)
    ne = int(off[-1])
    ib = (ctypes.c_uint32 * max(ne, 1)).from_address(ctypes.cast(g.ids, ctypes.c_void_p).value)
    ib.owner = mem
    ids = np.frombuffer(ib, dtype=np.uint32)[:ne]
    return off, ids


def _as(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Device:
    def __init__(self, device=0):
        self.lib = _ffi.load()
        h = ctypes.c_void_p()
        st = self.lib.fk_open(device, ctypes.byref(h))
        if st != _ffi.FK_OK:
            raise FkError(st, f"fk_open({device}) failed (no gfx950 device?)")
        self.h = h
        self.n_total = 0
        self.d = 0

    def close(self):
        if self.h:
            self.lib.fk_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != _ffi.FK_OK:
            msg = self.lib.fk_last_error(self.h)
            raise FkError(st, f"{what}: {msg.decode() if msg else ''}")

    # -- dataset -----------------------------------------------------------
    def upload(self, rows):
        rows = _as(rows, np.float32)
        assert rows.ndim == 2
        self._check(
            self.lib.fk_dataset_upload(self.h, ptr(rows), rows.strides[0], rows.shape[0], rows.shape[1]),
            "fk_dataset_upload",
        )
        self.n_total, self.d = rows.shape

    # -- unit entry points -------------------------------------------------
    def pair_distances(self, a, b):
        a = _as(a, np.uint32)
        b = _as(b, np.uint32)
        out = np.empty(len(a), np.float32)
        self._check(self.lib.fk_pair_distances(self.h, ptr(a), ptr(b), len(a), ptr(out)), "fk_pair_distances")
        return out

    def gram_pair_distances(self, a, b, flags=0):
        """MFMA Gram-form distances of resident row pairs: (dist, lo, hi,
        mode), lo <= canonical FP32 distance <= hi."""
        a = _as(a, np.uint32)
        b = _as(b, np.uint32)
        n = len(a)
        out, lo, hi = (np.empty(n, np.float32) for _ in range(3))
        mode = ctypes.c_uint32()
        self._check(self.lib.fk_gram_pair_distances(self.h, flags, ptr(a), ptr(b), n, ptr(out), ptr(lo), ptr(hi),
                                                    ctypes.byref(mode)), "fk_gram_pair_distances")
        return out, lo, hi, mode.value

    def screen_param_count(self):
        return int(self.lib.fk_screen_param_count(self.d))

    def screen_grad(self, weights, a, b, dist, inv_mu):
        """Gradient and loss of one training batch of the learned screen."""
        w = _as(weights, np.float32)
        a, b, dist = _as(a, np.uint32), _as(b, np.uint32), _as(dist, np.float32)
        g = np.empty_like(w)
        loss = ctypes.c_double()
        self._check(self.lib.fk_screen_grad(self.h, ptr(w), ptr(a), ptr(b), ptr(dist), len(a), inv_mu, ptr(g),
                                            ctypes.byref(loss)), "fk_screen_grad")
        return g, loss.value

    def screen_predict(self, weights, a, b, inv_mu):
        w = _as(weights, np.float32)
        a, b = _as(a, np.uint32), _as(b, np.uint32)
        out = np.empty(len(a), np.float32)
        self._check(self.lib.fk_screen_predict(self.h, ptr(w), ptr(a), ptr(b), len(a), inv_mu, ptr(out)),
                    "fk_screen_predict")
        return out

    def join_distances(self, ids, list_off, n_new):
        ids = _as(ids, np.uint32)
        list_off = _as(list_off, np.uint64)
        n_new = _as(n_new, np.uint32)
        nn = n_new.astype(np.uint64)
        T = np.diff(list_off)
        total = int(np.sum(nn * (nn - (nn > 0)) // 2 + nn * (T - nn)))
        out = np.empty(max(total, 1), np.float32)
        npairs = ctypes.c_uint64()
        self._check(
            self.lib.fk_join_distances(self.h, ptr(ids), ptr(list_off), ptr(n_new), len(n_new), ptr(out),
                                       ctypes.byref(npairs)),
            "fk_join_distances",
        )
        assert npairs.value == total
        return out[:total]

    def prune(self, pg_type, pools, pool_len, q, oracle_size, params, reversed_=False, noalloc=False):
        """Occlusion prune of the given pools (fk_prune, or fk_prune_noalloc
        when noalloc: the same result with no device allocation)."""
        pools = _as(pools, np.uint64)
        pool_len = _as(pool_len, np.uint32)
        n_tasks, stride = pools.shape
        q = None if q is None else _as(q, np.uint32)
        R = params.nsg_R
        out_ids = np.zeros((n_tasks, R), np.uint32)
        out_d = np.zeros((n_tasks, R), np.float32)
        out_len = np.zeros(n_tasks, np.uint32)
        cc = ctypes.c_uint64()
        p = _ffi.params_struct(params)
        fn, name = (self.lib.fk_prune_noalloc, "fk_prune_noalloc") if noalloc else (self.lib.fk_prune, "fk_prune")
        self._check(
            fn(self.h, pg_type, ptr(pools), ptr(pool_len), stride, ptr(q), n_tasks, oracle_size, ctypes.byref(p),
               1 if reversed_ else 0, ptr(out_ids), ptr(out_d), R, ptr(out_len), ctypes.byref(cc)),
            name,
        )
        return out_ids, out_d, out_len, cc.value

    def knn(self, queries, K, n_used=None, self_ids=None):
        queries = _as(queries, np.float32)
        n_q = queries.shape[0]
        n_used = self.n_total if n_used is None else n_used
        self_ids = None if self_ids is None else _as(self_ids, np.uint32)
        ids = np.empty((n_q, K), np.uint32)
        d = np.empty((n_q, K), np.float32)
        self._check(
            self.lib.fk_knn_bruteforce(self.h, ptr(queries), n_q, n_used, K, ptr(self_ids), ptr(ids), ptr(d)),
            "fk_knn_bruteforce",
        )
        return ids, d

    def nndescent(self, params, n_used=None):
        n_used = self.n_total if n_used is None else n_used
        L = min(params.L, n_used - 1)
        pools = np.zeros((n_used, L), np.uint64)
        plen = np.zeros(n_used, np.uint32)
        st = FkStats()
        p = _ffi.params_struct(params)
        self._check(
            self.lib.fk_nndescent(self.h, n_used, ctypes.byref(p), ptr(pools), ptr(plen), L, ctypes.byref(st)),
            "fk_nndescent",
        )
        return pools, plen, st.to_dict()

    def build(self, pg_type, params, n_used=None):
        """Full build; returns (offsets, ids, ep, width, stats dict)."""
        n_used = self.n_total if n_used is None else n_used
        g = FkGraph()
        p = _ffi.params_struct(params)
        self._check(self.lib.fk_build(self.h, pg_type, n_used, ctypes.byref(p), ctypes.byref(g)), "fk_build")
        off, ids = _graph_arrays(self.lib, g)
        return off, ids, int(g.ep), int(g.width), g.stats.to_dict()

    def build_batch(self, pg_type, params, set_off):
        """Batched build of the resident row ranges [set_off[s], set_off[s+1]);
        returns one (offsets, ids, ep, width) per set (ids relative to the
        set) and the batch's statistics."""
        set_off = _as(set_off, np.uint64)
        n = len(set_off) - 1
        gs = (FkGraph * n)()
        p = _ffi.params_struct(params)
        self._check(self.lib.fk_build_batch(self.h, pg_type, n, ptr(set_off), ctypes.byref(p), gs), "fk_build_batch")
        out, st = [], None
        try:
            for g in gs:
                off = np.ctypeslib.as_array(g.offsets, shape=(g.n + 1,)).copy()
                ne = int(off[-1])
                ids = np.ctypeslib.as_array(g.ids, shape=(max(ne, 1),))[:ne].copy()
                out.append((off, ids, int(g.ep), int(g.width)))
            st = gs[0].stats.to_dict()
        finally:
            for g in gs:
                self.lib.fk_graph_free(ctypes.byref(g))
        return out, st

    # -- step entry points (one step of the loop on a supplied state) -----
    def step_link(self, pg_type, params, state):
        """One Link step on `state` (a StepState): prune, reverse edges,
        reversed prune; fills the out-lists; returns the statistics."""
        st = FkStats()
        p = _ffi.params_struct(params)
        s = state.struct()
        self._check(self.lib.fk_step_link(self.h, pg_type, ctypes.byref(p), ctypes.byref(s), ctypes.byref(st)),
                    "fk_step_link")
        return st.to_dict()

    def step_update_nhoods(self, state):
        s = state.struct()
        self._check(self.lib.fk_step_update_nhoods(self.h, ctypes.byref(s)), "fk_step_update_nhoods")

    def step_bridgeview(self, params, state):
        """One BridgeView pass on `state` at state.loopcount; updates its
        pools and flags; returns the statistics."""
        st = FkStats()
        p = _ffi.params_struct(params)
        s = state.struct()
        self._check(self.lib.fk_step_bridgeview(self.h, ctypes.byref(p), ctypes.byref(s), ctypes.byref(st)),
                    "fk_step_bridgeview")
        return st.to_dict()

    def step_entry_point(self, params, state, init_node):
        """ComputeEp on the state's pools and out-lists from init_node's pool."""
        ep = ctypes.c_uint32()
        p = _ffi.params_struct(params)
        s = state.struct()
        self._check(self.lib.fk_step_entry_point(self.h, ctypes.byref(p), ctypes.byref(s), init_node,
                                                 ctypes.byref(ep)), "fk_step_entry_point")
        return ep.value

    def step_connectivity(self, params, state, ep):
        """The connectivity pass on the state's out-lists and pools from
        entry point ep; attached edges appended to the out-lists; returns the
        statistics (n_attached, d_conn)."""
        st = FkStats()
        p = _ffi.params_struct(params)
        s = state.struct()
        self._check(self.lib.fk_step_connectivity(self.h, ctypes.byref(p), ctypes.byref(s), ep, ctypes.byref(st)),
                    "fk_step_connectivity")
        return st.to_dict()

    def step_stop_stats(self, params, state, ctrl_ids, gt):
        """(delta, recall) of the stop rule on the state's pools."""
        ctrl_ids = _as(ctrl_ids, np.uint32)
        gt = _as(gt, np.uint32)
        delta, recall = ctypes.c_float(), ctypes.c_float()
        p = _ffi.params_struct(params)
        s = state.struct()
        self._check(self.lib.fk_step_stop_stats(self.h, ctypes.byref(p), ctypes.byref(s), len(ctrl_ids),
                                                ptr(ctrl_ids), ptr(gt), gt.shape[1] if gt.ndim == 2 else 0,
                                                ctypes.byref(delta), ctypes.byref(recall)), "fk_step_stop_stats")
        return delta.value, recall.value

    def search(self, offsets, ids, ep, queries, search_L, K):
        offsets = _as(offsets, np.uint64)
        ids = _as(ids, np.uint32)
        queries = _as(queries, np.float32)
        n = len(offsets) - 1
        out = np.empty((queries.shape[0], K), np.uint32)
        nd = ctypes.c_uint64()
        self._check(
            self.lib.fk_search(self.h, ptr(offsets), ptr(ids), n, ep, ptr(queries), queries.shape[0], search_L, K,
                               ptr(out), ctypes.byref(nd)),
            "fk_search",
        )
        return out, nd.value
