"""On-disk formats of the reference (all little-endian).

lshkit     u32 elem_size (4), u32 rows, u32 cols, rows x cols float32
           (src/kgraph-data.h:262-287)
NSG index  u32 width, u32 ep, then per node u32 k and k u32 ids
           (SaveNsg, src/kgraph.cpp:1704-1724)
CSV log    SaveResult, src/kgraph.cpp:1725-1807 (append mode)
ivecs      per row i32 k then k i32 (ground-truth files)
"""
import struct

import numpy as np


def write_lshkit(path, data):
    data = np.ascontiguousarray(data, dtype=np.float32)
    with open(path, "wb") as f:
        f.write(struct.pack("<III", 4, data.shape[0], data.shape[1]))
        f.write(data.tobytes())


def read_lshkit(path):
    with open(path, "rb") as f:
        hdr = f.read(12)
        if len(hdr) != 12:
            raise IOError(path)
        elem, rows, cols = struct.unpack("<III", hdr)
        if elem != 4:
            raise IOError(f"{path}: element size {elem} != 4")
        data = np.fromfile(f, dtype=np.float32, count=rows * cols)
    if data.size != rows * cols:
        raise IOError(f"{path}: truncated")
    return data.reshape(rows, cols)


def save_nsg(path, width, ep, offsets, ids):
    offsets = np.asarray(offsets, dtype=np.uint64)
    ids = np.asarray(ids, dtype=np.uint32)
    n = len(offsets) - 1
    deg = np.diff(offsets).astype(np.uint32)
    with open(path, "wb") as f:
        f.write(struct.pack("<II", width, ep))
        # interleave [k, ids...] per node
        buf = np.empty(n + len(ids), dtype=np.uint32)
        pos = offsets[:-1].astype(np.int64) + np.arange(n, dtype=np.int64)
        buf[pos] = deg
        mask = np.ones(len(buf), bool)
        mask[pos] = False
        buf[mask] = ids
        f.write(buf.tobytes())


def load_nsg(path):
    raw = np.fromfile(path, dtype=np.uint32)
    width, ep = int(raw[0]), int(raw[1])
    offs = [0]
    ids = []
    p = 2
    while p < len(raw):
        k = int(raw[p])
        ids.append(raw[p + 1:p + 1 + k])
        offs.append(offs[-1] + k)
        p += 1 + k
    ids = np.concatenate(ids) if ids else np.zeros(0, np.uint32)
    return width, ep, np.asarray(offs, np.uint64), ids


def write_ivecs(path, a):
    a = np.asarray(a, dtype=np.int32)
    with open(path, "wb") as f:
        for row in a:
            f.write(struct.pack("<i", len(row)))
            f.write(row.tobytes())


def read_ivecs(path):
    raw = np.fromfile(path, dtype=np.int32)
    k = raw[0]
    return raw.reshape(-1, k + 1)[:, 1:]


def _fmt(v):
    """std::ostream default formatting of a float/double (6 significant digits)."""
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    s = "%.6g" % v
    return s


def save_result_csv(path, pg_type, params, st):
    """SaveResult's rows (src/kgraph.cpp:1725-1807), appended."""
    nl = int(st["n_loops"])
    nb = max(nl - 1, 0)
    rows = []
    rows.append("PG_type,K,L,S,R,I,searchL,searchK,nsgR,massqS,tau,angle,step")
    rows.append(",".join(_fmt(v) for v in [pg_type, params.K, params.L, params.S, params.R, params.iterations,
                                           params.search_L, params.search_K, params.nsg_R, params.massq_S,
                                           float(params.tau), float(params.angle), params.step]) + ",")

    def row(name, vals):
        rows.append(name + "," + "".join(_fmt(v) + "," for v in vals))

    row("kgraph time", st["kgraph_time"][: st["n_kgraph"]])
    row("recall", st["recall"][: st["n_recall"]])
    row("prune time", st["prune_time"][:nl])
    row("inter time", st["inter_time"][:nl])
    row("tree_grow time", st["tree_grow_time"][:nl] if pg_type != 2 else [])
    row("buildnsg time", st["buildnsg_time"][:nl])
    row("search time", st["search_time"][:nb])
    row("con time", st["con_time"][:nl])
    row("prune cc", st["prune_cc"][:nl])
    row("search cc", st["search_cc"][:nb])
    rows.append("AOD," + str(int(st["aod"])))
    rows.append("final_prune_time,-1")
    with open(path, "a") as f:
        f.write("\n".join(rows) + "\n")
