"""The Gram-form (MFMA) local join against the plain FP32 join.

Bars: on integer data the Gram distance is the canonical FP32 distance bit for
bit (exact form); on float data every Gram distance comes with bounds that
contain the canonical distance, and its error against FP64 is reported; the
NN-descent pools and the whole NSG build with FK_FLAG_GRAM are bit-identical
to the plain FP32 path's, in every operand format (the recheck makes the
candidate sets equal, so pools hold only exact FP32 distances).
"""
import sys

import numpy as np
import pytest

import oracle
from fastkcna_amd import FK_PG_ALPHAPG, FK_PG_HNSW_LAYER, FK_PG_NSG, FK_PG_TAUMNG, IndexParams, _ffi
from fastkcna_amd.datasets import float_mixture, integer_mixture, unit_mixture

pytestmark = pytest.mark.gpu

GRAM = _ffi.FK_FLAG_GRAM
BF16 = _ffi.FK_FLAG_GRAM_BF16
F16 = _ffi.FK_FLAG_GRAM_F16


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def data(name):
    if name == "int128":
        return integer_mixture(3000, 128, 16, seed=21)
    if name == "float960":
        return float_mixture(1200, 960, 8, seed=22)
    if name == "unit96":
        return unit_mixture(3000, 96, 16, seed=23, sigma=0.25)
    if name == "odd100":
        return np.random.default_rng(24).normal(size=(2000, 100)).astype(np.float32)
    if name == "odd20":
        return (np.random.default_rng(25).normal(size=(2000, 20)) * 1e3).astype(np.float32)
    raise KeyError(name)


def pairs_with_near_duplicates(x, n, rng):
    a = rng.integers(0, len(x), n)
    b = rng.integers(0, len(x), n)
    b[: n // 8] = a[: n // 8]  # identical rows: the worst cancellation
    return a, b


@pytest.mark.parametrize("name", ["int128", "odd20"])
def test_gram_exact_form_is_canonical(device, name):
    x = data(name)
    if name == "odd20":  # integer-valued but signed: exact form needs |x|^2 < 2^22
        x = np.rint(x / 64.0).astype(np.float32)
    say("gram exact", name, x.shape)
    device.upload(x)
    rng = np.random.default_rng(3)
    a, b = pairs_with_near_duplicates(x, 20000, rng)
    g, lo, hi, mode = device.gram_pair_distances(a, b)
    assert mode == _ffi.FK_GRAM_EXACT
    c, c64 = oracle.pair_distances(x, a, b, with_f64=True)
    assert np.array_equal(g.view(np.uint32), c.view(np.uint32))
    assert np.array_equal(lo, g) and np.array_equal(hi, g)
    assert np.array_equal(g.astype(np.float64), c64)


@pytest.mark.parametrize("name,flags,mode,max_rel", [
    ("float960", 0, _ffi.FK_GRAM_BF16, 2e-1),
    ("unit96", 0, _ffi.FK_GRAM_F16, 2e-2),
    ("odd100", 0, _ffi.FK_GRAM_BF16, 1e-1),
    ("float960", F16, _ffi.FK_GRAM_F16, 2e-2),
    ("int128", BF16, _ffi.FK_GRAM_BF16, 1e-1),
    ("unit96", BF16, _ffi.FK_GRAM_BF16, 1e-1),
])
def test_gram_bounds_contain_canonical(device, name, flags, mode, max_rel):
    x = data(name)
    say("gram bounds", name, flags)
    device.upload(x)
    rng = np.random.default_rng(4)
    a, b = pairs_with_near_duplicates(x, 40000, rng)
    g, lo, hi, m = device.gram_pair_distances(a, b, flags)
    assert m == mode
    c, c64 = oracle.pair_distances(x, a, b, with_f64=True)
    assert np.all(lo <= c) and np.all(c <= hi)
    assert np.all(lo <= hi)
    # error of the Gram form against FP64, over pairs of distinct rows
    nz = c64 > 0
    rel = np.abs(g[nz].astype(np.float64) - c64[nz]) / c64[nz]
    say(f"  {name} mode {_ffi.GRAM_MODE_NAMES[m]}: rel err vs FP64 median {np.median(rel):.2e} "
        f"p99 {np.quantile(rel, 0.99):.2e} max {rel.max():.2e}")
    assert rel.max() < max_rel


def nnd_params(**kw):
    p = IndexParams(K=40, L=40, S=10, R=50, iterations=5, nsg_R=16, search_L=40, search_K=40)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("name,flags", [
    ("int128", GRAM),           # exact form
    ("int128", GRAM | BF16),    # bounded BF16 on the same data
    ("float960", GRAM | BF16),  # bounded BF16
    ("unit96", GRAM | F16),     # bounded FP16
    ("odd100", GRAM | F16),     # bounded FP16, dpad % 16 == 8
])
def test_gram_nndescent_pools_identical(device, name, flags):
    x = data(name)
    say("gram nnd", name, flags)
    device.upload(x)
    p = nnd_params()
    pools0, len0, st0 = device.nndescent(p)
    pools0b, len0b, _ = device.nndescent(p)
    assert np.array_equal(pools0, pools0b) and np.array_equal(len0, len0b), "plain NN-descent is not repeatable"
    p.flags = flags
    pools1, len1, st1 = device.nndescent(p)
    assert np.array_equal(len0, len1)
    assert np.array_equal(pools0, pools1)
    assert st1["d_gram"] == st0["d_join"] == st1["d_join"]
    if flags == GRAM and name == "int128":
        assert st1["gram_mode"] == _ffi.FK_GRAM_EXACT and st1["d_recheck"] == 0
    else:
        assert 0 < st1["d_recheck"] <= st1["d_gram"]
    say(f"  {name}: mode {_ffi.GRAM_MODE_NAMES[st1['gram_mode']]}, {st1['d_gram']} Gram pairs, "
        f"{st1['d_recheck']} rechecked ({st1['d_recheck'] / max(st1['d_gram'], 1):.3f})")
    if flags != GRAM:
        # the format picked from the data: the Gram join only on exact-form
        # rows, the plain join (same pools) on the bounded forms
        p.flags = GRAM
        pools2, len2, st2 = device.nndescent(p)
        assert np.array_equal(len0, len2) and np.array_equal(pools0, pools2)
        assert (st2["d_gram"] > 0) == (st2["gram_mode"] == _ffi.FK_GRAM_EXACT)


@pytest.mark.parametrize("S,R,K", [(20, 150, 100), (40, 300, 180)])
def test_gram_join_tiled_lists_identical(device, S, R, K):
    """Long sample lists (above the LDS staging area: 77 FP32 rows, 301 byte
    rows at d = 128) take the tiled path of the Gram join; the pools must
    still equal the plain join's."""
    x = data("int128")
    say("gram tiled lists", S, R, K)
    device.upload(x)
    p = nnd_params(K=K, L=K, S=S, R=R, iterations=4)
    pools0, len0, _ = device.nndescent(p)
    p.flags = GRAM | BF16
    pools1, len1, st1 = device.nndescent(p)
    assert np.array_equal(len0, len1) and np.array_equal(pools0, pools1)
    p.flags = GRAM
    pools2, len2, _ = device.nndescent(p)
    assert np.array_equal(len0, len2) and np.array_equal(pools0, pools2)


@pytest.mark.parametrize("name,pg", [("int128", FK_PG_NSG), ("float960", FK_PG_TAUMNG)])
def test_gram_full_build_identical(device, name, pg):
    x = data(name)
    say("gram build", name, pg)
    device.upload(x)
    p = nnd_params(tau=0.01)
    off0, ids0, ep0, w0, st0 = device.build(pg, p)
    p.flags = GRAM | (0 if name == "int128" else BF16)
    off1, ids1, ep1, w1, st1 = device.build(pg, p)
    assert ep0 == ep1 and w0 == w1
    assert np.array_equal(off0, off1) and np.array_equal(ids0, ids1)
    assert st1["d_gram"] > 0


@pytest.mark.parametrize("d", [96, 128, 256])
@pytest.mark.parametrize("pg", [FK_PG_HNSW_LAYER, FK_PG_ALPHAPG])
def test_gram_bound_pass_build_identical(device, d, pg):
    """FP16-form rows of widths 96, 128 and 256 with FK_FLAG_GRAM: BridgeView
    drops targets whose low-precision bound (8-bit rows at widths 96 and
    256, 16-bit ones at 128) puts them beyond the pool's last distance, and the
    Gram-form prune may run; the graph, the distance counts and the
    BridgeView pairs evaluated equal the plain build's, and the bound must
    have dropped some targets."""
    x = unit_mixture(5000, d, 16, seed=26 + d, sigma=0.3)
    say("gram bound pass", d, pg)
    device.upload(x)
    p = nnd_params(K=40, L=40, S=10, R=40, iterations=4, nsg_R=24, search_L=40, search_K=60, angle=60)
    off0, ids0, ep0, w0, st0 = device.build(pg, p)
    p.flags = GRAM
    off1, ids1, ep1, w1, st1 = device.build(pg, p)
    assert st1["gram_mode"] == _ffi.FK_GRAM_F16
    assert ep0 == ep1 and w0 == w1
    assert np.array_equal(off0, off1) and np.array_equal(ids0, ids1)
    for k in ("d_join", "d_prune", "d_prune_rev", "d_bv", "d_bv_computed", "d_ep"):
        assert st1[k] == st0[k], k
    assert st0["d_bv_rejected"] == 0 and 0 < st1["d_bv_rejected"] < st1["d_bv_computed"]
    say(f"  rejected {st1['d_bv_rejected']} of {st1['d_bv_computed']} computed BridgeView pairs; "
        f"prune variant {st1['prune_variant'] - 1}")


@pytest.mark.parametrize("pg", [FK_PG_HNSW_LAYER, FK_PG_NSG])
def test_bound_pass_8bit_rows_unnormalised_with_duplicates(device, pg):
    """The 8-bit bound pass at width 96 on rows that are not unit-normalised
    (FP16 form forced; per-row scales from 0.02 to 50) with exact and near
    duplicates (pairs at distance 0 and ~1e-6 of the row norm): the graph and
    counts equal the plain build's and some targets are dropped."""
    rng = np.random.default_rng(27)
    x = float_mixture(4000, 96, 16, seed=27)
    x *= np.exp(rng.uniform(np.log(0.02), np.log(50.0), size=(len(x), 1))).astype(np.float32)
    dup = rng.choice(len(x), 400, replace=False)
    x[dup[:200]] = x[dup[200:]]
    x[dup[:100]] *= np.float32(1.0 + 1e-6)
    say("8-bit bound pass, unnormalised rows with duplicates", pg)
    device.upload(x)
    p = nnd_params(K=40, L=40, S=10, R=40, iterations=4, nsg_R=24, search_L=40, search_K=60, angle=60)
    off0, ids0, ep0, w0, st0 = device.build(pg, p)
    p.flags = GRAM | F16
    off1, ids1, ep1, w1, st1 = device.build(pg, p)
    assert st1["gram_mode"] == _ffi.FK_GRAM_F16
    assert ep0 == ep1 and w0 == w1
    assert np.array_equal(off0, off1) and np.array_equal(ids0, ids1)
    for k in ("d_join", "d_prune", "d_prune_rev", "d_bv", "d_bv_computed", "d_ep"):
        assert st1[k] == st0[k], k
    assert 0 < st1["d_bv_rejected"] < st1["d_bv_computed"]
    say(f"  rejected {st1['d_bv_rejected']} of {st1['d_bv_computed']}")


def exact_rows(kind):
    """Exact-form rows of the three octet row kinds: bytes (values in
    [0, 255], d = 128: byte chains), signed (d = 128: FP32 dot form) and
    d = 96 (FP32 dot form)."""
    if kind == "u8_128":
        return integer_mixture(4000, 128, 32, seed=11)
    if kind == "s8_128":
        return integer_mixture(4000, 128, 32, seed=12) - 128.0
    if kind == "u8_100":  # dpad 104: the byte-row join's last 32-byte step is partial
        return integer_mixture(4000, 100, 32, seed=14)
    if kind == "u8_20":   # dpad 24: less than one step
        return integer_mixture(4000, 20, 32, seed=15)
    return integer_mixture(4000, 96, 32, seed=13)


@pytest.mark.parametrize("kind", ["u8_128", "s8_128", "u8_96"])
@pytest.mark.parametrize("pg", [FK_PG_NSG, FK_PG_TAUMNG, FK_PG_ALPHAPG])
@pytest.mark.parametrize("reversed_", [False, True])
def test_exact_form_prune_bitexact(device, pg, reversed_, kind):
    """With FK_FLAG_GRAM on exact-form rows the pruner takes its occlusion
    distances as |a|^2 + |b|^2 - 2 a.b (from byte rows when the values are
    in [0, 255] and d = 128); ids, kept distances and distance counts must
    equal the CPU oracle's (canonical FP32 distances)."""
    from test_gpu_parity import make_pools
    x = exact_rows(kind)
    # duplicate rows: zero distances take the 0/0 and x/0 paths of the alpha rule
    x[100:140] = x[200:240]
    say("exact prune", pg, reversed_, kind)
    device.upload(x)
    rng = np.random.default_rng(3)
    p = IndexParams(nsg_R=24, search_K=90, tau=300.0, angle=63)
    pools, plen, q = make_pools(x, 400, 120, rng)
    osz = len(x) - 50
    c = oracle.prune(pg, x, pools, plen, q, osz, p, reversed_=reversed_)
    g0 = device.prune(pg, pools, plen, q, osz, p, reversed_=reversed_)
    p.flags = GRAM
    g = device.prune(pg, pools, plen, q, osz, p, reversed_=reversed_)
    bad0 = [t for t in range(400) if g0[2][t] != c[2][t] or not np.array_equal(g0[0][t, :c[2][t]], c[0][t, :c[2][t]])]
    bad = [t for t in range(400) if g[2][t] != c[2][t] or not np.array_equal(g[0][t, :c[2][t]], c[0][t, :c[2][t]])]
    say(f"  differing tasks: plain {len(bad0)} {bad0[:5]}, exact form {len(bad)} {bad[:5]}")
    for t in bad[:2]:
        say("   task", t, "q", q[t], "len", plen[t], "oracle", c[0][t, :c[2][t]].tolist(), "gpu", g[0][t, :g[2][t]].tolist())
    assert np.array_equal(g[2], c[2])
    for t in range(400):
        k = c[2][t]
        assert np.array_equal(g[0][t, :k], c[0][t, :k]), t
        assert np.array_equal(g[1][t, :k].view(np.uint32), c[1][t, :k].view(np.uint32)), t
    assert g[3] == c[3]


@pytest.mark.parametrize("kind", ["u8_128", "s8_128", "u8_96", "u8_100", "u8_20"])
@pytest.mark.parametrize("pg", [FK_PG_ALPHAPG, FK_PG_HNSW_LAYER])
def test_gram_full_build_identical_exact_form(device, pg, kind):
    """Exact form end to end (join, prunes, BridgeView, every octet row
    kind; the byte-row join at dpad 128, 96, 104 and 24): the same graph as
    the plain FP32 build."""
    x = exact_rows(kind)[:3000]
    device.upload(x)
    p = nnd_params(angle=60)
    off0, ids0, ep0, w0, _ = device.build(pg, p)
    p.flags = GRAM
    off1, ids1, ep1, w1, st1 = device.build(pg, p)
    assert st1["gram_mode"] == _ffi.FK_GRAM_EXACT
    assert ep0 == ep1 and w0 == w1
    assert np.array_equal(off0, off1) and np.array_equal(ids0, ids1)
