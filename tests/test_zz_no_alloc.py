"""fk_prune_noalloc (the prune with its buffers in a static device scratch
area, no device allocation) against fk_prune through the C ABI: identical
ids, distances, lengths and distance counts, on two calls in a row (the
second smaller, so stale contents of the shared scratch from the first call
would show), for the NSG, tau-MNG and alpha-PG rules and the reversed form,
with the octet prune and with the Gram-form prune (FK_FLAG_GRAM, prune
variant 3: its 16-bit rows prepared by the fk_prune call before).
Also: a call above the scratch maximum fails with FK_E_INVALID."""
import sys

import numpy as np
import pytest

import oracle
from fastkcna_amd import FK_PG_ALPHAPG, FK_PG_NSG, FK_PG_TAUMNG, FkError, IndexParams, _ffi
from fastkcna_amd.datasets import integer_mixture, unit_mixture

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def pools_for(x, n_tasks, L, seed):
    rng = np.random.default_rng(seed)
    n = len(x)
    q = rng.choice(n, n_tasks, replace=False).astype(np.uint32)
    pools = np.zeros((n_tasks, L), np.uint64)
    plen = np.zeros(n_tasks, np.uint32)
    for t, qq in enumerate(q):
        cand = rng.choice(n, L, replace=False)
        d = oracle.pair_distances(x, np.full(L, qq), cand)
        order = np.lexsort((cand, d))
        pools[t] = oracle.pack(d[order], cand[order])
        plen[t] = rng.integers(1, L + 1)
    return pools, plen, q


def same(a, b):
    ids_a, d_a, len_a, cc_a = a
    ids_b, d_b, len_b, cc_b = b
    assert np.array_equal(len_a, len_b)
    for t in range(len(len_a)):
        k = len_a[t]
        assert np.array_equal(ids_a[t, :k], ids_b[t, :k]), t
        assert np.array_equal(d_a[t, :k].view(np.uint32), d_b[t, :k].view(np.uint32)), t
    assert cc_a == cc_b


@pytest.mark.parametrize("gram", [False, True])
@pytest.mark.parametrize("pg", [FK_PG_NSG, FK_PG_TAUMNG, FK_PG_ALPHAPG])
@pytest.mark.parametrize("reversed_", [False, True])
@pytest.mark.parametrize("name", ["int128", "unit96"])
def test_noalloc_matches_fk_prune(device, pg, reversed_, name, gram):
    x = integer_mixture(3000, 128, 16, seed=21) if name == "int128" else unit_mixture(2500, 96, 8, seed=22,
                                                                                      sigma=0.25)
    say("noalloc", pg, reversed_, name, "gram" if gram else "octet")
    device.upload(x)
    tau = 300.0 if name == "int128" else 0.05
    p = IndexParams(nsg_R=24, search_K=100, tau=tau, angle=63)
    if gram:
        p.flags = _ffi.FK_FLAG_GRAM | _ffi.FK_FLAG_PRUNE_V3
    big = pools_for(x, 600, 120, seed=5)
    small = pools_for(x, 150, 80, seed=6)
    for pools, plen, q in (big, small):
        ref = device.prune(pg, pools, plen, q, len(x) - 40, p, reversed_=reversed_)
        got = device.prune(pg, pools, plen, q, len(x) - 40, p, reversed_=reversed_, noalloc=True)
        same(got, ref)


def test_noalloc_rejects_calls_above_the_maximum(device):
    x = integer_mixture(2000, 128, 16, seed=23)
    say("noalloc limit")
    device.upload(x)
    p = IndexParams(nsg_R=16, search_K=64)
    n_tasks, stride = 70_000, 128  # 70,000 * 128 * 8 bytes of pools alone exceed 64 MiB
    pools = np.zeros((n_tasks, stride), np.uint64)
    pools[:, 0] = oracle.pack(np.float32([1.0]), [1])[0]
    plen = np.ones(n_tasks, np.uint32)
    q = np.zeros(n_tasks, np.uint32)
    with pytest.raises(FkError) as e:
        device.prune(FK_PG_NSG, pools, plen, q, len(x), p, noalloc=True)
    assert "FK_E_INVALID" in str(e.value) and "maximum" in str(e.value)
    # the handle still works afterwards
    small = pools_for(x, 50, 64, seed=7)
    same(device.prune(FK_PG_NSG, *small, len(x), p, noalloc=True), device.prune(FK_PG_NSG, *small, len(x), p))
