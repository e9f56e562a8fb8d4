"""Device memory reuse inside one handle.

Buffers a build releases stay with the handle and are handed to later
builds with whatever the earlier build left in them.  Bar: builds of
different sizes and datasets, alternated on one handle, give graphs
bit-identical to the same builds on fresh handles (each graph's first build
on its own handle sees only fresh allocations); the graph arrays returned
without a copy stay valid after the handle builds again.
"""
import sys

import numpy as np
import pytest

from fastkcna_amd import FK_PG_NSG, FK_PG_TAUMNG, Device, IndexParams
from fastkcna_amd.datasets import integer_mixture, unit_mixture

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def params(**kw):
    p = IndexParams(K=40, L=40, S=10, R=50, iterations=5, nsg_R=16, search_L=40, search_K=40, tau=0.01)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


CASES = [
    ("int 3000x128 NSG", lambda: integer_mixture(3000, 128, 8, seed=81), FK_PG_NSG, {}),
    ("unit 5000x96 tau-MNG", lambda: unit_mixture(5000, 96, 16, seed=82, sigma=0.3), FK_PG_TAUMNG, {}),
    ("int 2000x128 NSG, L=60", lambda: integer_mixture(2000, 128, 8, seed=83), FK_PG_NSG, {"L": 60, "K": 60,
                                                                                         "search_K": 60}),
]


def fresh_build(x, pg, p):
    dev = Device(0)
    try:
        dev.upload(x)
        off, ids, ep, w, _ = dev.build(pg, p)
        return off.copy(), ids.copy(), ep, w
    finally:
        dev.close()


def test_reused_handle_matches_fresh_handles():
    ref = {}
    for name, make, pg, kw in CASES:
        say("fresh", name)
        ref[name] = fresh_build(make(), pg, params(**kw))
    dev = Device(0)
    try:
        kept = []
        for rep in range(2):
            for name, make, pg, kw in (CASES if rep == 0 else CASES[::-1]):
                say("reused handle", rep, name)
                dev.upload(make())
                off, ids, ep, w, _ = dev.build(pg, params(**kw))
                r_off, r_ids, r_ep, r_w = ref[name]
                assert ep == r_ep and w == r_w, name
                assert np.array_equal(off, r_off) and np.array_equal(ids, r_ids), name
                kept.append((name, off, ids))
        # the arrays of every earlier build still hold their graph
        for name, off, ids in kept:
            assert np.array_equal(off, ref[name][0]) and np.array_equal(ids, ref[name][1]), name
    finally:
        dev.close()
