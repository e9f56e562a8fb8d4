"""Batched builds (fk_build_batch): many independent datasets in one pass.

Bar: every set's graph (offsets, ids, entry point, width) is bit-identical to
the graph fk_build gives for that set's rows uploaded alone -- per-set random
streams, controls and stop rule, BridgeView rounds, entry point and
connectivity -- for each graph family, for sets of uneven size, for sets of
different widths zero-padded to the widest, and with the Gram-form join.
"""
import sys

import numpy as np
import pytest

from fastkcna_amd import FK_PG_ALPHAPG, FK_PG_HNSW_LAYER, FK_PG_NSG, FK_PG_TAUMNG, IndexParams, _ffi
from fastkcna_amd.datasets import integer_mixture, unit_mixture

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def params(**kw):
    p = IndexParams(K=40, L=40, S=10, R=50, iterations=5, nsg_R=16, search_L=40, search_K=40, angle=60, tau=5.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def standalone(device, x, pg, p):
    device.upload(x)
    off, ids, ep, width, st = device.build(pg, p)
    return off, ids, ep, width, st


def check_batch(device, sets, pg, p, width=None):
    d = width or max(s.shape[1] for s in sets)
    rows = np.zeros((sum(len(s) for s in sets), d), np.float32)
    set_off = np.zeros(len(sets) + 1, np.uint64)
    for k, s in enumerate(sets):
        set_off[k + 1] = set_off[k] + len(s)
        rows[set_off[k]:set_off[k + 1], :s.shape[1]] = s
    device.upload(rows)
    got, st = device.build_batch(pg, p, set_off)
    assert len(got) == len(sets)
    for k, s in enumerate(sets):
        off, ids, ep, w, _ = standalone(device, s, pg, p)
        g_off, g_ids, g_ep, g_w = got[k]
        assert g_ep == ep and g_w == w, (k, g_ep, ep, g_w, w)
        assert np.array_equal(g_off, off), k
        assert np.array_equal(g_ids, ids), k
    return st


@pytest.mark.parametrize("pg", [FK_PG_NSG, FK_PG_ALPHAPG, FK_PG_TAUMNG, FK_PG_HNSW_LAYER])
def test_batch_equals_standalone_builds(device, pg):
    rng = np.random.default_rng(7)
    sizes = [1200, 3100, 450, 2200, 1700]
    sets = [integer_mixture(n, 128, 8, seed=40 + k) for k, n in enumerate(sizes)]
    say("batch", pg, sizes)
    st = check_batch(device, sets, pg, params())
    say(f"  nnd iterations {st['nnd_iterations']}, bridgeview rounds {st['bv_rounds'][:2]}")


def test_batch_unit_rows_mixed_widths(device):
    """Sets of width 96 and 128 (C4's unit-normalised mixtures) uploaded
    padded to 128: each graph equals its own-width standalone build."""
    sets = [unit_mixture(1500, 96, 16, seed=50, sigma=0.3), unit_mixture(2600, 128, 16, seed=51, sigma=0.3),
            unit_mixture(900, 96, 16, seed=52, sigma=0.3)]
    say("batch mixed widths")
    check_batch(device, sets, FK_PG_ALPHAPG, params(angle=60))


def test_batch_with_gram_join(device):
    sets = [integer_mixture(n, 128, 8, seed=60 + k) for k, n in enumerate([2000, 800, 2500])]
    say("batch gram")
    check_batch(device, sets, FK_PG_NSG, params(flags=_ffi.FK_FLAG_GRAM))


def test_batch_rejects_bad_sets(device):
    x = integer_mixture(1000, 32, 4, seed=70)
    device.upload(x)
    with pytest.raises(_ffi.FkError):  # a set no larger than controls = 100
        device.build_batch(FK_PG_NSG, params(), np.array([0, 900, 1000], np.uint64))
    with pytest.raises(_ffi.FkError):  # beyond the resident rows
        device.build_batch(FK_PG_NSG, params(), np.array([0, 500, 1200], np.uint64))
    with pytest.raises(_ffi.FkError):
        device.build_batch(FK_PG_NSG, params(flags=_ffi.FK_FLAG_SCREEN), np.array([0, 500, 1000], np.uint64))
