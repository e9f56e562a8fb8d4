"""The query evaluator and the entry point against the oracle through the C
ABI.

fk_search (the batched beam search every recall figure uses) against the
oracle's CPU restatement of the same best-first loop (itself pinned by
tests/test_oracle_steps.py): identical result ids and distance counts on GPU
graphs of several datasets and beam widths.  ComputeEp on a supplied state
from a given start node: the same entry point as the oracle (float centroid
in row order, get_neighbors).  And on a full HNSW-layer build: no
out-neighbour of the entry point is closer to the centroid than it.
"""
import sys

import numpy as np
import pytest

import oracle
from fastkcna_amd import FK_PG_HNSW_LAYER, FK_PG_NSG, IndexParams
from fastkcna_amd.datasets import float_mixture, integer_mixture, unit_mixture
from test_gpu_steps import start_state, params

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


GRAPH_DATA = {
    "int128": lambda: integer_mixture(4000, 128, 16, seed=61, sigma=30.0),
    "unit96": lambda: unit_mixture(3000, 96, 8, seed=62, sigma=0.2),
    "float960": lambda: float_mixture(800, 960, 4, seed=63, sigma=0.2),
    "odd20": lambda: np.random.default_rng(64).normal(size=(2000, 20)).astype(np.float32),
}


@pytest.mark.parametrize("name", list(GRAPH_DATA))
def test_search_matches_oracle(device, name):
    x = GRAPH_DATA[name]()
    say("search", name, x.shape)
    device.upload(x)
    p = IndexParams(K=40, L=40, S=10, R=40, iterations=5, nsg_R=16, search_L=40, search_K=40)
    off, ids, ep, _, _ = device.build(FK_PG_NSG, p)
    rng = np.random.default_rng(65)
    qs = (x[rng.choice(len(x), 300, replace=False)]
          + rng.normal(0, 0.1 * float(x.std()), size=(300, x.shape[1]))).astype(np.float32)
    for L in (10, 40, 120):
        g, gn = device.search(off, ids, ep, qs, L, 10)
        c, cn = oracle.search(x, off, ids, ep, qs, L, 10)
        assert np.array_equal(g, c), (name, L)
        assert gn == cn, (name, L, gn, cn)
    say("search", name, "distances at L 120", cn)


@pytest.mark.parametrize("name", ["unit96", "float960", "odd20"])
def test_entry_point_step_matches_oracle(device, name):
    x, s = start_state(device, name)
    p = params(name)
    oracle.step_link(FK_PG_NSG, x, p, s)
    assert (s.pool_len >= p.search_L).all()
    rng = np.random.default_rng(66)
    for init in rng.choice(s.n, 12, replace=False):
        g = device.step_entry_point(p, s, int(init))
        c = oracle.step_entry_point(x, p, s, int(init))
        assert g == c, (name, init, g, c)


def test_build_entry_point_is_a_local_minimum(device):
    """ComputeEp's search expands its best result, so no out-neighbour of ep
    is closer to the centroid (float row-order sums over all rows, divided by
    the row count).  HNSW-layer builds add no connectivity edges after it."""
    x = unit_mixture(5000, 96, 8, seed=67, sigma=0.2)
    device.upload(x)
    center = (np.cumsum(x, axis=0, dtype=np.float32)[-1] / np.float32(len(x))).astype(np.float32)
    p = IndexParams(K=40, L=40, S=10, R=40, iterations=5, nsg_R=16, search_L=40, search_K=40)
    off, ids, ep, _, _ = device.build(FK_PG_HNSW_LAYER, p)
    nb = ids[off[ep]:off[ep + 1]]
    assert len(nb) > 0
    dep = oracle.l2(center, x[ep])
    for v in nb:
        assert dep <= oracle.l2(center, x[v]), v
