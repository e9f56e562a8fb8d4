"""Synthetic workloads of the benchmark configurations (NumPy, seeded PCG64).

C1  100,000 x 128, 64 centres ~ U[0,128]^128, x = clip(round(c + N(0, 20^2)), 0, 255)
C2  1,000,000 x 128, same law with 1,024 centres
C3  1,000,000 x 960, 256 centres ~ U[0,1]^960, sigma 0.05, clamped at 0
C4  256 sets, n log-uniform in [2e3, 2e5], d in {96, 128, 256}, 16 unit centres,
    sigma 0.3, unit-normalised
C5  10,000,000 x 96, 4,096 unit centres, sigma 0.25, unit-normalised

Queries are drawn from the same mixture after the base rows.  Generation is
chunked so that peak host memory stays near the output size.  The float
mixtures (C3-C5) are drawn in float32, chunk k of FLOAT_CHUNK rows from its
own stream SeedSequence(seed, spawn_key=(k,)) (centres from the stream of
`seed` itself), so the rows do not depend on how many threads fill them.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

FLOAT_CHUNK = 1 << 14


def _chunked_rows(n, d, seed, make_chunk, chunk=FLOAT_CHUNK):
    """n x d float32 rows; chunk k (rows [k chunk, ...)) is make_chunk(rng_k,
    out_view) with rng_k seeded by SeedSequence(seed, spawn_key=(k,))."""
    out = np.empty((n, d), dtype=np.float32)
    starts = list(range(0, n, chunk))

    def fill(k):
        s = starts[k]
        make_chunk(np.random.default_rng(np.random.SeedSequence(seed, spawn_key=(k,))), out[s:s + chunk])

    workers = max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(fill, range(len(starts))))
    return out


def integer_mixture(n, d, n_centers, seed, sigma=20.0, lo=0.0, hi=128.0, chunk=1 << 18):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(lo, hi, size=(n_centers, d))
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        lab = rng.integers(0, n_centers, size=m)
        x = centers[lab] + rng.normal(0.0, sigma, size=(m, d))
        np.clip(np.rint(x), 0, 255, out=x)
        out[s:s + m] = x
    return out


def float_mixture(n, d, n_centers, seed, sigma=0.05):
    """Mixture of n_centers isotropic Gaussians (sigma) with centres uniform
    in [0, 1]^d, clamped at 0."""
    centers = np.random.default_rng(seed).uniform(0.0, 1.0, size=(n_centers, d)).astype(np.float32)

    def make(rng, o):
        lab = rng.integers(0, n_centers, size=len(o))
        rng.standard_normal(out=o, dtype=np.float32)
        o *= np.float32(sigma)
        o += centers[lab]
        np.maximum(o, 0.0, out=o)

    return _chunked_rows(n, d, seed, make)


def unit_mixture(n, d, n_centers, seed, sigma):
    """Unit-normalised mixture: centres N(0, I) normalised, rows c + N(0,
    sigma^2 I) normalised."""
    c = np.random.default_rng(seed).normal(size=(n_centers, d))
    c = (c / np.linalg.norm(c, axis=1, keepdims=True)).astype(np.float32)

    def make(rng, o):
        lab = rng.integers(0, n_centers, size=len(o))
        rng.standard_normal(out=o, dtype=np.float32)
        o *= np.float32(sigma)
        o += c[lab]
        o /= np.linalg.norm(o, axis=1, keepdims=True)

    return _chunked_rows(n, d, seed, make)


def c1(n=100_000, n_queries=1_000):
    x = integer_mixture(n + n_queries, 128, 64, seed=0)
    return x[:n], x[n:]


def c2(n=1_000_000, n_queries=10_000):
    x = integer_mixture(n + n_queries, 128, 1024, seed=1)
    return x[:n], x[n:]


def c3(n=1_000_000, n_queries=10_000):
    x = float_mixture(n + n_queries, 960, 256, seed=2)
    return x[:n], x[n:]


def c5(n=10_000_000, n_queries=10_000):
    x = unit_mixture(n + n_queries, 96, 4096, seed=4, sigma=0.25)
    return x[:n], x[n:]


def c4_sets(n_sets=256, n_queries=1_000, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_sets):
        n = int(np.floor(np.exp(rng.uniform(np.log(2_000), np.log(200_000)))))
        d = int(rng.choice([96, 128, 256]))
        x = unit_mixture(n + n_queries, d, 16, seed=seed * 100_003 + s, sigma=0.3)
        out.append((x[:n], x[n:]))
    return out


# Build parameters of each configuration (reference CLI flags).
def params_c1(IndexParams):
    return IndexParams(K=200, L=200, S=10, R=100, iterations=6, nsg_R=40, search_L=40, search_K=200, angle=60)


def params_c2(IndexParams):
    return IndexParams(K=500, L=500, S=12, R=100, iterations=6, nsg_R=50, search_L=80, search_K=500, angle=60)


def params_c3(IndexParams):
    return IndexParams(K=500, L=500, S=12, R=100, iterations=6, nsg_R=50, search_L=80, search_K=500, tau=10)


def params_c4(IndexParams):
    # "alpha = 1.2" read as the reference's --angle in degrees (its only
    # alpha-PG parameter); search_K = K as for the other configurations
    return IndexParams(K=100, L=100, S=10, R=50, iterations=6, nsg_R=32, search_L=40, search_K=100, angle=1.2)


def params_c5(IndexParams, level0=True):
    # fasthnsw -K 200 -L 200 -S 10 -R 100 -I 6 --nsg_R 32: level 0 is built
    # with nsg_R doubled; search_K = K as for the other configurations
    return IndexParams(K=200, L=200, S=10, R=100, iterations=6, nsg_R=64 if level0 else 32, search_L=40,
                       search_K=200, angle=60)

