"""Synthetic workloads of the benchmark configurations (NumPy, seeded PCG64).

C1  100,000 x 128, 64 centres ~ U[0,128]^128, x = clip(round(c + N(0, 20^2)), 0, 255)
C2  1,000,000 x 128, same law with 1,024 centres
C3  1,000,000 x 960, 256 centres ~ U[0,1]^960, sigma 0.05, clamped at 0
C4  256 sets, n log-uniform in [2e3, 2e5], d in {96, 128, 256}, 16 unit centres,
    sigma 0.3, unit-normalised
C5  10,000,000 x 96, 4,096 unit centres, sigma 0.25, unit-normalised

Queries are drawn from the same mixture after the base rows.  Generation is
chunked so that peak host memory stays near the output size.
"""
import numpy as np


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


def float_mixture(n, d, n_centers, seed, sigma=0.05, chunk=1 << 16):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0.0, 1.0, size=(n_centers, d))
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        lab = rng.integers(0, n_centers, size=m)
        x = centers[lab] + rng.normal(0.0, sigma, size=(m, d))
        np.maximum(x, 0.0, out=x)
        out[s:s + m] = x
    return out


def unit_mixture(n, d, n_centers, seed, sigma, chunk=1 << 16):
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(n_centers, d))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        lab = rng.integers(0, n_centers, size=m)
        x = c[lab] + rng.normal(0.0, sigma, size=(m, d))
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        out[s:s + m] = x
    return out


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
