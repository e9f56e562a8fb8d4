"""The error models behind BridgeView's bound pass, checked on the CPU.

The bound pass drops a target when a rigorous lower bound on its distance
exceeds the pool's last distance, so a target it drops must truly be beyond
it; an error model that under-states an error would drop pool members and
change the graph.  The GPU tests compare whole builds; these tests check the
two models directly, with the kernels' FP32 arithmetic restated in numpy
float32 on random, scaled and adversarial rows (exact and near duplicates,
rows at very different norms), against float64 distances.

8-bit rows (widths 96 and 256): q = rint(x / s) with s = max|x| / 127, the error norm
eps = |x - s q| (float64, rounded up by 1e-6) and n = |s q|^2 stored as
float32; the kernel forms dt = (n_x + n_y) - 2 (s_x s_y) (q_x . q_y) in FP32
with the integer dot exact, f = 2^-20 (n_x + n_y), and drops the target
when (dt - f)(1 - 2^-18) > (T + (eps_x + eps_y)(1 + 2^-20))^2 (1 + 2^-18).
Claim pinned here: whenever sqrt(max(dt - f, 0)) > T + eps_x + eps_y, the
float64 distance |x - y| exceeds T (the triangle inequality through the
quantised rows)."""
import numpy as np
import pytest

UP20 = np.float32(1.0 + 2.0 ** -20)
UP18 = np.float32(1.0 + 2.0 ** -18)
DN18 = np.float32(1.0 - 2.0 ** -18)
FREL = np.float32(2.0 ** -20)


def quantise8(x):
    """Per-row 8-bit copy as the row-copy kernel makes it."""
    x = x.astype(np.float32)
    m = np.max(np.abs(x), axis=1)
    s = (m / np.float32(127.0)).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.where(s[:, None] > 0, np.rint(x / s[:, None]), 0.0)
    q = np.clip(q, -127, 127).astype(np.int32)
    d = x.astype(np.float64) - s[:, None].astype(np.float64) * q
    eps = (np.sqrt(np.sum(d * d, axis=1)) * (1.0 + 1e-6)).astype(np.float32)
    n = (s.astype(np.float64) ** 2 * np.sum(q.astype(np.int64) ** 2, axis=1)).astype(np.float32)
    return q, s, eps, n


def lower_sqrt(q, s, eps, n, i, j):
    """sqrt(max(dt - f, 0)) - eps_i - eps_j with the kernel's FP32 dt and f
    (the quantity the kernel compares against T)."""
    dot = np.float32(np.dot(q[i].astype(np.int64), q[j].astype(np.int64)))
    nx = np.float32(n[i] + n[j])
    dt = np.float32(nx - np.float32(2.0) * np.float32(np.float32(s[i] * s[j]) * dot))
    f = np.float32(FREL * nx)
    return np.sqrt(max(float(dt) - float(f), 0.0)) - float(eps[i]) - float(eps[j])


def rows(kind, rng, n=600, d=96):
    if kind == "unit":
        x = rng.normal(size=(n, d))
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    elif kind == "scaled":
        x = rng.normal(size=(n, d)) * np.exp(rng.uniform(-4, 4, size=(n, 1)))
    elif kind == "sparse":
        x = rng.normal(size=(n, d)) * (rng.random((n, d)) < 0.05)
    else:  # clustered, with exact and near duplicates
        c = rng.normal(size=(8, d))
        x = c[rng.integers(0, 8, n)] + 0.01 * rng.normal(size=(n, d))
        x[: n // 10] = x[n // 10: n // 5]
        x[: n // 20] *= 1.0 + 1e-6
    return x.astype(np.float32)


@pytest.mark.parametrize("d", [96, 256])
@pytest.mark.parametrize("kind", ["unit", "scaled", "sparse", "clustered"])
def test_8bit_lower_bound_never_exceeds_the_distance(kind, d):
    rng = np.random.default_rng(97)
    x = rows(kind, rng, n=600 if d == 96 else 300, d=d)
    q, s, eps, n = quantise8(x)
    x64 = x.astype(np.float64)
    worst = -np.inf
    pairs = [(i, j) for i in range(0, len(x), 3) for j in range(len(x)) if j != i][:60000]
    for i, j in pairs:
        true = np.sqrt(np.sum((x64[i] - x64[j]) ** 2))
        lo = lower_sqrt(q, s, eps, n, i, j)
        worst = max(worst, lo - true)
        assert lo <= true * (1 + 1e-6) + 1e-12, (kind, i, j, lo, true)
    # and the bound is not vacuous: it is within a few error norms of the truth
    assert worst > -4.0 * float(np.max(eps)) - 1e-3


def test_8bit_error_norm_covers_each_row():
    rng = np.random.default_rng(98)
    for kind in ("unit", "scaled", "sparse", "clustered"):
        x = rows(kind, rng)
        q, s, eps, n = quantise8(x)
        err = np.sqrt(np.sum((x.astype(np.float64) - s[:, None].astype(np.float64) * q) ** 2, axis=1))
        assert np.all(eps.astype(np.float64) >= err)
        assert np.all(np.abs(q) <= 127)
        # |x - s q| <= s / 2 per value, so eps <= sqrt(d) s / 2 (rounded up)
        assert np.all(eps <= np.sqrt(x.shape[1]) * s / 2 * (1 + 1e-5) + 1e-30)


def test_8bit_drop_rule_implies_beyond_threshold():
    """The kernel's square-root-free drop rule, with T from the threshold as
    the kernel makes it (sqrt(thr / (1 - gam)) rounded up by 2^-20): every
    dropped pair's float64 squared distance exceeds thr / (1 - gam)."""
    rng = np.random.default_rng(99)
    x = rows("clustered", rng, n=400)
    q, s, eps, n = quantise8(x)
    x64 = x.astype(np.float64)
    gam = np.float32((96 + 16) * 2.0 ** -23)
    dropped = 0
    for i in range(0, 400, 7):
        d2 = np.sum((x64 - x64[i]) ** 2, axis=1)
        for thr in np.quantile(d2, [0.01, 0.05, 0.2]).astype(np.float32):
            T = np.float32(np.sqrt(np.float32(thr / (np.float32(1) - gam)))) * UP20
            for j in range(400):
                if j == i:
                    continue
                dot = np.float32(np.dot(q[i].astype(np.int64), q[j].astype(np.int64)))
                nx = np.float32(n[i] + n[j])
                dt = np.float32(nx - np.float32(2.0) * np.float32(np.float32(s[i] * s[j]) * dot))
                f = np.float32(FREL * nx)
                lo = np.float32(T + np.float32(np.float32(eps[i] + eps[j]) * UP20))
                if np.float32(np.float32(dt - f) * DN18) > np.float32(np.float32(lo * lo) * UP18):
                    dropped += 1
                    assert d2[j] > float(thr) / (1.0 - float(gam)), (i, j)
    assert dropped > 0
