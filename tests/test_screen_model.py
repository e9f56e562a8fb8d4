"""The float64 screen model used to check the GPU training kernels: its
analytic gradient against central finite differences (CPU only)."""
import numpy as np

from screen_ref import init_weights, loss_and_grad, n_params


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(40, 20))  # dpad 24: padding rows of P^T get zero gradient
    a = rng.integers(0, 40, 60)
    b = rng.integers(0, 40, 60)
    dist = np.sum((x[a] - x[b]) ** 2, axis=1)
    inv_mu = 1.0 / dist.mean()
    w = init_weights(20, rng)
    w[-129:-65] = rng.normal(0, 0.05, 64)  # b1: some units inactive
    w[-65:-1] += rng.normal(0, 0.6, 64)    # w2 of mixed sign
    _, g = loss_and_grad(w, x, a, b, dist, inv_mu)
    assert len(g) == n_params(20)
    idx = np.concatenate([rng.choice(24 * 32, 40, replace=False),           # P^T, padding included
                          24 * 32 + rng.choice(64 * 32 + 129, 60, replace=False)])
    eps = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        fd = (loss_and_grad(wp, x, a, b, dist, inv_mu)[0] - loss_and_grad(wm, x, a, b, dist, inv_mu)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    # padding dimensions of P^T never receive gradient
    assert np.all(g[20 * 32:24 * 32] == 0)


def test_initial_model_is_the_projection_estimate():
    """At initialisation (W1 >= 0, w2 = 1/2, b = 0) the network is a positive
    multiple of the projected squared distance, so its ranking of pairs is the
    random-projection estimate's."""
    from screen_ref import forward, unpack
    rng = np.random.default_rng(1)
    x = rng.normal(size=(50, 16))
    a, b = rng.integers(0, 50, 200), rng.integers(0, 50, 200)
    w = init_weights(16, rng)
    y = forward(w, x, a, b, 1.0)
    PT = unpack(w, 16)[0]
    z = (x[a] - x[b]) @ PT
    u = z * z
    W1 = unpack(w, 16)[1]
    assert np.allclose(y, 0.5 * (u @ W1.T).sum(1))
