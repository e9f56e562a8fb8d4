"""Float64 NumPy restatement of the learned distance screen (test helper).

Same model and loss as the GPU kernels: z = P (x_a - x_b), u = z^2 / mu,
h = relu(W1 u + b1), y = w2 . h + b2, loss = mean ((y - t) / t)^2 with
t = max(dist / mu, 0.01).  The
weight vector layout is P^T [dpad][32] | W1 [64][32] | b1 [64] | w2 [64] | b2.
Its analytic gradient is checked against central finite differences in
test_screen_model.py, so it pins the GPU gradient independently of the
formulas written here.
"""
import numpy as np

Z, H = 32, 64


def dpad_of(d):
    return (d + 7) // 8 * 8


def unpack(w, dpad):
    w = np.asarray(w, np.float64)
    o = 0
    PT = w[o:o + dpad * Z].reshape(dpad, Z); o += dpad * Z
    W1 = w[o:o + H * Z].reshape(H, Z); o += H * Z
    b1 = w[o:o + H]; o += H
    w2 = w[o:o + H]; o += H
    b2 = w[o]
    return PT, W1, b1, w2, b2


def n_params(d):
    return dpad_of(d) * Z + H * Z + 2 * H + 1


def init_weights(d, rng):
    dp = dpad_of(d)
    w = np.zeros(n_params(d))
    w[:dp * Z] = rng.normal(0, 1 / np.sqrt(Z), dp * Z)
    w[dp * Z:dp * Z + H * Z] = rng.uniform(0, 2 / Z, H * Z)
    w[dp * Z + H * Z + H:dp * Z + H * Z + 2 * H] = 0.5
    return w


def _delta(x, a, b):
    d = x.shape[1]
    dp = dpad_of(d)
    delta = np.zeros((len(a), dp))
    delta[:, :d] = x[a].astype(np.float64) - x[b].astype(np.float64)
    return delta


def forward(w, x, a, b, inv_mu):
    dp = dpad_of(x.shape[1])
    PT, W1, b1, w2, b2 = unpack(w, dp)
    z = _delta(x, a, b) @ PT
    u = z * z * inv_mu
    hp = u @ W1.T + b1
    h = np.maximum(hp, 0)
    return h @ w2 + b2


def loss_and_grad(w, x, a, b, dist, inv_mu):
    dp = dpad_of(x.shape[1])
    PT, W1, b1, w2, b2 = unpack(w, dp)
    delta = _delta(x, a, b)
    z = delta @ PT
    u = z * z * inv_mu
    hp = u @ W1.T + b1
    h = np.maximum(hp, 0)
    y = h @ w2 + b2
    t = np.maximum(np.asarray(dist, np.float64) * inv_mu, 1e-2)
    r = (y - t) / t
    n = len(a)
    loss = np.mean(r * r)
    dy = 2 * r / (t * n)
    g_w2 = h.T @ dy
    g_b2 = dy.sum()
    dh = np.outer(dy, w2) * (hp > 0)
    g_W1 = dh.T @ u
    g_b1 = dh.sum(0)
    du = dh @ W1
    dz = du * 2 * z * inv_mu
    g_PT = delta.T @ dz
    return loss, np.concatenate([g_PT.ravel(), g_W1.ravel(), g_b1, g_w2, [g_b2]])
