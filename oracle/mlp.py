"""CPU restatement of the tile-cost MLP (test infrastructure only).

Same parameter layout as the library's model: one FP32 vector
[W1 (in x H), b1 (H), W2 (H x H), b2 (H), w3 (H), b3 (1)], ReLU hidden
layers, mean-squared-error loss.  Gradients are hand-derived backprop; the
tests pin them against central finite differences in float64.
"""
import numpy as np

H = 64


def layout(n_in):
    sizes = [("W1", n_in * H), ("b1", H), ("W2", H * H), ("b2", H), ("w3", H), ("b3", 1)]
    out, off = {}, 0
    for name, sz in sizes:
        out[name] = (off, sz)
        off += sz
    return out, off


def unpack(p, n_in):
    lay, _ = layout(n_in)
    g = lambda k: p[lay[k][0]:lay[k][0] + lay[k][1]]
    return (g("W1").reshape(n_in, H), g("b1"), g("W2").reshape(H, H), g("b2"), g("w3"), g("b3")[0])


def forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p, X.shape[1])
    z1 = X @ W1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0)
    return h2 @ w3 + b3, (z1, h1, z2, h2)


def loss(p, X, T):
    y, _ = forward(p, X)
    return float(np.mean((y - T) ** 2))


def gradient(p, X, T):
    n_in = X.shape[1]
    W1, b1, W2, b2, w3, b3 = unpack(p, n_in)
    y, (z1, h1, z2, h2) = forward(p, X)
    n = X.shape[0]
    dy = 2.0 * (y - T) / n
    g_w3 = h2.T @ dy
    g_b3 = dy.sum()
    d2 = np.outer(dy, w3) * (z2 > 0)
    g_W2 = h1.T @ d2
    g_b2 = d2.sum(axis=0)
    d1 = (d2 @ W2.T) * (z1 > 0)
    g_W1 = X.T @ d1
    g_b1 = d1.sum(axis=0)
    return np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3]]), float(np.mean((y - T) ** 2))
