"""Float64 numpy reference of the cost-model MLP (n_in -> H -> H -> 1, ReLU, mean
squared error), the checker of the HIP training and inference kernels.

Parameter layout (the same as the library's): W1[n_in][H], b1[H], W2[H][H], b2[H],
W3[H], b3[1], concatenated.
"""
import numpy as np


def unpack(p, n_in, H):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + n_in * H].reshape(n_in, H); o += n_in * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]; o += 1
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def forward(p, X, n_in, H):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, H)
    X = np.asarray(X, dtype=np.float64)
    h1 = np.maximum(X @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    return h2 @ W3 + b3, (X, h1, h2)


def loss_grad(p, X, y, n_in, H):
    """mean((pred - y)^2) and its gradient (same layout as p)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, H)
    pred, (X, h1, h2) = forward(p, X, n_in, H)
    n = len(pred)
    d = pred - np.asarray(y, dtype=np.float64)
    loss = float(np.mean(d * d))
    dpred = 2.0 * d / n
    gW3 = h2.T @ dpred
    gb3 = dpred.sum()
    d2 = np.outer(dpred, W3) * (h2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ W2.T) * (h1 > 0)
    gW1 = X.T @ d1
    gb1 = d1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g


def num_params(n_in, H):
    return n_in * H + H + H * H + H + H + 1
