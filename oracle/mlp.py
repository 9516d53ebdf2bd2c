"""CPU restatement of the learned planner's network -- TEST INFRASTRUCTURE ONLY.

The planner's cost model is a 3-layer MLP, in -> H -> H -> 1, ReLU after the
two hidden layers, loss L = 1/(2B) * sum_r (out_r - t_r)^2, Adam
(beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected).  Parameters are one flat
vector in the order W1[H][in], b1[H], W2[H][H], b2[H], W3[1][H], b3[1].
The reference has no learned planner, so this module follows no reference
file; it is pinned in the tests by central finite differences of the loss
(float64) and by Adam's closed form on its first step.

Only tests/ use this module, as the checker of the HIP kernels.
"""
import numpy as np


def param_count(in_features, hidden):
    return hidden * in_features + hidden + hidden * hidden + hidden + hidden + 1


def unpack(p, in_features, hidden):
    p = np.asarray(p, dtype=np.float64)
    H, n = hidden, in_features
    o = 0
    W1 = p[o:o + H * n].reshape(H, n); o += H * n
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H].reshape(1, H); o += H
    b3 = p[o:o + 1]; o += 1
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def forward(p, X, in_features, hidden):
    """Returns (out[B], cache) in float64."""
    W1, b1, W2, b2, W3, b3 = unpack(p, in_features, hidden)
    X = np.asarray(X, dtype=np.float64)
    h1 = np.maximum(X @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    out = (h2 @ W3.T + b3)[:, 0]
    return out, (X, h1, h2)


def loss(p, X, t, in_features, hidden):
    out, _ = forward(p, X, in_features, hidden)
    e = out - np.asarray(t, dtype=np.float64)
    return float(e @ e / (2.0 * len(e)))


def gradients(p, X, t, in_features, hidden):
    """Backpropagation by hand: returns (flat gradient, loss) in float64."""
    W1, b1, W2, b2, W3, b3 = unpack(p, in_features, hidden)
    out, (X, h1, h2) = forward(p, X, in_features, hidden)
    B = len(out)
    e = out - np.asarray(t, dtype=np.float64)
    dout = (e / B)[:, None]                   # [B,1]
    gW3 = dout.T @ h2                         # [1,H]
    gb3 = dout.sum(0)
    d2 = (dout @ W3) * (h2 > 0)               # [B,H]
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (h1 > 0)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return g, float(e @ e / (2.0 * B))


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = 1 - b1 ** step
    c2 = 1 - b2 ** step
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
