"""CPU restatement of the learned-warm-start network (test infrastructure only:
imported by tests/ to check the GPU kernels; never by the product).

The network has no counterpart in the reference (its scan reuse is the
balanced subspace start, src/solve/mod.rs:708-727 and :2419-2444), so this file
restates the library's own definition: in → h1 (tanh) → h2 (tanh) → out
(linear), parameters packed as W1[h1][in], b1, W2[h2][h1], b2, W3[out][h2], b3,
loss L = 1/(2n) Σ_s ‖y_s − t_s‖², full-batch Adam (β1 = 0.9, β2 = 0.999,
ε = 1e-8, bias-corrected). Pinned by central differences of `loss` in
tests/test_mlp_oracle.py (the gradient) and by closed forms (tanh(0) = 0 layers,
a single Adam step moves each parameter by -lr·sign(g)).
"""

import numpy as np


def offsets(shape):
    i, h1, h2, o = shape
    off = {}
    k = 0
    for name, size in (("W1", h1 * i), ("b1", h1), ("W2", h2 * h1), ("b2", h2), ("W3", o * h2), ("b3", o)):
        off[name] = (k, size)
        k += size
    off["count"] = k
    return off


def unpack(shape, p):
    i, h1, h2, o = shape
    off = offsets(shape)
    get = lambda n: p[off[n][0]:off[n][0] + off[n][1]]
    return (get("W1").reshape(h1, i), get("b1"), get("W2").reshape(h2, h1), get("b2"), get("W3").reshape(o, h2),
            get("b3"))


def forward(shape, p, x):
    W1, b1, W2, b2, W3, b3 = unpack(shape, np.asarray(p, dtype=np.float64))
    a1 = np.tanh(x @ W1.T + b1)
    a2 = np.tanh(a1 @ W2.T + b2)
    return a2 @ W3.T + b3, (a1, a2)


def loss(shape, p, x, t):
    y, _ = forward(shape, p, x)
    return 0.5 * np.sum((y - t) ** 2) / x.shape[0]


def gradient(shape, p, x, t):
    """Analytic gradient of `loss` (backpropagation), float64."""
    n = x.shape[0]
    W1, b1, W2, b2, W3, b3 = unpack(shape, np.asarray(p, dtype=np.float64))
    y, (a1, a2) = forward(shape, p, x)
    d3 = (y - t) / n
    d2 = (d3 @ W3) * (1 - a2 ** 2)
    d1 = (d2 @ W2) * (1 - a1 ** 2)
    g = [d1.T @ x, d1.sum(0), d2.T @ a1, d2.sum(0), d3.T @ a2, d3.sum(0)]
    return np.concatenate([a.ravel() for a in g])


def adam(shape, p, x, t, epochs, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """Full-batch Adam; returns (parameters, loss of every epoch before its update)."""
    p = np.asarray(p, dtype=np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    hist = []
    for ep in range(1, epochs + 1):
        hist.append(loss(shape, p, x, t))
        g = gradient(shape, p, x, t)
        m = beta1 * m + (1 - beta1) * g
        v = beta2 * v + (1 - beta2) * g * g
        p -= lr * (m / (1 - beta1 ** ep)) / (np.sqrt(v / (1 - beta2 ** ep)) + eps)
    return p, np.array(hist)
