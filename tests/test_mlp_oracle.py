"""Pins the CPU restatement of the warm-start network (oracle/mlp.py): the
analytic gradient against central differences, and closed forms."""

import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import mlp  # noqa: E402

SHAPE = (2, 6, 5, 3)


def _data(seed=0, n=7):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(mlp.offsets(SHAPE)["count"]) * 0.5
    x = rng.uniform(-1, 1, (n, SHAPE[0]))
    t = rng.standard_normal((n, SHAPE[3])) * 0.3
    return p, x, t


def test_gradient_matches_central_differences():
    p, x, t = _data()
    g = mlp.gradient(SHAPE, p, x, t)
    h = 1e-6
    num = np.empty_like(p)
    for k in range(p.size):
        e = np.zeros_like(p)
        e[k] = h
        num[k] = (mlp.loss(SHAPE, p + e, x, t) - mlp.loss(SHAPE, p - e, x, t)) / (2 * h)
    np.testing.assert_allclose(g, num, rtol=1e-6, atol=1e-9)


def test_a_dropped_term_is_caught():
    # a gradient without the tanh derivative of the first layer must fail
    p, x, t = _data(1)
    g = mlp.gradient(SHAPE, p, x, t)
    W1, b1, W2, b2, W3, b3 = mlp.unpack(SHAPE, p)
    y, (a1, a2) = mlp.forward(SHAPE, p, x)
    d3 = (y - t) / x.shape[0]
    d2 = (d3 @ W3) * (1 - a2 ** 2)
    d1_wrong = d2 @ W2
    assert not np.allclose((d1_wrong.T @ x).ravel(), g[:W1.size], rtol=1e-3)


def test_closed_forms():
    p, x, t = _data(2)
    off = mlp.offsets(SHAPE)
    z = p.copy()
    z[:off["W3"][0]] = 0.0  # hidden activations vanish: y = b3
    y, (a1, a2) = mlp.forward(SHAPE, z, x)
    assert np.all(a1 == 0) and np.all(a2 == 0)
    np.testing.assert_allclose(y, np.broadcast_to(z[off["b3"][0]:], y.shape))
    # the first Adam step moves every parameter with a nonzero gradient by lr·sign(g)
    g = mlp.gradient(SHAPE, p, x, t)
    p1, hist = mlp.adam(SHAPE, p, x, t, 1, 0.01, eps=0.0)
    mask = np.abs(g) > 0
    np.testing.assert_allclose((p - p1)[mask], 0.01 * np.sign(g[mask]), rtol=1e-12)
    assert hist[0] == mlp.loss(SHAPE, p, x, t)


def test_adam_decreases_the_loss():
    p, x, t = _data(3)
    _, hist = mlp.adam(SHAPE, p, x, t, 200, 0.01)
    assert hist[-1] < 0.5 * hist[0]
