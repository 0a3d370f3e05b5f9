"""Pins the float64 oracle of the predictor network (CPU): gradients against central
finite differences, the forward pass against the affine map it reduces to when every
ReLU is active, and Adam's first step against its closed form."""
import numpy as np

from oracle import mlp_oracle as M


def _random(seed, n=37, scale=0.5):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, scale, M.N_PARAMS)
    x = rng.normal(0, 1, (n, M.N_IN))
    t = rng.normal(0, 1, n)
    return p, x, t


def test_param_count():
    assert M.N_PARAMS == 64 * 16 + 64 + 64 * 64 + 64 + 64 + 1 == 5313


def test_gradient_matches_finite_differences():
    p, x, t = _random(1)
    g, _, _ = M.gradient(p, x, t)
    rng = np.random.default_rng(2)
    # every bias and a sample of weights of every tensor
    idx = list(range(1024, 1088)) + list(range(5184, 5313))
    idx += list(rng.choice(1024, 40, replace=False)) + list(1088 + rng.choice(4096, 60, replace=False))
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (M.loss(p + e, x, t) - M.loss(p - e, x, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_planted_mistake_is_caught():
    """A transposed W2 in the backward pass would fail the check above."""
    p, x, t = _random(3)
    w = M.unpack(p)
    y, (x_, a1, h1, a2, h2) = M.forward(p, x)
    dy = 2.0 * (y - t) / len(y)
    dh2 = np.outer(dy, w["W3"]) * (a2 > 0)
    wrong_dh1 = (dh2 @ w["W2"].T) * (a1 > 0)
    right = M.gradient(p, x, t)[0][:1024].reshape(64, 16)
    assert not np.allclose(wrong_dh1.T @ x_, right)


def test_all_active_network_is_affine():
    rng = np.random.default_rng(4)
    d = {"W1": rng.uniform(0.01, 0.1, (64, 16)), "b1": rng.uniform(1, 2, 64),
         "W2": rng.uniform(0.01, 0.1, (64, 64)), "b2": rng.uniform(1, 2, 64),
         "W3": rng.normal(0, 1, 64), "b3": np.array([0.7])}
    p = M.pack(d)
    x = rng.uniform(0, 1, (9, 16))
    y, _ = M.forward(p, x)
    A = d["W3"] @ d["W2"] @ d["W1"]
    c = d["W3"] @ (d["W2"] @ d["b1"] + d["b2"]) + d["b3"][0]
    assert np.allclose(y, x @ A + c, rtol=1e-12, atol=1e-12)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(5)
    p = rng.normal(0, 1, 50)
    g = rng.normal(0, 1, 50)
    p1, m, v = M.adam(p, g, np.zeros(50), np.zeros(50), 1, 0.01)
    assert np.allclose(p1, p - 0.01 * g / (np.abs(g) + 1e-8))
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)
