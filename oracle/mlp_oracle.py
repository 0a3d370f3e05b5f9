"""CPU oracle of the node-growth predictor's network (test infrastructure only).

Float64 numpy restatement of the FP32 MLP 16 -> 64 -> 64 -> 1 (ReLU) whose HIP
kernels live in the library: forward, the gradient of the mean squared error and
one Adam step.  Used only by tests to check the device kernels; nothing in the
product imports it.  Pinned in tests by central finite differences of the loss and
by closed forms (a network whose ReLUs are all active is an affine map).

Parameter layout (flat): W1[64][16] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3.
"""
import numpy as np

N_IN, N_HID = 16, 64
SIZES = [("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
         ("W3", (N_HID,)), ("b3", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SIZES])


def forward(p, x):
    """y [n] and the hidden pre-activations."""
    w = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    a1 = x @ w["W1"].T + w["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(a2, 0.0)
    y = h2 @ w["W3"] + w["b3"][0]
    return y, (x, a1, h1, a2, h2)


def loss(p, x, t):
    y, _ = forward(p, x)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def gradient(p, x, t):
    """(dL/dp flat, y, L) for L = mean((y - t)^2)."""
    w = unpack(p)
    y, (x, a1, h1, a2, h2) = forward(p, x)
    n = len(y)
    dy = 2.0 * (y - np.asarray(t, dtype=np.float64)) / n
    g = {}
    g["W3"] = dy @ h2
    g["b3"] = np.array([dy.sum()])
    dh2 = np.outer(dy, w["W3"]) * (a2 > 0)
    g["W2"] = dh2.T @ h1
    g["b2"] = dh2.sum(axis=0)
    dh1 = (dh2 @ w["W2"]) * (a1 > 0)
    g["W1"] = dh1.T @ x
    g["b1"] = dh1.sum(axis=0)
    return pack(g), y, float(np.mean((y - t) ** 2))


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
