"""ORACLE for the frontier-size predictor — test infrastructure only.

The 264 -> 64 -> 64 -> 1 ReLU MLP and the gradient of its mean squared error,
in float64 numpy, written out from the chain rule.  Pinned in the tests by
central finite differences; the HIP kernels (ngap_amd/csrc/mlp.hip) are
checked against it.
"""
import numpy as np

IN, HID = 264, 64


def split(p):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for shape in ((IN, HID), (HID,), (HID, HID), (HID,), (HID,), (1,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    return out


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = split(p)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    return h2 @ W3 + b3[0], (x, h1, h2)


def loss(p, x, t):
    y, _ = forward(p, x)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def gradient(p, x, t):
    W1, b1, W2, b2, W3, b3 = split(p)
    y, (x, h1, h2) = forward(p, x)
    n = x.shape[0]
    dy = 2.0 * (y - np.asarray(t, np.float64)) / n          # dL/dy
    gW3 = h2.T @ dy
    gb3 = np.array([dy.sum()])
    d2 = np.outer(dy, W3) * (h2 > 0)                         # dL/d(pre-activation 2)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (h1 > 0)
    gW1 = x.T @ d1
    gb1 = d1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, gb3]), \
        float(np.mean((y - t) ** 2))
