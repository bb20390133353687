"""numpy restatement of the initial-guess network of the learned solve (test
infrastructure: the GPU forward and backward kernels are checked against it).

Network: six 3 x 3 x 3 convolutions, 3 -> 32 -> 32 -> 32 -> 32 -> 32 -> 1
channels, GELU (0.5 x (1 + erf(x / sqrt 2))) after the first five, bias on
every layer.  Fields are periodic in x and y and padded with zeros beyond the
walls in z, like the pressure operator.  Arrays are (C, nx, ny, nz); weights
are (27, Cout, Cin) with tap t = (dx + 1) + 3 (dy + 1) + 9 (dz + 1) weighing
the input at (i + dx, j + dy, k + dz).
"""
import math

import numpy as np
from scipy.special import erf

CHANNELS = (3, 32, 32, 32, 32, 32, 1)
TAPS = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def shift(a, dx, dy, dz):
    """a at (i + dx, j + dy, k + dz): periodic in x, y, zero beyond the z walls."""
    out = np.roll(a, (-dx, -dy), axis=(1, 2))
    if dz:
        z = np.zeros_like(out)
        if dz > 0:
            z[..., :-dz] = out[..., dz:]
        else:
            z[..., -dz:] = out[..., :dz]
        out = z
    return out


def conv(x, w, b):
    """y[co] = b[co] + sum_t sum_ci w[t, co, ci] x[ci](p + tap_t)."""
    cout = w.shape[1]
    y = np.zeros((cout,) + x.shape[1:], dtype=x.dtype)
    for t, (dx, dy, dz) in enumerate(TAPS):
        xs = shift(x, dx, dy, dz)
        y += np.tensordot(w[t], xs, axes=(1, 0))
    return y + b.reshape(-1, 1, 1, 1)


def conv_backward(x, w, dy):
    """Gradients of conv with respect to x, w and b for the output gradient dy."""
    dx_ = np.zeros_like(x)
    dw = np.zeros_like(w)
    for t, (sx, sy, sz) in enumerate(TAPS):
        xs = shift(x, sx, sy, sz)
        dw[t] = np.tensordot(dy, xs, axes=([1, 2, 3], [1, 2, 3]))
        # the input at p + tap received w[t] dy(p): move dy back by the tap
        dx_ += shift(np.tensordot(w[t].T, dy, axes=(1, 0)), -sx, -sy, -sz)
    db = dy.sum(axis=(1, 2, 3))
    return dx_, dw, db


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def init_params(seed=0, dtype=np.float64):
    rng = np.random.default_rng(seed)
    params = []
    for cin, cout in zip(CHANNELS[:-1], CHANNELS[1:]):
        scale = math.sqrt(2.0 / (27 * cin))
        params.append((rng.standard_normal((27, cout, cin)).astype(dtype) * scale, np.zeros(cout, dtype=dtype)))
    return params


def forward(params, x):
    """Returns the output and the pre-activations of every layer."""
    pre = []
    a = x
    for n, (w, b) in enumerate(params):
        z = conv(a, w, b)
        pre.append((a, z))
        a = gelu(z) if n < len(params) - 1 else z
    return a, pre


def backward(params, pre, dout):
    """Gradients of all weights and biases for the output gradient dout."""
    grads = [None] * len(params)
    d = dout
    for n in range(len(params) - 1, -1, -1):
        a, z = pre[n]
        if n < len(params) - 1:
            d = d * gelu_grad(z)
        da, dw, db = conv_backward(a, params[n][0], d)
        grads[n] = (dw, db)
        d = da
    return grads


def loss_and_grads(params, x, target):
    """Mean squared error over the voxels and its gradients."""
    out, pre = forward(params, x)
    diff = out - target
    n = diff.size
    loss = float(np.sum(diff * diff)) / n
    grads = backward(params, pre, 2.0 * diff / n)
    return loss, grads
