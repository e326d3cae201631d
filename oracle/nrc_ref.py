"""CPU restatement of the radiance-cache MLP (test infrastructure only).

It mirrors the arithmetic of bulbit_amd/csrc/nrc.hip: the 74-value input
encoding padded to 80, bf16 rounding (round to nearest even) of the inputs,
weights and every layer's activations and back-propagated gradients, fp32
products, ReLU hidden layers, a linear 3-channel output, the relative L2
loss (y - t)^2 / (y^2 + 0.01) averaged over rows and channels, and Adam
(beta1 0.9, beta2 0.99, eps 1e-8).  The cache has no counterpart in Bulbit;
`forward64`/`loss64` are the same network in float64 without rounding, for
finite-difference checks of the gradient formulas.

Two shapes are built, 4 hidden layers of width 64 and 5 of width 128; every
function that takes flat weights reads the shape from their count.
"""
from __future__ import annotations

import numpy as np

K_IN, WIDTH, HIDDEN, OUT_COLS = 80, 64, 4, 16
SHAPES = ((64, 4), (128, 5))  # (hidden width, hidden layers)


def layer_shapes(width=WIDTH, hidden=HIDDEN):
    return [(K_IN if l == 0 else width, OUT_COLS if l == hidden else width) for l in range(hidden + 1)]


def param_count(width=WIDTH, hidden=HIDDEN):
    return sum(k * n for k, n in layer_shapes(width, hidden))


def shape_of(w):
    """(width, hidden) of a flat parameter vector"""
    for width, hidden in SHAPES:
        if len(w) == param_count(width, hidden):
            return width, hidden
    raise ValueError(f"no network shape has {len(w)} parameters")


def split(w):
    """flat k-major parameters -> list of [k][n] matrices (dtype kept)"""
    w = np.asarray(w)
    out, off = [], 0
    for k, n in layer_shapes(*shape_of(w)):
        out.append(w[off:off + k * n].reshape(k, n))
        off += k * n
    return out


def bf16(x):
    """float32 -> bf16 (round to nearest even) -> float32"""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def encode(raw):
    """raw [n][16] -> [n][80]: sin/cos(2^f pi p_a) frequency-major (first 64),
    direction, normal, roughness, albedo, zero padding"""
    raw = np.asarray(raw, np.float64)
    n = len(raw)
    e = np.zeros((n, K_IN), np.float64)
    k = 0
    for f in range(12):
        for a in range(3):
            if k >= 64:
                break
            arg = np.pi * (2.0 ** f) * raw[:, a]
            e[:, k], e[:, k + 1] = np.sin(arg), np.cos(arg)
            k += 2
    e[:, 64:74] = raw[:, 3:13]
    return e


def forward(w, raw):
    """y [n][3] and the bf16 inputs of every layer, as the GPU computes them"""
    W = [bf16(m) for m in split(np.asarray(w, np.float32))]
    hidden = len(W) - 1
    h = bf16(encode(raw).astype(np.float32))
    hs = [h]
    for l in range(hidden):
        h = bf16(np.maximum(h @ W[l], 0.0).astype(np.float32))
        hs.append(h)
    y = (h @ W[hidden])[:, :3].astype(np.float32)
    return y, hs


def loss_and_grad(w, raw, target):
    """relative L2 loss and its gradient with respect to the flat weights"""
    y, hs = forward(w, raw)
    W = [bf16(m) for m in split(np.asarray(w, np.float32))]
    hidden = len(W) - 1
    n = len(raw)
    d = y - target
    den = y * y + 0.01
    loss = float(np.sum(d * d / den) / (3 * n))
    dz = np.zeros((n, OUT_COLS), np.float32)
    dz[:, :3] = bf16((2.0 * d / den / (3 * n)).astype(np.float32))
    grads = [None] * (hidden + 1)
    for l in range(hidden, -1, -1):
        grads[l] = (hs[l].T.astype(np.float64) @ dz.astype(np.float64)).astype(np.float32)
        if l == 0:
            break
        dh = dz @ W[l].T
        dz = bf16(np.where(hs[l] > 0, dh, 0.0).astype(np.float32))
    return loss, np.concatenate([g.ravel() for g in grads])


def forward64(w, raw):
    W = split(np.asarray(w, np.float64))
    hidden = len(W) - 1
    h = encode(raw)
    for l in range(hidden):
        h = np.maximum(h @ W[l].astype(np.float64), 0.0)
    return (h @ W[hidden].astype(np.float64))[:, :3]


def loss64(w, raw, target):
    y = forward64(w, raw)
    return float(np.mean((y - target) ** 2 / (y * y + 0.01)))


def grad64(w, raw, target):
    """analytic gradient in float64 with the loss denominator held constant
    (the stop-gradient of NRC's relative loss)"""
    W = split(np.asarray(w, np.float64))
    hidden = len(W) - 1
    h = encode(raw)
    hs = [h]
    for l in range(hidden):
        h = np.maximum(h @ W[l], 0.0)
        hs.append(h)
    y = (h @ W[hidden])[:, :3]
    n = len(raw)
    dz = np.zeros((n, OUT_COLS))
    dz[:, :3] = 2.0 * (y - target) / (y * y + 0.01) / (3 * n)
    grads = [None] * (hidden + 1)
    for l in range(hidden, -1, -1):
        grads[l] = hs[l].T @ dz
        if l == 0:
            break
        dz = np.where(hs[l] > 0, dz @ W[l].T, 0.0)
    return np.concatenate([g.ravel() for g in grads])


def adam(w, m, v, g, step, lr):
    """one Adam step on the used parameters (the 13 padded output columns stay 0)"""
    m = 0.9 * m + 0.1 * g
    v = 0.99 * v + 0.01 * g * g
    c1, c2 = 1.0 - 0.9 ** step, 1.0 - 0.99 ** step
    w = w - lr * (m / c1) / (np.sqrt(v / c2) + 1e-8)
    return w, m, v
