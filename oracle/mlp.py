"""FP64 restatement of the warm-start MLP (forward, loss, backward, Adam).

TEST INFRASTRUCTURE ONLY: the checker the HIP training and inference kernels are compared
against; only tests/ may import it. The network has no counterpart in the reference
(Nikita-Lysiuk/Fluid-Engine has no learned component); its definition is the one in
include/fluid_engine.h (fe_learn_*):

    z1 = x W1 + b1, h1 = gelu(z1); z2 = h1 W2 + b2, h2 = gelu(z2); z3 = h2 W3 + b3,
    h3 = gelu(z3); y = h3 . w4 + b4;  gelu(z) = z (1 + erf(z / sqrt 2)) / 2;
    L = sum_b (y_b - t_b)^2 / B.

`bf16=True` rounds at the points where the HIP kernels hold BF16 values (inputs, W1..W3,
w4, the activations h1..h3 and the backward dZ operands of the matrix products), with
round-to-nearest-even; all other arithmetic stays FP64. Pinned in tests by central finite
differences of the FP64 loss and by closed forms (zero weights, a single active unit).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

N_IN = 12
HIDDEN = 64
N_PARAMS = N_IN * HIDDEN + HIDDEN + 2 * (HIDDEN * HIDDEN + HIDDEN) + HIDDEN + 1
N_NORM = 2 * N_IN + 2


def bf16(a) -> np.ndarray:
    """Round to bfloat16 (round to nearest even, via float32), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    assert p.shape == (N_PARAMS,)
    o = 0
    out = []
    for shape in ((N_IN, HIDDEN), (HIDDEN,), (HIDDEN, HIDDEN), (HIDDEN,), (HIDDEN, HIDDEN), (HIDDEN,), (HIDDEN,),
                  ()):
        size = int(np.prod(shape)) if shape else 1
        out.append(p[o:o + size].reshape(shape) if shape else p[o])
        o += size
    return out


def pack(W1, b1, W2, b2, W3, b3, w4, b4):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3, w4, [b4]]).astype(np.float64)


def forward(p, x, bf16_points=False):
    """Network output y (B,) and the intermediates of the backward pass."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(p)
    r = bf16 if bf16_points else (lambda a: np.asarray(a, dtype=np.float64))
    x = r(np.asarray(x, dtype=np.float64).reshape(-1, N_IN))
    Ws = [r(W1), r(W2), r(W3)]
    bs = [b1, b2, b3]
    hs = [x]
    zs = []
    for W, b in zip(Ws, bs):
        z = hs[-1] @ W + b
        zs.append(z)
        hs.append(r(gelu(z)))
    w4r = r(w4)
    y = hs[-1] @ w4r + b4
    return y, (Ws, w4r, hs, zs)


def loss_grad(p, x, t, bf16_points=False):
    """(sum of squared errors, gradient of L = SSE / B w.r.t. the packed parameters)."""
    y, (Ws, w4r, hs, zs) = forward(p, x, bf16_points)
    t = np.asarray(t, dtype=np.float64).reshape(-1)
    B = y.shape[0]
    r = bf16 if bf16_points else (lambda a: a)
    e = y - t
    dy = 2.0 * e / B
    dw4 = hs[3].T @ dy
    db4 = dy.sum()
    dz = np.outer(dy, w4r) * gelu_grad(zs[2])
    dWs = [None] * 3
    dbs = [None] * 3
    for layer in (2, 1, 0):
        dzr = r(dz)
        dbs[layer] = dz.sum(axis=0)
        dWs[layer] = hs[layer].T @ dzr
        if layer > 0:
            dz = (dzr @ Ws[layer].T) * gelu_grad(zs[layer - 1])
    g = pack(dWs[0], dbs[0], dWs[1], dbs[1], dWs[2], dbs[2], dw4, db4)
    return float(e @ e), g


def init_params(seed: int = 0, scale: float = 1.0):
    """Glorot-uniform weights and small random biases (a test initialisation)."""
    rng = np.random.default_rng(seed)
    a1 = np.sqrt(6.0 / (N_IN + HIDDEN))
    a2 = np.sqrt(6.0 / (2 * HIDDEN))
    a4 = np.sqrt(6.0 / (HIDDEN + 1))
    W1 = rng.uniform(-a1, a1, (N_IN, HIDDEN)) * scale
    W2 = rng.uniform(-a2, a2, (HIDDEN, HIDDEN)) * scale
    W3 = rng.uniform(-a2, a2, (HIDDEN, HIDDEN)) * scale
    w4 = rng.uniform(-a4, a4, HIDDEN) * scale
    b1, b2, b3 = (rng.uniform(-0.1, 0.1, HIDDEN) for _ in range(3))
    return pack(W1, b1, W2, b2, W3, b3, w4, rng.uniform(-0.1, 0.1))


def adam_step(p, m1, m2, g, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns new (p, m1, m2)."""
    m1 = beta1 * m1 + (1 - beta1) * g
    m2 = beta2 * m2 + (1 - beta2) * g * g
    bc1 = 1 - beta1 ** step
    bc2 = 1 - beta2 ** step
    return p - lr * (m1 / bc1) / (np.sqrt(m2 / bc2) + eps), m1, m2
