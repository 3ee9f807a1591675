"""FP64 restatement of the closure network and its analytic backward pass.

TEST INFRASTRUCTURE ONLY: the checker for the matrix-core training and
inference kernels (lb_amd/csrc/closure.hip, mlp.hpp).  The network has no
counterpart in the reference; its definition is the one documented in
include/lbm_closure.h: 12 -> 64 -> 64 -> 1, SiLU, head C = c softplus(z)/ln 2,
loss sum w (C - y)^2, gradient of loss / batch.  Pinned by central finite
differences (tests/test_closure.py).
"""
from __future__ import annotations

import numpy as np

NF, NH = 12, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NF
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NPARAM = OFF_B3 + 1
LN2 = np.log(2.0)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NH, NF), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, w3, [b3]])


def bf16(a):
    """Round to BF16 (nearest even) and back, elementwise."""
    a = np.asarray(a, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def sigmoid(h):
    return 0.5 * (1.0 + np.tanh(0.5 * h))


def silu(h):
    return h * sigmoid(h)


def dsilu(h):
    s = sigmoid(h)
    return s * (1.0 + h * (1.0 - s))


def softplus(z):
    return np.where(z > 20, z, np.log1p(np.exp(np.minimum(z, 20))))


def forward(p, X, c, round_bf16=False):
    """Coefficient C for rows of X (n, 12).  round_bf16 evaluates the network
    as the matrix-core inference does: in the scaled variable g = -log2(e) h
    (logistic 1 / (1 + 2^g)), with the FP32 products -log2(e) W1 rounded to
    BF16, FP32 -log2(e) b1, -log2(e) b2 and -ln(2) w3, and the inputs and
    the scaled layer-1 activations rounded to BF16.  Without rounding the
    scaling cancels and this is the plain network."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    if not round_bf16:
        H1 = X @ W1.T + b1
        A1 = silu(H1)
        H2 = A1 @ W2.T + b2
        A2 = silu(H2)
        z = A2 @ w3 + b3
        return c * softplus(z) / LN2, (X, H1, A1, H2, A2, z)
    f32 = np.float32
    k2, kn = f32(-1.4426950408889634), f32(-0.69314718055994531)

    def scaled(k, a):   # FP32 product k * a
        return (k * np.asarray(a, dtype=f32)).astype(np.float64)

    def gsilu(g):       # g / (1 + 2^g) = -log2(e) SiLU(h)
        return g / (1.0 + np.exp2(g))

    G1 = bf16(X) @ bf16(scaled(k2, W1)).T + scaled(k2, b1)
    A1 = bf16(gsilu(G1))
    G2 = A1 @ bf16(W2).T + scaled(k2, b2)
    z = gsilu(G2) @ scaled(kn, w3) + b3
    return c * softplus(z) / LN2, (X, G1, A1, G2, gsilu(G2), z)


def loss_grad(p, X, y, w, c):
    """(sum w (C - y)^2, d(sum w (C - y)^2 / n) / dp)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    C, (X, H1, A1, H2, A2, z) = forward(p, X, c)
    n = X.shape[0]
    w = np.ones(n) if w is None else np.asarray(w, dtype=np.float64)
    e = C - np.asarray(y, dtype=np.float64)
    loss = float(np.sum(w * e * e))
    dz = 2.0 * w * e / n * c * sigmoid(z) / LN2
    gw3 = A2.T @ dz
    gb3 = dz.sum()
    dH2 = np.outer(dz, w3) * dsilu(H2)
    gW2 = dH2.T @ A1
    gb2 = dH2.sum(0)
    dH1 = (dH2 @ W2) * dsilu(H1)
    gW1 = dH1.T @ X
    gb1 = dH1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gw3, gb3)


def init_params(seed):
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((NH, NF)) / np.sqrt(NF)
    W2 = rng.standard_normal((NH, NH)) / np.sqrt(NH)
    return pack(W1, np.zeros(NH), W2, np.zeros(NH), np.zeros(NH), 0.0)
