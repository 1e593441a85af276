"""FP64 numpy restatement of the learned warm-start network and its training step.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU MLP
kernels (mujoco-mlx-cpp_amd/csrc/mlp.hip). The engine never imports it.

The reference has no learned warm start (its solver starts from qacc = 0,
reference src/compat/mkx_kernels/solver.hpp:592), so there is no reference
algorithm to follow. What pins this file: the backward pass is checked against
central finite differences of `loss`, and `adamw` against the closed form of
the first two Adam steps (tests/test_mlp_oracle.py).

Network: h1 = silu(x W1 + b1), h2 = silu(h1 W2 + b2), y = h2 W3 + b3;
loss = mean over rows and outputs of (y - t)^2. Parameters are packed in the
order W1 b1 W2 b2 W3 b3, row-major.
"""
from __future__ import annotations

import numpy as np


def shapes(nin: int, nh: int, nout: int):
    return [(nin, nh), (nh,), (nh, nh), (nh,), (nh, nout), (nout,)]


def unpack(p: np.ndarray, nin: int, nh: int, nout: int):
    out, o = [], 0
    for s in shapes(nin, nh, nout):
        k = int(np.prod(s))
        out.append(p[o:o + k].reshape(s))
        o += k
    assert o == p.size
    return out


def pack(blocks) -> np.ndarray:
    return np.concatenate([np.asarray(b, np.float64).ravel() for b in blocks])


def silu(z):
    return z / (1.0 + np.exp(-z))


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) and back to float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, x, nin, nh, nout, round_bf16=False):
    """Returns (y, cache). With round_bf16 the weights, inputs and hidden
    activations are rounded to bfloat16 the way the GPU kernels store them."""
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float64), nin, nh, nout)
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    x = r(x)
    z1 = x @ r(W1) + b1
    h1 = r(silu(z1))
    z2 = h1 @ r(W2) + b2
    h2 = r(silu(z2))
    y = h2 @ r(W3) + b3
    return y, (x, z1, h1, z2, h2)


def loss(p, x, t, nin, nh, nout):
    y, _ = forward(p, x, nin, nh, nout)
    return float(np.mean((y - t) ** 2))


def grad(p, x, t, nin, nh, nout, round_bf16=False):
    """Gradient of the mean squared error w.r.t. the packed parameters."""
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float64), nin, nh, nout)
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    y, (x, z1, h1, z2, h2) = forward(p, x, nin, nh, nout, round_bf16)
    n = y.shape[0]
    dy = 2.0 * (y - t) / (n * nout)
    dy_r = r(dy)
    gW3 = h2.T @ dy_r
    gb3 = dy.sum(0)
    dh2 = dy_r @ r(W3).T
    s2 = 1.0 / (1.0 + np.exp(-z2))
    dz2 = dh2 * s2 * (1.0 + z2 * (1.0 - s2))
    dz2_r = r(dz2)
    gW2 = h1.T @ dz2_r
    gb2 = dz2.sum(0)
    dh1 = dz2_r @ r(W2).T
    s1 = 1.0 / (1.0 + np.exp(-z1))
    dz1 = dh1 * s1 * (1.0 + z1 * (1.0 - s1))
    dz1_r = r(dz1)
    gW1 = x.T @ dz1_r
    gb1 = dz1.sum(0)
    return pack([gW1, gb1, gW2, gb2, gW3, gb3])


def adamw(p, g, m, v, step, lr, wd, b1=0.9, b2=0.999, eps=1e-8):
    """One AdamW step (decoupled weight decay, bias-corrected moments); step counts from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    p = p - lr * (mh / (np.sqrt(vh) + eps) + wd * p)
    return p, m, v
