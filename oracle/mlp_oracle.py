"""MLP oracle (numpy, FP64 accumulation) -- TEST INFRASTRUCTURE ONLY.

Restates the learned warm start's network for checking libtbec's BF16-MFMA kernels:
dense layers n_in -> width^depth (GELU, erf form) -> n_out, mean-squared-error loss.
Every GEMM operand is rounded to BF16 (nearest even) as the device stages it, the
products are accumulated in FP64 here (FP32 on the device).  Parameters are laid out per
layer as W (fan_in x fan_out, row-major) then b (fan_out), as in include/tbec.h.

Pinned in tests/test_mlp.py: GELU and its derivative against finite differences in FP64,
the backward pass against central differences of this module's own FP64 loss with BF16
rounding switched off.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf


def bf16(x, enabled=True):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    if not enabled:
        return np.asarray(x, dtype=np.float64)
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def unpack(params, dims):
    out, off = [], 0
    for l in range(len(dims) - 1):
        nw = dims[l] * dims[l + 1]
        W = params[off:off + nw].reshape(dims[l], dims[l + 1])
        off += nw
        b = params[off:off + dims[l + 1]]
        off += dims[l + 1]
        out.append((W.astype(np.float64), b.astype(np.float64)))
    return out


def loss_grad(params, dims, X, T, round_bf16=True):
    """(loss, flat gradient, output) of the mean-squared error."""
    layers = unpack(params, dims)
    r = lambda a: bf16(a, round_bf16)  # noqa: E731
    H = [np.asarray(X, np.float64)]
    Z = []
    for l, (W, b) in enumerate(layers):
        z = r(H[-1]) @ r(W) + b
        if l < len(layers) - 1:
            Z.append(z)
            H.append(gelu(z))
        else:
            Y = z
    n = Y.size
    loss = float(np.sum((Y - T) ** 2) / n)
    dZ = 2.0 * (Y - T) / n
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        gW = r(H[l]).T @ r(dZ)
        gb = dZ.sum(axis=0)
        grads[l] = (gW, gb)
        if l == 0:
            break
        dH = r(dZ) @ r(W).T
        dZ = dH * gelu_grad(Z[l - 1])
    flat = np.concatenate([np.concatenate([g.ravel(), b.ravel()]) for g, b in grads])
    return loss, flat, Y
