"""FP64 numpy restatement of the surrogate network and its loss gradient.

TEST INFRASTRUCTURE ONLY (the FP64 reference of the GPU training kernels): only
tests/ import it.  Same architecture and loss as the GPU path
(include/bhk_surrogate.h): features padded to 32 -> depth x width SiLU ->
16 outputs; loss = sum_rows [sum_c w (y - t)^2 + 0.01 (y_conf - log10(err + 1e-12))^2] / n
with err = sum_c w (y - t)^2 (the confidence column excluded, its target
detached).  The GPU keeps activations and incoming gradients in BF16; this
restatement evaluates in FP64 on the same BF16-rounded weights and inputs.
"""
from __future__ import annotations

import numpy as np

CONF = 11


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray, shapes):
    out, off = [], 0
    for (o, i) in shapes:
        W = params[off:off + o * i].reshape(o, i)
        off += o * i
        b = params[off:off + o]
        off += o
        out.append((W, b))
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, shapes, X):
    layers = unpack(params, shapes)
    h = np.asarray(X, dtype=np.float64)  # the input layer runs in FP32 (unquantised)
    Zs, Hs = [], [h]
    for l, (W, b) in enumerate(layers):
        Wl = W if l == 0 else bf16_round(W)
        z = h @ Wl.T + b
        if l < len(layers) - 1:
            Zs.append(z)
            h = silu(z)
            Hs.append(h)
        else:
            return z, Zs, Hs


def loss_and_grad(params, shapes, X, T, Wt):
    layers = unpack(params, shapes)
    Y, Zs, Hs = forward(params, shapes, X)
    n = X.shape[0]
    mask = np.ones(Y.shape[1], dtype=bool)
    mask[CONF] = False
    d = (Y - T) * mask
    err = (Wt * d * d).sum(axis=1)
    tgt = np.log10(err + 1e-12)
    dconf = Y[:, CONF] - tgt
    loss = (err + 0.01 * dconf ** 2).sum() / n
    dY = 2.0 * Wt * d / n
    dY[:, CONF] = 2.0 * 0.01 * dconf / n
    grads = []
    g = dY
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        grads.append((g.T @ Hs[l], g.sum(axis=0)))
        if l > 0:
            g = (g @ (W if l == 0 else bf16_round(W))) * silu_grad(Zs[l - 1])
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])
    return loss, flat
