"""NumPy oracle of the learned smoothing-length seed network (test
infrastructure only: used by tests/ to check the GPU forward and backward
kernels; the product never imports it).

Network: 28 -> 64 -> 64 -> 1, SiLU, evaluated the way the GPU kernels do:
weights and hidden activations rounded to bfloat16 (round to nearest even),
products and sums in float32 or wider.  The float64 evaluation here bounds the
GPU's float32 accumulation-order differences.

Parameter blob order (float32, 6081 entries), shared with meshoid_amd/seed.py:
    W1 (64 x 28, row-major: out x in), b1 (64), W2 (64 x 64), b2 (64),
    w3 (64), b3 (1)
Inputs are normalised first: x = (features - mu) / sd.
"""
from __future__ import annotations

import numpy as np

NIN, NH = 28, 64
SIZES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("w3", (NH,)),
         ("b3", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(P):
    out, o = {}, 0
    for name, shp in SIZES:
        n = int(np.prod(shp))
        out[name] = np.asarray(P[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(a), np.nan, out)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(P, F, mu, sd):
    """Returns (y, cache): y = predicted log(h / cell) per row."""
    p = unpack(P)
    x = bf16((np.asarray(F, np.float64) - mu) / sd)
    W1, W2, w3 = bf16(p["W1"]), bf16(p["W2"]), bf16(p["w3"])
    z1 = x @ W1.T + p["b1"]
    a1 = bf16(silu(z1))
    z2 = a1 @ W2.T + p["b2"]
    a2 = bf16(silu(z2))
    y = a2 @ w3 + p["b3"][0]
    return y, (x, W1, W2, w3, z1, a1, z2, a2)


def loss_and_grad(P, F, t, mu, sd):
    """Mean squared error of y against t and its gradient w.r.t. the blob.
    Straight-through bfloat16: the rounding is treated as identity in the
    backward pass, as the GPU kernels do."""
    y, (x, W1, W2, w3, z1, a1, z2, a2) = forward(P, F, mu, sd)
    n = len(t)
    r = y - np.asarray(t, np.float64)
    L = np.mean(r * r)
    dy = 2.0 * r / n
    g = {}
    g["w3"] = a2.T @ dy
    g["b3"] = np.array([dy.sum()])
    dz2 = np.outer(dy, w3) * dsilu(z2)
    g["W2"] = dz2.T @ a1
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ W2) * dsilu(z1)
    g["W1"] = dz1.T @ x
    g["b1"] = dz1.sum(0)
    grad = np.concatenate([g[name].ravel() for name, _ in SIZES])
    return L, grad
