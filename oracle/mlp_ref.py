"""CPU reference of the learned restructure scheduler's MLP (16 -> 64 -> 64 -> 1, ReLU).

TEST INFRASTRUCTURE ONLY (checker for the HIP inference / training kernels).

Numerics restated from the kernels' contract: weights are rounded to bfloat16
(round-to-nearest-even) for the forward pass, layer inputs are rounded to bfloat16 before
each of the two hidden matrix products, accumulation and biases are fp32 (here float64, so
the comparison tolerance covers accumulation order), the output layer takes the fp32
hidden activations with bf16-rounded weights. Gradients are those of the MSE loss
mean((y - t)^2) with respect to the fp32 master parameters, passing straight through the
bf16 rounding. Parameter packing: W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3.
"""
from __future__ import annotations

import numpy as np

NIN, NH = 16, 64
SIZES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH,)), ("b3", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, np.float64).ravel()
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES]).astype(np.float32)


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float64."""
    a = np.asarray(x, np.float32).copy()
    u = a.view(np.uint32).astype(np.uint64)
    nan = np.isnan(a)
    lsb = (u >> np.uint64(16)) & np.uint64(1)
    u = (u + np.uint64(0x7FFF) + lsb) & np.uint64(0xFFFF0000)
    r = u.astype(np.uint32).view(np.float32).astype(np.float64)
    r[nan] = np.nan
    return r


def forward(params, x, keep=False, quant=True):
    """quant=False drops the bf16 rounding (exact float64 MLP, for finite-difference checks)."""
    P = unpack(params)
    q = bf16 if quant else (lambda a: np.asarray(a, np.float64))
    W1, W2, W3 = q(P["W1"]), q(P["W2"]), q(P["W3"])
    xq = q(x)
    z1 = xq @ W1.T + P["b1"]
    h1 = np.maximum(z1, 0.0)
    if quant:
        h1 = h1.astype(np.float32).astype(np.float64)
    h1q = q(h1)
    z2 = h1q @ W2.T + P["b2"]
    h2 = np.maximum(z2, 0.0)
    if quant:
        h2 = h2.astype(np.float32).astype(np.float64)
    y = h2 @ W3 + P["b3"][0]
    if keep:
        return y, dict(xq=xq, z1=z1, h1q=h1q, z2=z2, h2=h2, W2=W2, W3=W3)
    return y


def loss_and_grad(params, x, t, quant=True):
    y, c = forward(params, x, keep=True, quant=quant)
    n = x.shape[0]
    r = y - np.asarray(t, np.float64)
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n
    g = {"W3": c["h2"].T @ dy, "b3": np.array([dy.sum()])}
    dz2 = np.outer(dy, c["W3"]) * (c["z2"] > 0)
    g["W2"] = dz2.T @ c["h1q"]
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ c["W2"]) * (c["z1"] > 0)
    g["W1"] = dz1.T @ c["xq"]
    g["b1"] = dz1.sum(0)
    return loss, pack(g).astype(np.float64)


def adam_step(params, m, v, grad, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2.0 / NIN), (NH, NIN)), "b1": np.zeros(NH),
         "W2": rng.normal(0, np.sqrt(2.0 / NH), (NH, NH)), "b2": np.zeros(NH),
         "W3": rng.normal(0, np.sqrt(1.0 / NH), NH), "b3": np.zeros(1)}
    return pack(d)
