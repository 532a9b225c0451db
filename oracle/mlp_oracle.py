"""CPU oracle of the learned free-surface correction's MLP (test
infrastructure only: tests/ use it as the checker; nothing in the product
imports it).

The network has no counterpart in the reference.  It is restated here from
its definition: y = W3 gelu(W2 gelu(W1 x + b1) + b2) + b3 with
gelu(z) = z Phi(z) (Hendrycks & Gimpel 2016, the erf form), loss
sum (y - t)^2 / (2 B), and its gradient by the chain rule.  The device
kernels keep weights, inputs, activations and the output gradient in bfloat16
(round to nearest even) and accumulate in FP32; `emulate=True` applies the
same roundings, accumulating in FP64, so the two differ by accumulation
order and precision only.  The restatement itself is pinned by tests against
central differences and against torch.autograd in FP64.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

NI, NH, NO = 192, 128, 64
SHAPES = [("W1", (NI, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NO)), ("b3", (NO,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unflatten(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def flatten(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """round float32 values to bfloat16 (nearest even), returned as float64"""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def _r(a, emulate):
    return bf16(a) if emulate else np.asarray(a, dtype=np.float64)


def forward(p, x, emulate=True):
    """returns y and the cache of the backward pass"""
    P = unflatten(p)
    x0 = _r(x, emulate)
    W1, W2, W3 = (_r(P[k], emulate) for k in ("W1", "W2", "W3"))
    z1 = x0 @ W1 + P["b1"]
    h1 = _r(gelu(z1), emulate)
    z2 = h1 @ W2 + P["b2"]
    h2 = _r(gelu(z2), emulate)
    y = h2 @ W3 + P["b3"]
    return y, dict(x0=x0, z1=z1, h1=h1, z2=z2, h2=h2, W1=W1, W2=W2, W3=W3)


def loss_and_grad(p, x, t, emulate=True):
    y, c = forward(p, x, emulate)
    B = x.shape[0]
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.sum(d * d) / (2.0 * B))
    dy = _r(d / B, emulate)
    g = {"W3": c["h2"].T @ dy, "b3": dy.sum(0)}
    dh2 = dy @ c["W3"].T
    dz2 = _r(dh2 * gelu_grad(c["z2"]), emulate)
    g["W2"], g["b2"] = c["h1"].T @ dz2, dz2.sum(0)
    dh1 = dz2 @ c["W2"].T
    dz1 = _r(dh1 * gelu_grad(c["z1"]), emulate)
    g["W1"], g["b1"] = c["x0"].T @ dz1, dz1.sum(0)
    return loss, flatten(g)
